// A host-memory backend for the device mirror (tpls_amd/csrc/mirror.cpp), so
// that its coherence protocol and fault handling run on machines without a
// GPU: the "device" copy is plain host memory that a test can read and write
// as a kernel would.
#include <cstdlib>
#include <cstring>
#include <string>

#include "mirror.hpp"

namespace {

std::string g_err;

void* dev_alloc(size_t n) { return std::calloc(1, n ? n : 1); }
void dev_release(void* p) { std::free(p); }
int dev_upload(void* d, const void* h, size_t n)
{
    std::memcpy(d, h, n);
    return 0;
}
int dev_download(void* h, const void* d, size_t n)
{
    std::memcpy(h, d, n);
    return 0;
}
int dev_fail(int code, const char* msg)
{
    g_err = msg;
    return code;
}

struct Install {
    Install() { tpls::mirror::set_backend({dev_alloc, dev_release, dev_upload, dev_download, dev_fail}); }
} g_install;

}  // namespace

extern "C" {

void* mh_device_of(const void* h, long long n) { return tpls::mirror::device_of(h, size_t(n)); }
int mh_to_device(const void* h, long long n) { return tpls::mirror::to_device(h, size_t(n)); }
int mh_written(const void* h, long long n) { return tpls::mirror::written(h, size_t(n)); }
int mh_host_current(const void* h, long long n) { return tpls::mirror::host_current(h, size_t(n)) ? 1 : 0; }
int mh_host_access(const void* p, long long n, int write) { return tpls::mirror::host_access(p, size_t(n), write != 0); }
const char* mh_error() { return g_err.c_str(); }

}  // extern "C"
