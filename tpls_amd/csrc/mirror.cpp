// Device mirrors of host arrays: the attach/detach C ABI, the per-granule
// coherence states and the fault handler that keeps them (the protocol is
// described in mirror.hpp).  Plain C++ and POSIX: the device side comes from
// the Backend that the HIP library installs.
#include "mirror.hpp"

#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "tpls_core.h"

namespace tpls {
namespace mirror {
namespace {

enum : uint8_t { CLEAN = 0, DIRTY = 1, STALE = 2 };

constexpr int MAX_ENTRIES = 1024;

struct Entry {
    uintptr_t h = 0, n = 0;    // the attached host bytes [h, h + n)
    uintptr_t p0 = 0, p1 = 0;  // its whole pages [p0, p1), the protected part
    char* d = nullptr;         // device copy of the n bytes
    size_t gsz = 0, ng = 0;    // granule bytes, granule count over [p0, p1)
    uint8_t* st = nullptr;     // per-granule state
    // the end bytes [h, p0) and [p1, h + n) as the device copy holds them, so
    // that an upload of unchanged end bytes is skipped
    char* shadow = nullptr;
};

Backend g_be{};
bool g_have_be = false;
std::recursive_mutex g_mu;  // the table and every state change
thread_local int t_depth = 0;  // how deep this thread is in the lock
Entry* g_tab[MAX_ENTRIES];  // sorted by h, non-overlapping
int g_n = 0;
std::atomic<int> g_count{0};
size_t g_page = 0;
size_t g_granule = size_t(1) << 20;
struct sigaction g_prev;
bool g_installed = false;
// counters: bytes up, bytes down, faults resolved, uploads, downloads
std::atomic<long long> g_ctr[5];

size_t page()
{
    if (!g_page) g_page = size_t(sysconf(_SC_PAGESIZE));
    return g_page;
}

int fail(int code, const char* msg)
{
    if (g_have_be && g_be.fail) return g_be.fail(code, msg);
    return code;
}

// The lock, counting this thread's depth in it so that a fault taken while
// the thread holds it (a host access from inside the mirror itself) is
// reported instead of being resolved under a half-made state change.
struct Lock {
    Lock()
    {
        g_mu.lock();
        ++t_depth;
    }
    ~Lock()
    {
        --t_depth;
        g_mu.unlock();
    }
};

void say(const char* msg)
{
    ssize_t r = write(2, msg, strlen(msg));
    (void)r;
}

int protect(uintptr_t a, uintptr_t b, int prot)
{
    if (b <= a) return 0;
    return mprotect(reinterpret_cast<void*>(a), b - a, prot);
}

int up(Entry* e, uintptr_t a, uintptr_t b)
{
    if (b <= a) return 0;
    if (g_be.upload(e->d + (a - e->h), reinterpret_cast<const void*>(a), b - a)) return -1;
    g_ctr[0] += static_cast<long long>(b - a);
    g_ctr[3] += 1;
    return 0;
}

int down(Entry* e, uintptr_t a, uintptr_t b)
{
    if (b <= a) return 0;
    if (g_be.download(reinterpret_cast<void*>(a), e->d + (a - e->h), b - a)) return -1;
    g_ctr[1] += static_cast<long long>(b - a);
    g_ctr[4] += 1;
    return 0;
}

// The unprotected end bytes: uploaded when they differ from what the device
// copy holds, downloaded after every device write.
int ends_up(Entry* e)
{
    const size_t nh = e->p0 - e->h, nt = e->h + e->n - e->p1;
    if (nh && std::memcmp(e->shadow, reinterpret_cast<const void*>(e->h), nh) != 0) {
        if (up(e, e->h, e->p0)) return -1;
        std::memcpy(e->shadow, reinterpret_cast<const void*>(e->h), nh);
    }
    if (nt && std::memcmp(e->shadow + nh, reinterpret_cast<const void*>(e->p1), nt) != 0) {
        if (up(e, e->p1, e->h + e->n)) return -1;
        std::memcpy(e->shadow + nh, reinterpret_cast<const void*>(e->p1), nt);
    }
    return 0;
}

int ends_down(Entry* e)
{
    const size_t nh = e->p0 - e->h, nt = e->h + e->n - e->p1;
    if (down(e, e->h, e->p0) || down(e, e->p1, e->h + e->n)) return -1;
    std::memcpy(e->shadow, reinterpret_cast<const void*>(e->h), nh);
    std::memcpy(e->shadow + nh, reinterpret_cast<const void*>(e->p1), nt);
    return 0;
}

uintptr_t gbeg(const Entry* e, size_t g) { return e->p0 + g * e->gsz; }
uintptr_t gend(const Entry* e, size_t g) { return std::min<uintptr_t>(e->p1, e->p0 + (g + 1) * e->gsz); }

// Index of the first entry with h > a, the table being sorted by h.
int upper(uintptr_t a)
{
    int lo = 0, hi = g_n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (g_tab[mid]->h <= a) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

Entry* containing(uintptr_t a)
{
    const int i = upper(a) - 1;
    if (i < 0) return nullptr;
    Entry* e = g_tab[i];
    return a < e->h + e->n ? e : nullptr;
}

Entry* exact(const void* host, size_t bytes)
{
    const uintptr_t h = reinterpret_cast<uintptr_t>(host);
    Entry* e = containing(h);
    return e && e->h == h && e->n == bytes ? e : nullptr;
}

// Runs of granules in one state: f(first, last + 1).
template <class F>
int runs(Entry* e, size_t g0, size_t g1, uint8_t state, F&& f)
{
    for (size_t g = g0; g < g1;) {
        if (e->st[g] != state) {
            ++g;
            continue;
        }
        size_t q = g;
        while (q < g1 && e->st[q] == state) ++q;
        if (int rc = f(g, q)) return rc;
        g = q;
    }
    return 0;
}

// STALE granules [g0, g1) made current on the host, left CLEAN (read-only).
int fetch(Entry* e, size_t g0, size_t g1)
{
    return runs(e, g0, g1, STALE, [&](size_t a, size_t b) {
        const uintptr_t x = gbeg(e, a), y = gend(e, b - 1);
        if (protect(x, y, PROT_READ | PROT_WRITE)) return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
        if (down(e, x, y)) return fail(TPLS_ERR_HIP, "mirror: device-to-host copy failed");
        if (protect(x, y, PROT_READ)) return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
        std::fill(e->st + a, e->st + b, uint8_t(CLEAN));
        return 0;
    });
}

void on_fault(int sig, siginfo_t* si, void* uc);

void install()
{
    struct sigaction cur;
    if (sigaction(SIGSEGV, nullptr, &cur) == 0 && (cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == on_fault) return;
    struct sigaction sa;
    std::memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_fault;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigemptyset(&sa.sa_mask);
    struct sigaction prev;
    if (sigaction(SIGSEGV, &sa, &prev) == 0) {
        // a handler installed after ours and then replaced by ours again is
        // not chained to itself
        if (!((prev.sa_flags & SA_SIGINFO) && prev.sa_sigaction == on_fault)) g_prev = prev;
        g_installed = true;
    }
}

void forward(int sig, siginfo_t* si, void* uc)
{
    if (g_installed && (g_prev.sa_flags & SA_SIGINFO) && g_prev.sa_sigaction) {
        g_prev.sa_sigaction(sig, si, uc);
        return;
    }
    if (g_installed && !(g_prev.sa_flags & SA_SIGINFO) && g_prev.sa_handler != SIG_DFL &&
        g_prev.sa_handler != SIG_IGN) {
        g_prev.sa_handler(sig);
        return;
    }
    // the default action: the faulting access runs again and ends the process
    signal(sig, SIG_DFL);
}

// A protection fault inside an attached array.  Returns false when the
// address is not in the protected part of any attached array.
bool resolve(uintptr_t a)
{
    if (t_depth > 0) {
        say("tpls mirror: host access to an attached array from inside the mirror\n");
        return false;
    }
    Lock lk;
    Entry* e = containing(a);
    if (!e || a < e->p0 || a >= e->p1) return false;
    const size_t g = (a - e->p0) / e->gsz;
    const uintptr_t x = gbeg(e, g), y = gend(e, g);
    switch (e->st[g]) {
    case STALE:
        // the device wrote it: fetch, and leave it read-only so that a write
        // faults once more and marks it DIRTY
        if (fetch(e, g, g + 1)) {
            say("tpls mirror: could not fetch an attached array's data from the device\n");
            std::abort();
        }
        break;
    case CLEAN:
        // read-only pages fault only on a write
        if (protect(x, y, PROT_READ | PROT_WRITE)) {
            say("tpls mirror: mprotect failed\n");
            std::abort();
        }
        e->st[g] = DIRTY;
        break;
    default:
        // DIRTY pages are writable; another thread changed the protection
        // between this thread's access and the fault
        protect(x, y, PROT_READ | PROT_WRITE);
        break;
    }
    g_ctr[2] += 1;
    return true;
}

void on_fault(int sig, siginfo_t* si, void* uc)
{
    const int saved = errno;
    const bool ours = si && si->si_code == SEGV_ACCERR && g_count.load() > 0 &&
                      resolve(reinterpret_cast<uintptr_t>(si->si_addr));
    errno = saved;
    if (!ours) forward(sig, si, uc);
}

void drop(Entry* e)
{
    if (e->d && g_have_be) g_be.release(e->d);
    delete[] e->st;
    delete[] e->shadow;
    delete e;
}

}  // namespace

// Called from a static initialiser of the library that provides the device
// (before any attach, single-threaded): no lock, whose own initialisation
// may not have run yet.
void set_backend(const Backend& b)
{
    g_be = b;
    g_have_be = true;
}

bool any() { return g_count.load() > 0; }

void* device_of(const void* host, size_t bytes)
{
    if (!any()) return nullptr;
    Lock lk;
    Entry* e = exact(host, bytes);
    return e ? e->d : nullptr;
}

int to_device(const void* host, size_t bytes)
{
    Lock lk;
    Entry* e = exact(host, bytes);
    if (!e) return fail(TPLS_ERR_ARG, "mirror: not an attached array");
    install();
    if (ends_up(e)) return fail(TPLS_ERR_HIP, "mirror: host-to-device copy failed");
    return runs(e, 0, e->ng, DIRTY, [&](size_t a, size_t b) {
        const uintptr_t x = gbeg(e, a), y = gend(e, b - 1);
        // read-only first, so that a host write during the copy faults
        if (protect(x, y, PROT_READ)) return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
        if (up(e, x, y)) return fail(TPLS_ERR_HIP, "mirror: host-to-device copy failed");
        std::fill(e->st + a, e->st + b, uint8_t(CLEAN));
        return 0;
    });
}

int written(const void* host, size_t bytes)
{
    Lock lk;
    Entry* e = exact(host, bytes);
    if (!e) return fail(TPLS_ERR_ARG, "mirror: not an attached array");
    if (e->ng) {
        if (protect(e->p0, e->p1, PROT_NONE)) return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
        std::fill(e->st, e->st + e->ng, uint8_t(STALE));
    }
    if (ends_down(e)) return fail(TPLS_ERR_HIP, "mirror: device-to-host copy failed");
    return 0;
}

bool host_current(const void* host, size_t bytes)
{
    Lock lk;
    Entry* e = exact(host, bytes);
    if (!e) return true;
    return std::find(e->st, e->st + e->ng, uint8_t(STALE)) == e->st + e->ng;
}

int host_access(const void* p, size_t bytes, bool write)
{
    if (!any() || !bytes) return 0;
    Lock lk;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p), b = a + bytes;
    for (int i = std::max(0, upper(a) - 1); i < g_n && g_tab[i]->h < b; ++i) {
        Entry* e = g_tab[i];
        const uintptr_t x = std::max(a, e->p0), y = std::min(b, e->p1);
        if (x >= y) continue;
        const size_t g0 = (x - e->p0) / e->gsz, g1 = (y - 1 - e->p0) / e->gsz + 1;
        if (int rc = fetch(e, g0, g1)) return rc;
        if (write) {
            if (protect(gbeg(e, g0), gend(e, g1 - 1), PROT_READ | PROT_WRITE))
                return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
            std::fill(e->st + g0, e->st + g1, uint8_t(DIRTY));
        }
    }
    return 0;
}

}  // namespace mirror
}  // namespace tpls

using namespace tpls::mirror;

extern "C" {

int tpls_mirror_attach(void* host, long long bytes)
{
    if (!host || bytes <= 0) return fail(TPLS_ERR_ARG, "mirror: null array or empty size");
    if (!g_have_be) return fail(TPLS_ERR_NOGPU, "mirror: no device backend");
    Lock lk;
    const uintptr_t h = reinterpret_cast<uintptr_t>(host), n = uintptr_t(bytes);
    // an array already attached with the same bounds is left as it is
    const int i = upper(h);
    if (i > 0 && g_tab[i - 1]->h == h && g_tab[i - 1]->n == n) return TPLS_OK;
    if ((i > 0 && g_tab[i - 1]->h + g_tab[i - 1]->n > h) || (i < g_n && g_tab[i]->h < h + n))
        return fail(TPLS_ERR_ARG, "mirror: the array overlaps an attached array of other bounds");
    if (g_n == MAX_ENTRIES) return fail(TPLS_ERR_ARG, "mirror: too many attached arrays");
    Entry* e = new Entry;
    const uintptr_t pg = page();
    e->h = h;
    e->n = n;
    e->p0 = (h + pg - 1) / pg * pg;
    e->p1 = (h + n) / pg * pg;
    if (e->p1 <= e->p0) e->p0 = e->p1 = h + n;  // no whole page: never protected
    e->gsz = std::max<size_t>(pg, g_granule / pg * pg);
    e->ng = (e->p1 - e->p0 + e->gsz - 1) / e->gsz;
    e->st = new uint8_t[std::max<size_t>(1, e->ng)];
    std::fill(e->st, e->st + std::max<size_t>(1, e->ng), uint8_t(CLEAN));
    const size_t nh = e->p0 - h, nt = h + n - e->p1;
    e->shadow = new char[std::max<size_t>(1, nh + nt)];
    std::memcpy(e->shadow, host, nh);
    std::memcpy(e->shadow + nh, reinterpret_cast<const void*>(e->p1), nt);
    e->d = static_cast<char*>(g_be.alloc(n));
    if (!e->d) {
        drop(e);
        return fail(TPLS_ERR_NOMEM, "mirror: device allocation failed");
    }
    if (up(e, h, h + n)) {
        drop(e);
        return fail(TPLS_ERR_HIP, "mirror: host-to-device copy failed");
    }
    install();
    if (protect(e->p0, e->p1, PROT_READ)) {
        drop(e);
        return fail(TPLS_ERR_ARG, "mirror: mprotect failed (is the array in writable memory?)");
    }
    std::memmove(g_tab + i + 1, g_tab + i, sizeof(Entry*) * size_t(g_n - i));
    g_tab[i] = e;
    ++g_n;
    g_count.store(g_n);
    return TPLS_OK;
}

int tpls_mirror_detach(void* host)
{
    Lock lk;
    const uintptr_t h = reinterpret_cast<uintptr_t>(host);
    const int i = upper(h) - 1;
    if (i < 0 || g_tab[i]->h != h) return fail(TPLS_ERR_ARG, "mirror: not an attached array");
    Entry* e = g_tab[i];
    if (int rc = fetch(e, 0, e->ng)) return rc;
    if (protect(e->p0, e->p1, PROT_READ | PROT_WRITE)) return fail(TPLS_ERR_ARG, "mirror: mprotect failed");
    std::memmove(g_tab + i, g_tab + i + 1, sizeof(Entry*) * size_t(g_n - i - 1));
    --g_n;
    g_count.store(g_n);
    drop(e);
    return TPLS_OK;
}

int tpls_mirror_detach_all(void)
{
    Lock lk;
    while (g_n > 0)
        if (int rc = tpls_mirror_detach(reinterpret_cast<void*>(g_tab[g_n - 1]->h))) return rc;
    return TPLS_OK;
}

int tpls_mirror_flush(void)
{
    Lock lk;
    for (int i = 0; i < g_n; ++i)
        if (int rc = fetch(g_tab[i], 0, g_tab[i]->ng)) return rc;
    return TPLS_OK;
}

int tpls_mirror_set_granule(long long bytes)
{
    if (bytes <= 0) return fail(TPLS_ERR_ARG, "mirror: granule must be positive");
    Lock lk;
    g_granule = size_t(bytes);
    return TPLS_OK;
}

int tpls_mirror_counters(long long* out)
{
    if (!out) return fail(TPLS_ERR_ARG, "null argument");
    for (int q = 0; q < 5; ++q) out[q] = g_ctr[q].load();
    out[5] = g_count.load();
    return TPLS_OK;
}

}  // extern "C"
