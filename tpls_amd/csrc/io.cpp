// The reference's text formats on the host side of the core (include/tpls_io.h):
// 3-D data files (io.F90:34-169), NAME VALUE options (options_utils.F90:56-374),
// the run configuration (main_ns_hybrid.F90:242-311 through configure.F90) and
// loading a case from a configuration directory (main_ns_hybrid.F90:380-393).
//
// Data files are large (a 513x129x257 pressure file holds 17 M records), so
// they are memory-mapped and parsed, or formatted, by several threads: the
// records are split at line boundaries, each thread counts the records that
// carry a value in its share, a prefix sum gives every share its first array
// index, and a second pass converts.  Values are written as the shortest
// decimal that reads back to the same double (std::to_chars), and read with
// correctly rounded conversion (std::from_chars), so files round-trip exactly.
// Coordinates and face averages of the channel outputs round as the
// reference's expressions do: no fused multiply-adds.
#pragma clang fp contract(off)

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/tpls_io.h"
#include "host.hpp"

using tpls::set_error;

struct tpls_options {
    std::vector<std::pair<std::string, std::string>> kv;
};

namespace {

// ---------------------------------------------------------------------------
// Fortran list-directed input, one item at a time.

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }
inline bool is_sep(char c) { return is_blank(c) || c == ',' || c == '/' || c == '\n'; }
inline char lower(char c) { return (c >= 'A' && c <= 'Z') ? char(c - 'A' + 'a') : c; }

enum Item { ITEM_VALUE, ITEM_NULL, ITEM_NONE, ITEM_BAD, ITEM_END };

// A decimal beyond the double range: an underflow rounds to the nearest
// subnormal or zero, an overflow is an input error (as the Fortran runtime
// treats them).
bool out_of_range(const char* s, double* out)
{
    const double v = std::strtod(s, nullptr);
    if (std::isinf(v)) return false;
    *out = v;
    return true;
}

// Real constant of Fortran input: [sign] mantissa [exponent], the exponent
// written E, D or Q with optional sign and digits, or a bare signed integer
// (1.5+3 is 1.5e3); also Inf, Infinity and NaN.  Converted with correct
// rounding.
bool parse_real(const char* b, const char* e, double* out)
{
    if (b == e) return false;
    // fast path: the forms std::from_chars takes as they are
    const char* q = b + (*b == '+');
    if (q < e && (q == b || (*q != '+' && *q != '-'))) {
        double v;
        auto r = std::from_chars(q, e, v, std::chars_format::general);
        if (r.ec == std::errc() && r.ptr == e) {
            // from_chars also takes inf/nan spellings Fortran takes
            *out = v;
            return true;
        }
        if (r.ec == std::errc::result_out_of_range && r.ptr == e) return out_of_range(std::string(q, e).c_str(), out);
    }
    // general form: normalise into "[-]digits[.digits]e[-]digits"
    char buf[128];
    size_t n = 0;
    const char* p = b;
    if (*p == '+' || *p == '-') {
        if (*p == '-') buf[n++] = '-';
        ++p;
    }
    int digits = 0, dots = 0;
    while (p < e && n < 100) {
        if (*p >= '0' && *p <= '9') {
            ++digits;
        } else if (*p == '.') {
            if (++dots > 1) return false;
        } else {
            break;
        }
        buf[n++] = *p++;
    }
    if (!digits || n >= 100) return false;
    if (p < e) {
        char c = lower(*p);
        if (c == 'e' || c == 'd' || c == 'q') ++p;
        else if (c != '+' && c != '-') return false;
        buf[n++] = 'e';
        if (p < e && (*p == '+' || *p == '-')) buf[n++] = *p++;
        int ed = 0;
        while (p < e && *p >= '0' && *p <= '9' && n < 120) {
            buf[n++] = *p++;
            ++ed;
        }
        if (!ed || p != e) return false;
    }
    buf[n] = 0;
    double v;
    auto r = std::from_chars(buf, buf + n, v, std::chars_format::general);
    if (r.ptr != buf + n) return false;
    if (r.ec == std::errc::result_out_of_range) return out_of_range(buf, out);
    if (r.ec != std::errc()) return false;
    *out = v;
    return true;
}

bool parse_int(const char* b, const char* e, long* out)
{
    const char* p = b;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p == e) return false;
    long v = 0;
    for (; p < e; ++p) {
        if (*p < '0' || *p > '9') return false;
        v = v * 10 + (*p - '0');
        if (v > 4294967296L) return false;
    }
    *out = neg ? -v : v;
    return true;
}

// Logical input: optional '.', then T or F, then anything up to a separator.
bool parse_logical(const char* b, const char* e, int* out)
{
    const char* p = b;
    if (p < e && *p == '.') ++p;
    if (p == e) return false;
    char c = lower(*p);
    if (c == 't') *out = 1;
    else if (c == 'f') *out = 0;
    else return false;
    return true;
}

// The next list-directed item of the record [p, e): its constant in
// [*tb, *te).  A repeat count r*c yields c (r*c with nothing after the star is
// a null value); a slash ends the input items (ITEM_END).  *p advances past
// the item and its separator.
Item next_item(const char*& p, const char* e, const char** tb, const char** te)
{
    while (p < e && is_blank(*p)) ++p;
    if (p >= e) return ITEM_NONE;
    if (*p == ',') {
        ++p;
        return ITEM_NULL;
    }
    if (*p == '/') {
        p = e;
        return ITEM_END;
    }
    const char* b = p;
    while (p < e && !is_sep(*p)) ++p;
    const char* t = b;
    const char* star = static_cast<const char*>(std::memchr(b, '*', size_t(p - b)));
    if (star) {
        for (const char* d = b; d < star; ++d)
            if (*d < '0' || *d > '9') return ITEM_BAD;
        if (star == b) return ITEM_BAD;
        t = star + 1;
    }
    *tb = t;
    *te = p;
    // consume the separator after the item: blanks, then one comma
    while (p < e && is_blank(*p)) ++p;
    if (p < e && *p == ',') ++p;
    return t == *te ? ITEM_NULL : ITEM_VALUE;
}

// ---------------------------------------------------------------------------
// Files.

struct MappedFile {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    void* map = nullptr;

    int open(const char* path)
    {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return set_error(TPLS_ERR_IO, std::string(path) + ": " + std::strerror(errno));
        struct stat st;
        if (fstat(fd, &st) != 0) return set_error(TPLS_ERR_IO, std::string(path) + ": " + std::strerror(errno));
        n = size_t(st.st_size);
        if (n == 0) {
            p = "";
            return TPLS_OK;
        }
        map = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            map = nullptr;
            return set_error(TPLS_ERR_IO, std::string(path) + ": mmap: " + std::strerror(errno));
        }
        madvise(map, n, MADV_SEQUENTIAL);
        madvise(map, n, MADV_WILLNEED);
        p = static_cast<const char*>(map);
        return TPLS_OK;
    }
    ~MappedFile()
    {
        if (map) munmap(map, n);
        if (fd >= 0) ::close(fd);
    }
};

bool file_exists(const char* path)
{
    struct stat st;
    return ::stat(path, &st) == 0;
}

inline const char* record_end(const char* p, const char* e)
{
    const char* q = static_cast<const char*>(std::memchr(p, '\n', size_t(e - p)));
    return q ? q : e;
}

int io_threads(size_t bytes)
{
    static const int env = [] {
        const char* s = std::getenv("TPLS_IO_THREADS");
        return s ? std::atoi(s) : 0;
    }();
    int hc = int(std::thread::hardware_concurrency());
    int t = env > 0 ? env : std::min(std::max(hc, 1), 16);
    const size_t per = size_t(2) << 20;  // at least 2 MiB of text per thread
    return int(std::max<size_t>(1, std::min<size_t>(size_t(t), bytes / per + 1)));
}

template <class F>
void parallel_for(int nt, F&& f)
{
    if (nt <= 1) {
        f(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) th.emplace_back([&f, t] { f(t); });
    f(0);
    for (auto& x : th) x.join();
}

struct DataHeader {
    int b[6];
    const char* body;  // first record after the bounds
};

// Two records of metadata, then the six bounds read list-directed (they may
// span records; the rest of the record holding the sixth is skipped).
int read_header(const MappedFile& f, const char* path, DataHeader* h)
{
    const char* p = f.p;
    const char* e = f.p + f.n;
    for (int r = 0; r < 2; ++r) {
        if (p >= e) return set_error(TPLS_ERR_FORMAT, std::string(path) + ": ends before the bounds record");
        p = record_end(p, e) + 1;
    }
    int got = 0;
    while (got < 6) {
        if (p >= e) return set_error(TPLS_ERR_FORMAT, std::string(path) + ": ends inside the bounds");
        const char* re = record_end(p, e);
        const char* q = p;
        while (got < 6) {
            const char *tb, *te;
            Item it = next_item(q, re, &tb, &te);
            if (it == ITEM_NONE) break;
            long v;
            if (it != ITEM_VALUE || !parse_int(tb, te, &v) || v < -2147483647L || v > 2147483647L)
                return set_error(TPLS_ERR_FORMAT, std::string(path) + ": bad bounds record");
            h->b[got++] = int(v);
        }
        p = re + 1;
    }
    h->body = std::min(p, e);
    return TPLS_OK;
}

long bound_count(const int* b)
{
    long n = 1;
    for (int d = 0; d < 3; ++d) n *= std::max(0, b[2 * d + 1] - b[2 * d] + 1);
    return n;
}

// Reads the records of [body, end) that carry values (blank records are
// skipped, as list-directed input skips them): record r gives the first
// ncols items, which sink(r, v, got) stores (got[c] is false for a null item,
// which leaves the value as it was; a slash ends the record's items early).
// Further items of a record are ignored.  Returns the number of value
// records (capped at count), or a negative error.  A record with fewer than
// ncols items is an error here (Fortran would read on into the next record).
template <class Sink>
long read_records(const char* body, const char* end, const char* path, int ncols, long count, Sink&& sink)
{
    const size_t bytes = size_t(end - body);
    const int nt = io_threads(bytes);
    std::vector<const char*> cut(nt + 1);
    cut[0] = body;
    cut[nt] = end;
    for (int t = 1; t < nt; ++t) {
        const char* c = body + bytes * size_t(t) / size_t(nt);
        c = std::max(c, cut[t - 1]);
        // a share starts at the beginning of a record
        if (c > body && c[-1] != '\n') c = std::min(record_end(c, end) + 1, end);
        cut[t] = c;
    }
    std::vector<long> nrec(nt, 0), first(nt + 1, 0);
    parallel_for(nt, [&](int t) {
        long n = 0;
        for (const char* p = cut[t]; p < cut[t + 1];) {
            const char* re = record_end(p, cut[t + 1]);
            for (const char* q = p; q < re; ++q)
                if (!is_blank(*q)) {
                    ++n;
                    break;
                }
            p = re + 1;
        }
        nrec[t] = n;
    });
    for (int t = 0; t < nt; ++t) first[t + 1] = first[t] + nrec[t];
    std::vector<long> bad(nt, -1);
    std::vector<int> why(nt, 0);
    parallel_for(nt, [&](int t) {
        long idx = first[t];
        double v[8];
        bool got[8];
        for (const char* p = cut[t]; p < cut[t + 1] && idx < count;) {
            const char* re = record_end(p, cut[t + 1]);
            const char* q = p;
            p = re + 1;
            bool blank = false;
            int c = 0;
            for (; c < ncols; ++c) {
                const char *tb, *te;
                Item it = next_item(q, re, &tb, &te);
                if (it == ITEM_END) break;  // a slash: the remaining items keep their values
                if (it == ITEM_NONE) {
                    if (c == 0) {
                        blank = true;  // list-directed input skips blank records
                        break;
                    }
                    bad[t] = idx;
                    why[t] = 1;
                    return;
                }
                got[c] = it == ITEM_VALUE;
                if (it == ITEM_BAD || (got[c] && !parse_real(tb, te, &v[c]))) {
                    bad[t] = idx;
                    return;
                }
            }
            if (blank) continue;
            for (int r = c; r < ncols; ++r) got[r] = false;
            sink(idx, v, got);
            ++idx;
        }
    });
    for (int t = 0; t < nt; ++t)
        if (bad[t] >= 0)
            return set_error(TPLS_ERR_FORMAT, std::string(path) + ": record " + std::to_string(bad[t] + 1) +
                                                  (why[t] ? " has too few values" : " is not a real number"));
    return std::min(first[nt], count);
}

long read_values(const char* body, const char* end, const char* path, double* out, long count)
{
    return read_records(body, end, path, 1, count, [&](long r, const double* v, const bool* got) {
        if (got[0]) out[r] = v[0];
    });
}

int write_all(FILE* fp, const char* data, size_t n, const char* path)
{
    if (n && std::fwrite(data, 1, n, fp) != n)
        return set_error(TPLS_ERR_IO, std::string(path) + ": write: " + std::strerror(errno));
    return TPLS_OK;
}

// count records of ncols values each, " v1 v2 ...\n" with every value the
// shortest decimal that reads back to the same double; values(r, v) gives
// record r.  Blocks of records are formatted in parallel and written in order.
template <class Values>
int write_records(FILE* fp, const char* path, long count, int ncols, Values&& values)
{
    const long block = std::max(1L, (1L << 18) / ncols);
    const int nt = io_threads(size_t(count) * size_t(ncols) * 24);
    const size_t cap = size_t(block) * size_t(ncols) * 26 + 64;
    std::vector<std::vector<char>> buf(nt, std::vector<char>(cap));
    std::vector<size_t> len(nt);
    int rc = TPLS_OK;
    for (long base = 0; base < count && !rc; base += block * nt) {
        parallel_for(nt, [&](int t) {
            const long b0 = base + t * block, b1 = std::min(count, b0 + block);
            char* p = buf[t].data();
            char* e = p + buf[t].size();
            double v[8];
            for (long r = b0; r < b1; ++r) {
                values(r, v);
                for (int c = 0; c < ncols; ++c) {
                    *p++ = ' ';
                    p = std::to_chars(p, e, v[c]).ptr;
                }
                *p++ = '\n';
            }
            len[t] = b0 < b1 ? size_t(p - buf[t].data()) : 0;
        });
        for (int t = 0; t < nt && !rc; ++t) rc = write_all(fp, buf[t].data(), len[t], path);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// Options.

std::string rtrim(std::string s)
{
    size_t n = s.size();
    while (n && s[n - 1] == ' ') --n;
    s.resize(n);
    return s;
}

std::pair<std::string, std::string>* find_option(tpls_options* o, const std::string& name)
{
    const std::string key = rtrim(name);
    for (auto& kv : o->kv)
        if (kv.first == key) return &kv;
    return nullptr;
}

const std::pair<std::string, std::string>* find_option(const tpls_options* o, const std::string& name)
{
    return find_option(const_cast<tpls_options*>(o), name);
}

void put(tpls_options* o, const std::string& name, const std::string& value)
{
    if (auto* kv = find_option(o, name)) {
        kv->second = rtrim(value);
        return;
    }
    o->kv.emplace_back(rtrim(name), rtrim(value));
}

// The value string a typed get reads: the option's first 60 characters
// (character(len=60) :: value_str, options_utils.F90:99-135).
int typed_value(const tpls_options* o, const char* name, std::string* s, int* ierr)
{
    if (!o || !name || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    const auto* kv = find_option(o, name);
    if (!kv) {
        *ierr = TPLS_OPTION_NOT_FOUND_ERR;
        return TPLS_OK;
    }
    *ierr = 0;
    *s = kv->second.substr(0, 60);
    return TPLS_OK;
}

int bad_value(const char* name, const std::string& v, const char* what, int* ierr)
{
    *ierr = TPLS_OPTION_FORMAT_ERR;
    return set_error(TPLS_ERR_FORMAT, std::string("option ") + name + ": '" + v + "' is not " + what);
}

// Fortran F20.10 / I10 output fields.
std::string f20_10(double v)
{
    char buf[512];
    if (std::isnan(v)) std::snprintf(buf, sizeof buf, "%20s", "NaN");
    else if (std::isinf(v)) std::snprintf(buf, sizeof buf, "%20s", v < 0 ? "-Inf" : "Inf");
    else std::snprintf(buf, sizeof buf, "%20.10f", v);
    std::string s(buf);
    if (s.size() > 20) s.assign(20, '*');
    return s;
}

std::string i10(int v)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%10d", v);
    std::string s(buf);
    if (s.size() > 10) s.assign(10, '*');
    return s;
}

int load_options_file(const char* path, tpls_options* o, int* ierr)
{
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    const char* p = f.p;
    const char* e = f.p + f.n;
    while (p < e) {
        const char* re = record_end(p, e);
        std::string line(p, re);
        p = re + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();  // CRLF files
        // read(...,'(A)') line into character(80), then adjustl
        if (line.size() > 80) line.resize(80);
        size_t lead = line.find_first_not_of(' ');
        if (lead == std::string::npos) continue;  // blank line
        line = line.substr(lead);
        if (line[0] == '#') continue;
        line.resize(80, ' ');
        size_t lt = line.find_last_not_of(' ') + 1;
        size_t sp = line.find(' ');
        const size_t split = sp == std::string::npos ? 0 : sp + 1;  // Fortran index()
        if (split < lt) {
            std::string name = split ? line.substr(0, split - 1) : std::string();
            std::string value = rtrim(line.substr(split));
            if (name.size() > 20) name.resize(20);
            if (value.size() > 60) value.resize(60);
            put(o, name, value);
        } else {
            *ierr = TPLS_OPTION_FORMAT_ERR;
            return TPLS_OK;
        }
    }
    *ierr = 0;
    return TPLS_OK;
}

int append_options_file(const char* path, const tpls_options* o, const char* mode)
{
    FILE* fp = std::fopen(path, mode);
    if (!fp) return set_error(TPLS_ERR_IO, std::string(path) + ": " + std::strerror(errno));
    std::string s;
    for (const auto& kv : o->kv) s += kv.first + " " + kv.second + "\n";
    int rc = write_all(fp, s.data(), s.size(), path);
    if (std::fclose(fp) != 0 && !rc) rc = set_error(TPLS_ERR_IO, std::string(path) + ": close failed");
    return rc;
}

// ---------------------------------------------------------------------------
// Configuration (configure.F90).  Each getter fails with the reference's
// code and message.

struct Cfg {
    const tpls_options* o;
    int* ierr;
    std::string msg;

    int missing(const char* key)
    {
        msg = std::string("Error: Missing value: ") + key;
        return *ierr = TPLS_OPTION_NOT_FOUND_ERR;
    }
    int invalid_i(const char* key, int v, const char* expected)
    {
        char b[32];
        std::snprintf(b, sizeof b, "%6d", v);
        msg = std::string("Error: Invalid value: ") + key + " " + b + " " + expected;
        return *ierr = TPLS_OPTION_INVALID_ERR;
    }
    int invalid_d(const char* key, double v, const char* expected)
    {
        msg = std::string("Error: Invalid value: ") + key + " " + f20_10(v) + " " + expected;
        return *ierr = TPLS_OPTION_INVALID_ERR;
    }
    // 0, a reference code (> 0), or a library error (< 0)
    int geti(const char* key, int* v, int min, const char* missing_key = nullptr)
    {
        int rc = tpls_get_integer_option(o, key, v, ierr);
        if (rc) return rc;
        if (*ierr) return missing(missing_key ? missing_key : key);
        if (*v < min) return invalid_i(key, *v, "Value must be >= 1");
        return 0;
    }
    int getd(const char* key, double* v, const char* missing_key = nullptr)
    {
        int rc = tpls_get_double_option(o, key, v, ierr);
        if (rc) return rc;
        if (*ierr) return missing(missing_key ? missing_key : key);
        return 0;
    }
    int getl(const char* key, int* v)
    {
        int rc = tpls_get_logical_option(o, key, v, ierr);
        if (rc) return rc;
        if (*ierr) return missing(key);
        return 0;
    }
};

#define CFG(x)                \
    do {                      \
        int r_ = (x);         \
        if (r_) return r_;    \
    } while (0)

// initial_config.dat: main_ns_hybrid.F90:263-285.
int initial_conditions(Cfg& c, tpls_params* prm, tpls_run_config* run)
{
    // get_grid_configuration (configure.F90:137-182)
    CFG(c.geti("maxl", &run->maxl, 1));
    CFG(c.geti("maxm", &run->maxm, 1));
    CFG(c.geti("maxn", &run->maxn, 1));
    // get_fluid_configuration (configure.F90:197-286), smooth width unscaled;
    // a missing smooth_width is reported under smooth_width_scale, as there
    CFG(c.getd("re", &prm->Re));
    CFG(c.getd("mu_minus", &prm->mu_minus));
    CFG(c.getd("mu_plus", &prm->mu_plus));
    CFG(c.getd("height", &prm->height));
    if (prm->height < 0 || prm->height > 1) return c.invalid_d("height", prm->height, "Value must be 0 .. 1");
    CFG(c.getd("dpdl", &run->dpdl));
    CFG(c.getd("scap", &prm->scap));
    CFG(c.getd("omega_max", &run->omega_max));
    CFG(c.getd("smooth_width", &prm->smooth_width, "smooth_width_scale"));
    // get_dgrid_configuration (configure.F90:363-395)
    CFG(c.getd("dx", &prm->dx));
    CFG(c.getd("dy", &prm->dy));
    CFG(c.getd("dz", &prm->dz));
    // get_timestep (configure.F90:294-315)
    CFG(c.getd("dt", &prm->dt));
    if (prm->dt <= 0) return c.invalid_d("dt", prm->dt, "Value must be > 0");
    // get_interface_detection_method (configure.F90:323-349): character(len=5)
    char idm[64];
    CFG(tpls_get_option(c.o, "idm", idm, sizeof idm, c.ierr));
    if (*c.ierr) return c.missing("idm");
    std::string s = rtrim(std::string(idm).substr(0, 5));
    if (s == "lsm") run->idm_dim = 0;
    else if (s == "dim") run->idm_dim = 1;
    else {
        c.msg = "Error: Invalid value: idm " + s + " Value must lsm or dim";
        return *c.ierr = TPLS_OPTION_INVALID_ERR;
    }
    if (run->idm_dim) {
        // get_dim_configuration (configure.F90:404-430)
        CFG(c.getd("pe", &run->Pe));
        CFG(c.getd("epn", &run->epn));
    }
    return 0;
}

// tpls_config.opt: main_ns_hybrid.F90:290-309.
int run_options(Cfg& c, tpls_params* prm, tpls_run_config* run)
{
    // get_process_configuration (configure.F90:508-602) for one process grid
    // holding the whole domain: the grid's own checks, not the process count
    CFG(c.getl("auto_decomp", &run->auto_decomp));
    if (!run->auto_decomp) {
        CFG(c.geti("num_procs_x", &run->num_procs_x, 1));
        CFG(c.geti("num_procs_y", &run->num_procs_y, 1));
        char b[160];
        if ((run->maxl - 1) % run->num_procs_x != 0) {
            std::snprintf(b, sizeof b, "Error: num_procs_x must divide maxl - 1. %6d does not divide %6d",
                          run->num_procs_x, run->maxl - 1);
            c.msg = b;
            return *c.ierr = TPLS_OPTION_INVALID_ERR;
        }
        if ((run->maxm - 1) % run->num_procs_y != 0) {
            std::snprintf(b, sizeof b, "Error: num_procs_y must divide maxm - 1. %6d does not divide %6d",
                          run->num_procs_y, run->maxm - 1);
            c.msg = b;
            return *c.ierr = TPLS_OPTION_INVALID_ERR;
        }
    } else {
        run->num_procs_x = run->num_procs_y = 1;
    }
    // get_solver_configuration (configure.F90:613-686)
    CFG(c.geti("mom_u", &prm->mom_u, 1));
    CFG(c.geti("mom_v", &prm->mom_v, 1));
    CFG(c.geti("mom_w", &prm->mom_w, 1));
    CFG(c.geti("levelset", &prm->levelset_its, 1));
    CFG(c.geti("num_timesteps", &run->n_timesteps, 1));
    // get_output_configuration (configure.F90:695-764)
    CFG(c.geti("phi_dat_frequency", &run->phi_dat_frequency, 1));
    CFG(c.getl("backup_text_format", &run->backup_text_format));
    CFG(c.getl("backup_hdf5_format", &run->backup_hdf5_format));
    CFG(c.geti("uvw_dat_frequency", &run->uvw_dat_frequency, 1));
    CFG(c.geti("backup_frequency", &run->backup_frequency, 1));
    // get_misc_configuration (configure.F90:471-489)
    CFG(c.getd("maxu", &run->maxu));
    // get_dim_solver_configuration (configure.F90:438-463)
    if (run->idm_dim) CFG(c.geti("max_iteration_dim", &run->max_iteration_dim, 1));
    return 0;
}

std::string join_path(const char* dir, const char* name)
{
    // get_filepath (io.F90:25-31): trim(directory)//'/'//trim(filename)
    return rtrim(dir) + "/" + rtrim(name);
}

struct OptionsHolder {
    tpls_options* o = nullptr;
    ~OptionsHolder() { tpls_options_destroy(o); }
};

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int tpls_save_data(const char* filename, const char* description, const double* data, int lx, int ly, int mx,
                   int my, int nx, int ny)
{
    if (!filename || !description) return set_error(TPLS_ERR_ARG, "null argument");
    const int b[6] = {lx, ly, mx, my, nx, ny};
    const long count = bound_count(b);
    if (count > 0 && !data) return set_error(TPLS_ERR_ARG, "null data");
    FILE* fp = std::fopen(filename, "w");
    if (!fp) return set_error(TPLS_ERR_IO, std::string(filename) + ": " + std::strerror(errno));
    char hdr[512];
    std::string head = std::string(" DATA: ") + description + "\n DIMENSIONS: Lmin Lmax Mmin Mmax Nmin Nmax\n";
    std::snprintf(hdr, sizeof hdr, "%12d%12d%12d%12d%12d%12d\n", lx, ly, mx, my, nx, ny);
    head += hdr;
    int rc = write_all(fp, head.data(), head.size(), filename);
    if (!rc) rc = write_records(fp, filename, count, 1, [&](long r, double* v) { v[0] = data[r]; });
    if (std::fclose(fp) != 0 && !rc) rc = set_error(TPLS_ERR_IO, std::string(filename) + ": close failed");
    return rc;
}

int tpls_data_bounds(const char* filename, int bounds[6])
{
    if (!filename || !bounds) return set_error(TPLS_ERR_ARG, "null argument");
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    DataHeader h;
    if ((rc = read_header(f, filename, &h))) return rc;
    std::memcpy(bounds, h.b, sizeof h.b);
    return TPLS_OK;
}

int tpls_load_data(const char* filename, double** data, int bounds[6], int* ierr)
{
    if (!filename || !data || !bounds || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    *data = nullptr;
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    DataHeader h;
    if ((rc = read_header(f, filename, &h))) return rc;
    const long count = bound_count(h.b);
    double* a = static_cast<double*>(std::calloc(size_t(std::max(count, 1L)), sizeof(double)));
    if (!a) return set_error(TPLS_ERR_NOMEM, "load_data: host allocation failed");
    long got = read_values(h.body, f.p + f.n, filename, a, count);
    if (got < 0) {
        std::free(a);
        return int(got);
    }
    std::memcpy(bounds, h.b, sizeof h.b);
    *data = a;
    *ierr = got < count ? 1 : 0;
    return TPLS_OK;
}

int tpls_load_data_into_array(const char* filename, double* data, int lx, int ly, int mx, int my, int nx, int ny,
                              int* ierr)
{
    if (!filename || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    DataHeader h;
    if ((rc = read_header(f, filename, &h))) return rc;
    if (h.b[0] != lx || h.b[1] != ly) {
        *ierr = 1;
        return TPLS_OK;
    }
    if (h.b[2] != mx || h.b[3] != my) {
        *ierr = 2;
        return TPLS_OK;
    }
    if (h.b[4] != nx || h.b[5] != ny) {
        *ierr = 3;
        return TPLS_OK;
    }
    const long count = bound_count(h.b);
    if (count > 0 && !data) return set_error(TPLS_ERR_ARG, "null data");
    long got = read_values(h.body, f.p + f.n, filename, data, count);
    if (got < 0) return int(got);
    if (got < count)
        return set_error(TPLS_ERR_FORMAT, std::string(filename) + ": holds " + std::to_string(got) + " of " +
                                              std::to_string(count) + " values");
    *ierr = 0;
    return TPLS_OK;
}

void tpls_free(void* p) { std::free(p); }

// ---- channel outputs and text backups ----

namespace {

FILE* open_write(const char* filename, int* rc)
{
    FILE* fp = std::fopen(filename, "w");
    *rc = fp ? TPLS_OK : set_error(TPLS_ERR_IO, std::string(filename) + ": " + std::strerror(errno));
    return fp;
}

int close_file(FILE* fp, const char* filename, int rc)
{
    if (std::fclose(fp) != 0 && !rc) rc = set_error(TPLS_ERR_IO, std::string(filename) + ": close failed");
    return rc;
}

// The Tecplot zone line the channel outputs start with.
std::string zone_line(int l, int m, int n, bool with_j)
{
    std::string s = " ZONE T=\"Floor\", I= " + std::to_string(l - 1) + " ";
    if (with_j) s += " J= " + std::to_string(m - 1) + " ";
    return s + " K= " + std::to_string(n - 1) + "  F=POINT\n";
}

bool grid_ok(int l, int m, int n) { return l >= 2 && m >= 2 && n >= 2; }

}  // namespace

int tpls_save_backup_info(const char* filename, int lx, int ly, int mx, int my, int nx, int ny, int dx, int dy,
                          int dz)
{
    if (!filename) return set_error(TPLS_ERR_ARG, "null argument");
    int rc;
    FILE* fp = open_write(filename, &rc);
    if (!fp) return rc;
    char b[256];
    std::snprintf(b, sizeof b, " DIMENSIONS: Lmin Lmax Mmin Mmax Nmin Nmax\n %d %d %d %d %d %d\n DIMENSIONS: Dx,Dy,Dz\n %d %d %d\n",
                  lx, ly, mx, my, nx, ny, dx, dy, dz);
    rc = write_all(fp, b, std::strlen(b), filename);
    return close_file(fp, filename, rc);
}

int tpls_channel_output_phi(const double* phi, int l, int m, int n, double dx, double dy, double dz,
                            const char* filename)
{
    if (!phi || !filename || !grid_ok(l, m, n)) return set_error(TPLS_ERR_ARG, "bad argument");
    int rc;
    FILE* fp = open_write(filename, &rc);
    if (!fp) return rc;
    const std::string head = " VARIABLES=\"X \",\"Y \",\"Z \",\"Phi\"\n" + zone_line(l, m, n, true);
    rc = write_all(fp, head.data(), head.size(), filename);
    const long ni = l - 1, nj = m - 1, nk = n - 1, sx = l + 1, sy = long(m + 1) * sx;
    if (!rc)
        rc = write_records(fp, filename, ni * nj * nk, 4, [&](long r, double* v) {
            const long i = r % ni + 1, j = (r / ni) % nj, k = r / (ni * nj);
            v[0] = dx / 2.0 + dx * double(i);
            v[1] = dy / 2.0 + dy * double(j);
            v[2] = dz / 2.0 + dz * double(k);
            v[3] = phi[(i + 1) + j * sx + (k + 1) * sy];
        });
    return close_file(fp, filename, rc);
}

int tpls_channel_input_phi(const char* filename, double* phi, int l, int m, int n, int* ierr)
{
    if (!filename || !phi || !ierr || !grid_ok(l, m, n)) return set_error(TPLS_ERR_ARG, "bad argument");
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    const char* p = f.p;
    const char* e = f.p + f.n;
    for (int r = 0; r < 2 && p < e; ++r) p = std::min(record_end(p, e) + 1, e);
    const long ni = l - 1, nj = m - 1, nk = n - 1, sx = l + 1, sy = long(m + 1) * sx;
    const long count = ni * nj * nk;
    long got = read_records(p, e, filename, 4, count, [&](long r, const double* v, const bool* g) {
        const long i = r % ni + 1, j = (r / ni) % nj, k = r / (ni * nj);
        if (g[3]) phi[(i + 1) + j * sx + (k + 1) * sy] = v[3];
    });
    if (got < 0) return int(got);
    *ierr = got < count ? 1 : 0;
    return TPLS_OK;
}

int tpls_channel_output_uvw(const double* u3, const double* v3, const double* w3, const double* p, int l, int m,
                            int n, double dx, double dy, double dz, const char* filename)
{
    if (!u3 || !v3 || !w3 || !p || !filename || !grid_ok(l, m, n)) return set_error(TPLS_ERR_ARG, "bad argument");
    int rc;
    FILE* fp = open_write(filename, &rc);
    if (!fp) return rc;
    const std::string head = " VARIABLES=\"X \",\"Y \",\"Z \",\"U\",\"V\",\"W\",\"P\"\n" + zone_line(l, m, n, true);
    rc = write_all(fp, head.data(), head.size(), filename);
    const long ni = l - 1, nj = m - 1, nk = n - 1, sx = l + 1, sy = long(m + 1) * sx;
    if (!rc)
        rc = write_records(fp, filename, ni * nj * nk, 7, [&](long r, double* v) {
            const long i = r % ni + 1, j = (r / ni) % nj, k = r / (ni * nj);
            const long c = i + j * sx + k * sy;  // u3, v3 (0:n-2) and w3 (0:n-1) share the plane stride
            v[0] = dx / 2.0 + dx * double(i);
            v[1] = dy / 2.0 + dy * double(j);
            v[2] = dz / 2.0 + dz * double(k);
            v[3] = (u3[c] + u3[c + 1]) / 2.0;
            v[4] = (v3[c] + v3[c + sx]) / 2.0;
            v[5] = (w3[c] + w3[c + sy]) / 2.0;
            v[6] = p[c + sy];
        });
    return close_file(fp, filename, rc);
}

int tpls_channel_output2d(const double* u3, const double* v3, const double* w3, const double* p, const double* phi,
                          const double* visc, int l, int m, int n, double dx, double dy, double dz, const char* filename)
{
    if (!u3 || !v3 || !w3 || !p || !phi || !visc || !filename || !grid_ok(l, m, n))
        return set_error(TPLS_ERR_ARG, "bad argument");
    (void)dy;
    int rc;
    FILE* fp = open_write(filename, &rc);
    if (!fp) return rc;
    const std::string head =
        " VARIABLES=\"X \",\"Z \",\"U\",\"V\",\"W\",\"P\",\"fi\",\"Viscosity\"\n" + zone_line(l, m, n, false);
    rc = write_all(fp, head.data(), head.size(), filename);
    const long ni = l - 1, nk = n - 1, sx = l + 1, sy = long(m + 1) * sx;
    const long j = (m - 1) / 2;  // the mid-channel plane
    if (!rc)
        rc = write_records(fp, filename, ni * nk, 8, [&](long r, double* v) {
            const long i = r % ni + 1, k = r / ni;
            const long c = i + j * sx + k * sy;
            v[0] = dx / 2.0 + dx * double(i);
            v[1] = dz / 2.0 + dz * double(k);
            v[2] = (u3[c] + u3[c + 1]) / 2.0;
            v[3] = (v3[c] + v3[c + sx]) / 2.0;
            v[4] = (w3[c] + w3[c + sy]) / 2.0;
            v[5] = p[c + sy];
            v[6] = phi[c + 1 + sy];
            v[7] = visc[c + sy];
        });
    return close_file(fp, filename, rc);
}

namespace {

// The five blocks of backup_channel / dataload_channel (io.F90:647-791):
// u, v and w with their two convection histories, phi with its two, then
// pressure and pres_old, each block over the whole array, i fastest.
struct BackupBlock {
    int ncols;
    long count;
    double* a[3];
};

std::vector<BackupBlock> backup_blocks(double* u2, double* v2, double* w2, double* c0u, double* c1u, double* c0v,
                                       double* c1v, double* c0w, double* c1w, double* pold, double* pres, double* phi,
                                       double* c0p, double* c1p, int l, int m, int n)
{
    const long plane = long(l + 1) * (m + 1);
    return {{3, plane * (n - 1), {u2, c0u, c1u}},
            {3, plane * (n - 1), {v2, c0v, c1v}},
            {3, plane * n, {w2, c0w, c1w}},
            {3, plane * (n + 1), {phi, c0p, c1p}},
            {2, plane * (n + 1), {pres, pold, nullptr}}};
}

}  // namespace

int tpls_backup_channel(const double* u2, const double* v2, const double* w2, const double* conv0_u,
                        const double* conv1_u, const double* conv0_v, const double* conv1_v, const double* conv0_w,
                        const double* conv1_w, const double* pres_old, const double* pres, const double* phi,
                        const double* conv0_phi, const double* conv1_phi, int l, int m, int n, const char* filename)
{
    const double* all[] = {u2, v2, w2, conv0_u, conv1_u, conv0_v, conv1_v, conv0_w, conv1_w,
                           pres_old, pres, phi, conv0_phi, conv1_phi};
    for (const double* a : all)
        if (!a) return set_error(TPLS_ERR_ARG, "null array");
    if (!filename || !grid_ok(l, m, n)) return set_error(TPLS_ERR_ARG, "bad argument");
    auto blocks = backup_blocks(const_cast<double*>(u2), const_cast<double*>(v2), const_cast<double*>(w2),
                                const_cast<double*>(conv0_u), const_cast<double*>(conv1_u),
                                const_cast<double*>(conv0_v), const_cast<double*>(conv1_v),
                                const_cast<double*>(conv0_w), const_cast<double*>(conv1_w),
                                const_cast<double*>(pres_old), const_cast<double*>(pres), const_cast<double*>(phi),
                                const_cast<double*>(conv0_phi), const_cast<double*>(conv1_phi), l, m, n);
    int rc;
    FILE* fp = open_write(filename, &rc);
    if (!fp) return rc;
    for (const auto& b : blocks) {
        if (rc) break;
        rc = write_records(fp, filename, b.count, b.ncols, [&](long r, double* v) {
            for (int c = 0; c < b.ncols; ++c) v[c] = b.a[c][r];
        });
    }
    return close_file(fp, filename, rc);
}

int tpls_dataload_channel(double* u2, double* v2, double* w2, double* conv0_u, double* conv1_u, double* conv0_v,
                          double* conv1_v, double* conv0_w, double* conv1_w, double* pres_old, double* pres,
                          double* phi, double* conv0_phi, double* conv1_phi, int l, int m, int n, const char* filename)
{
    double* all[] = {u2, v2, w2, conv0_u, conv1_u, conv0_v, conv1_v, conv0_w, conv1_w,
                     pres_old, pres, phi, conv0_phi, conv1_phi};
    for (double* a : all)
        if (!a) return set_error(TPLS_ERR_ARG, "null array");
    if (!filename || !grid_ok(l, m, n)) return set_error(TPLS_ERR_ARG, "bad argument");
    auto blocks = backup_blocks(u2, v2, w2, conv0_u, conv1_u, conv0_v, conv1_v, conv0_w, conv1_w, pres_old, pres, phi,
                                conv0_phi, conv1_phi, l, m, n);
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    // the blocks follow each other with no separator: find where each starts
    // by counting value records, then parse the blocks in turn
    const char* p = f.p;
    const char* e = f.p + f.n;
    for (const auto& b : blocks) {
        const char* q = p;
        long seen = 0;
        while (q < e && seen < b.count) {
            const char* re = record_end(q, e);
            for (const char* c = q; c < re; ++c)
                if (!is_blank(*c)) {
                    ++seen;
                    break;
                }
            q = std::min(re + 1, e);
        }
        if (seen < b.count)
            return set_error(TPLS_ERR_FORMAT, std::string(filename) + ": ends inside a backup block");
        long got = read_records(p, q, filename, b.ncols, b.count, [&](long r, const double* v, const bool* g) {
            for (int c = 0; c < b.ncols; ++c)
                if (g[c]) b.a[c][r] = v[c];
        });
        if (got < 0) return int(got);
        p = q;
    }
    return TPLS_OK;
}

// ---- options ----

int tpls_options_create(tpls_options** out)
{
    if (!out) return set_error(TPLS_ERR_ARG, "null argument");
    *out = new tpls_options();
    return TPLS_OK;
}

int tpls_options_destroy(tpls_options* o)
{
    delete o;
    return TPLS_OK;
}

int tpls_options_count(const tpls_options* o) { return o ? int(o->kv.size()) : 0; }

int tpls_options_entry(const tpls_options* o, int i, const char** name, const char** value)
{
    if (!o || !name || !value || i < 0 || i >= int(o->kv.size())) return set_error(TPLS_ERR_ARG, "bad argument");
    *name = o->kv[i].first.c_str();
    *value = o->kv[i].second.c_str();
    return TPLS_OK;
}

int tpls_has_option(const tpls_options* o, const char* name)
{
    return o && name && find_option(o, name) ? 1 : 0;
}

int tpls_get_option(const tpls_options* o, const char* name, char* value, size_t len, int* ierr)
{
    if (!o || !name || !value || !len || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    const auto* kv = find_option(o, name);
    if (!kv) {
        *ierr = TPLS_OPTION_NOT_FOUND_ERR;
        return TPLS_OK;
    }
    const size_t n = std::min(len - 1, kv->second.size());
    std::memcpy(value, kv->second.data(), n);
    value[n] = 0;
    *ierr = 0;
    return TPLS_OK;
}

int tpls_get_integer_option(const tpls_options* o, const char* name, int* value, int* ierr)
{
    std::string s;
    int rc = typed_value(o, name, &s, ierr);
    if (rc || *ierr) return rc;
    if (!value) return set_error(TPLS_ERR_ARG, "null argument");
    const char* p = s.data();
    const char *tb, *te;
    Item it = next_item(p, s.data() + s.size(), &tb, &te);
    if (it == ITEM_NULL || it == ITEM_END) return TPLS_OK;
    long v;
    if (it != ITEM_VALUE || !parse_int(tb, te, &v) || v < -2147483648L || v > 2147483647L)
        return bad_value(name, s, "an integer", ierr);
    *value = int(v);
    return TPLS_OK;
}

int tpls_get_double_option(const tpls_options* o, const char* name, double* value, int* ierr)
{
    std::string s;
    int rc = typed_value(o, name, &s, ierr);
    if (rc || *ierr) return rc;
    if (!value) return set_error(TPLS_ERR_ARG, "null argument");
    const char* p = s.data();
    const char *tb, *te;
    Item it = next_item(p, s.data() + s.size(), &tb, &te);
    if (it == ITEM_NULL || it == ITEM_END) return TPLS_OK;
    if (it != ITEM_VALUE || !parse_real(tb, te, value)) return bad_value(name, s, "a real number", ierr);
    return TPLS_OK;
}

int tpls_get_logical_option(const tpls_options* o, const char* name, int* value, int* ierr)
{
    std::string s;
    int rc = typed_value(o, name, &s, ierr);
    if (rc || *ierr) return rc;
    if (!value) return set_error(TPLS_ERR_ARG, "null argument");
    const char* p = s.data();
    const char *tb, *te;
    Item it = next_item(p, s.data() + s.size(), &tb, &te);
    if (it == ITEM_NULL || it == ITEM_END) return TPLS_OK;
    if (it != ITEM_VALUE || !parse_logical(tb, te, value)) return bad_value(name, s, "a logical", ierr);
    return TPLS_OK;
}

int tpls_put_option(tpls_options* o, const char* name, const char* value)
{
    if (!o || !name || !value) return set_error(TPLS_ERR_ARG, "null argument");
    put(o, name, value);
    return TPLS_OK;
}

int tpls_put_integer_option(tpls_options* o, const char* name, int value)
{
    if (!o || !name) return set_error(TPLS_ERR_ARG, "null argument");
    put(o, name, i10(value));
    return TPLS_OK;
}

int tpls_put_double_option(tpls_options* o, const char* name, double value)
{
    if (!o || !name) return set_error(TPLS_ERR_ARG, "null argument");
    put(o, name, f20_10(value));
    return TPLS_OK;
}

int tpls_put_logical_option(tpls_options* o, const char* name, int value)
{
    if (!o || !name) return set_error(TPLS_ERR_ARG, "null argument");
    put(o, name, value ? "T" : "F");
    return TPLS_OK;
}

int tpls_load_options(const char* filename, tpls_options** out, int* ierr)
{
    if (!filename || !out || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!file_exists(filename)) {
        *ierr = TPLS_OPTION_FILE_NOT_FOUND_ERR;
        return TPLS_OK;
    }
    auto* o = new tpls_options();
    int rc = load_options_file(filename, o, ierr);
    if (rc) {
        delete o;
        return rc;
    }
    *out = o;
    return TPLS_OK;
}

int tpls_save_options(const char* filename, const tpls_options* o, const char* const* hdr, int num_hdr_lines)
{
    if (!filename || !o || (num_hdr_lines > 0 && !hdr)) return set_error(TPLS_ERR_ARG, "null argument");
    FILE* fp = std::fopen(filename, "w");
    if (!fp) return set_error(TPLS_ERR_IO, std::string(filename) + ": " + std::strerror(errno));
    std::string s;
    for (int i = 0; i < num_hdr_lines; ++i) s += "# " + rtrim(hdr[i] ? hdr[i] : "") + "\n";
    int rc = write_all(fp, s.data(), s.size(), filename);
    if (std::fclose(fp) != 0 && !rc) rc = set_error(TPLS_ERR_IO, std::string(filename) + ": close failed");
    return rc ? rc : append_options_file(filename, o, "a");
}

int tpls_append_options(const char* filename, const tpls_options* o)
{
    if (!filename || !o) return set_error(TPLS_ERR_ARG, "null argument");
    // status='OLD', position='append': the file must exist
    if (!file_exists(filename)) return set_error(TPLS_ERR_IO, std::string(filename) + ": no such file");
    return append_options_file(filename, o, "a");
}

// ---- configuration ----

int tpls_read_config(const char* config_dir, const char* options_file, tpls_params* prm, tpls_run_config* run,
                     int* ierr)
{
    if (!config_dir || !options_file || !prm || !run || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    *ierr = 0;
    std::memset(run, 0, sizeof *run);
    auto fail = [&](int code, const std::string& msg) {
        *ierr = code;
        return set_error(TPLS_ERR_CONFIG, msg);
    };
    // get_configuration (configure.F90:39-88)
    if (!file_exists(config_dir))
        return fail(TPLS_CONFIG_DIR_NOT_FOUND_ERR,
                    "Error: Configuration directory not found: " + rtrim(config_dir));
    OptionsHolder opts, init;
    int e = 0;
    int rc = tpls_load_options(options_file, &opts.o, &e);
    if (rc) return rc;
    if (e == TPLS_OPTION_FILE_NOT_FOUND_ERR)
        return fail(e, "Error: Options file not found: " + rtrim(options_file));
    if (e) return fail(e, "Error: Loading options file: " + rtrim(options_file));
    // initial_config.dat (main_ns_hybrid.F90:250-261)
    const std::string ic = join_path(config_dir, "initial_config.dat");
    if ((rc = tpls_load_options(ic.c_str(), &init.o, &e))) return rc;
    if (e == TPLS_OPTION_FILE_NOT_FOUND_ERR)
        return fail(TPLS_CONFIG_FILE_NOT_FOUND_ERR, "Error: Initial conditions configuration file not found: " + ic);
    if (e) return fail(e, "Error: Loading initial conditions configurations file: " + ic);
    Cfg c{init.o, ierr, {}};
    if ((rc = initial_conditions(c, prm, run))) return rc < 0 ? rc : set_error(TPLS_ERR_CONFIG, c.msg);
    c.o = opts.o;
    if ((rc = run_options(c, prm, run))) return rc < 0 ? rc : set_error(TPLS_ERR_CONFIG, c.msg);
    // the interface method and its parameters, as the case reads them
    prm->idm_dim = run->idm_dim;
    prm->epn = run->idm_dim ? run->epn : 0.0;
    prm->Pe = run->idm_dim ? run->Pe : 0.0;
    prm->dim_its = run->idm_dim ? run->max_iteration_dim : 0;
    *ierr = 0;
    return TPLS_OK;
}

int tpls_case_load_config(tpls_case* c, const char* config_dir, int* ierr)
{
    if (!c || !config_dir || !ierr) return set_error(TPLS_ERR_ARG, "null argument");
    int maxl, maxm, maxn;
    int rc = tpls::case_dims(c, &maxl, &maxm, &maxn);
    if (rc) return rc;
    struct F {
        const char* file;
        int nz;  // upper n bound
    } files[6] = {{"initial_u.dat", maxn - 2},        {"initial_v.dat", maxn - 2},
                  {"initial_w.dat", maxn - 1},        {"initial_pressure.dat", maxn},
                  {"initial_phi.dat", maxn},          {"initial_viscosity.dat", maxn}};
    // every value is parsed or the load fails: no need to clear the buffers
    std::unique_ptr<double[]> a[6];
    const long plane = long(maxl + 1) * (maxm + 1);
    // load_configuration_data (configure.F90:94-122) for each file
    for (int q = 0; q < 6; ++q) {
        const std::string path = join_path(config_dir, files[q].file);
        if (!file_exists(path.c_str())) {
            *ierr = TPLS_CONFIG_FILE_NOT_FOUND_ERR;
            return set_error(TPLS_ERR_CONFIG, "Error: Configuration file not found: " + path);
        }
        a[q].reset(new (std::nothrow) double[size_t(plane * (files[q].nz + 1))]);
        if (!a[q]) return set_error(TPLS_ERR_NOMEM, "load_config: host allocation failed");
        if ((rc = tpls_load_data_into_array(path.c_str(), a[q].get(), 0, maxl, 0, maxm, 0, files[q].nz, ierr)))
            return rc;
        if (*ierr) {
            char b[64];
            std::snprintf(b, sizeof b, "%6d%6d%6d", maxl, maxm, files[q].nz);
            return set_error(TPLS_ERR_CONFIG, "Error: " + path + " data dimensions do not match " + b);
        }
    }
    *ierr = 0;
    return tpls_case_load(c, a[0].get(), a[1].get(), a[2].get(), a[3].get(), a[4].get());
}

}  // extern "C"
