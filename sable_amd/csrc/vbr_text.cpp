// Native reader and writer of the `.vbr` / `.vbrc` text lines
// (reference: utils/fileio.py:13-35 writer, :55-93 readers;
// utils/convert_real_to_vbr.py:84-95 the .vbrc writer).
//
// A file is a sequence of lines `name=[a,b,...]`.  The reference parses each
// token with Python's float() / int() and writes each value with str().
// Here the file is memory-mapped, every line's body is cut into per-thread
// chunks at commas, and the chunks are parsed concurrently with
// std::from_chars, which rounds decimal to binary like strtod (correctly
// rounded, as Python's float()), so the arrays are bit-identical to the
// reference's.  The writer formats doubles as Python's str() does (the
// shortest digits that round-trip, fixed notation for decimal exponents in
// [-4, 16) with at least one digit after the point, else d.ddde+XX) and
// integers in decimal, one chunk per thread, then writes the chunks in
// order.  Host code only; no device involved.
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <sched.h>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/sable.h"

namespace {

thread_local std::string t_err;

int text_fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

int default_threads(int nthreads)
{
    if (nthreads > 0) return nthreads;
    cpu_set_t s;
    if (sched_getaffinity(0, sizeof s, &s) == 0) return std::max(1, CPU_COUNT(&s));
    return std::max(1u, std::thread::hardware_concurrency());
}

template <typename F>
void parallel_for(int n, F f)
{
    if (n <= 1) {
        f(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(n - 1);
    for (int i = 1; i < n; ++i) th.emplace_back(f, i);
    f(0);
    for (auto &t : th) t.join();
}

// Chunks smaller than this are not worth a thread.
constexpr size_t kMinChunk = 1 << 20;

struct Line {
    const char *b = nullptr, *e = nullptr;   // body between '[' and ']'
    std::vector<const char *> cut;           // chunk starts (cut[0] = b), chunk ends at the next start or e
    std::vector<int64_t> first;              // index of each chunk's first element
    int64_t n = 0;
};

}  // namespace

struct SableText {
    int fd = -1;
    const char *base = nullptr;
    size_t size = 0;
    int nthreads = 1;
    std::vector<Line> lines;
};

extern "C" {

const char *sable_text_last_error(void)
{
    return t_err.c_str();
}

int sable_text_open(const char *path, int nlines, int nthreads, SableText **out, int64_t *counts)
{
    if (!path || !out || !counts || nlines <= 0) return text_fail(SABLE_ERR_INVALID, "bad arguments");
    *out = nullptr;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return text_fail(SABLE_ERR_INVALID, std::string("cannot open ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size <= 0) {
        close(fd);
        return text_fail(SABLE_ERR_INVALID, std::string("empty or unreadable file ") + path);
    }
    void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
        close(fd);
        return text_fail(SABLE_ERR_INVALID, std::string("cannot map ") + path);
    }
    madvise(m, (size_t)st.st_size, MADV_SEQUENTIAL);
    auto *h = new SableText;
    h->fd = fd;
    h->base = (const char *)m;
    h->size = (size_t)st.st_size;
    h->nthreads = default_threads(nthreads);
    const char *p = h->base, *end = h->base + h->size;
    for (int i = 0; i < nlines; ++i) {
        if (p >= end) {
            sable_text_close(h);
            return text_fail(SABLE_ERR_INVALID, "file has " + std::to_string(i) + " lines, " +
                                                    std::to_string(nlines) + " expected");
        }
        const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        const char *eq = (const char *)memchr(p, '=', (size_t)(le - p));
        const char *rb = le;
        while (rb > p && (rb[-1] == '\r' || rb[-1] == ' ')) --rb;
        if (!eq || eq + 1 >= rb || eq[1] != '[' || rb[-1] != ']') {
            sable_text_close(h);
            return text_fail(SABLE_ERR_INVALID, "line " + std::to_string(i + 1) + " is not name=[...]");
        }
        Line L;
        L.b = eq + 2;
        L.e = rb - 1;
        const size_t len = (size_t)(L.e - L.b);
        const int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)h->nthreads, len / kMinChunk));
        // chunk starts: nominal positions moved past the next comma
        L.cut.push_back(L.b);
        for (int c = 1; c < nt; ++c) {
            const char *q = L.b + len * c / nt;
            const char *cm = (const char *)memchr(q, ',', (size_t)(L.e - q));
            const char *s = cm ? cm + 1 : L.e;
            if (s > L.cut.back() && s < L.e) L.cut.push_back(s);
        }
        const int nc = (int)L.cut.size();
        std::vector<int64_t> commas(nc, 0);
        parallel_for(nc, [&](int c) {
            const char *s = L.cut[c], *t = c + 1 < nc ? L.cut[c + 1] : L.e;
            int64_t k = 0;
            for (const char *q = s; q < t; ++q) k += (*q == ',');
            commas[c] = k;
        });
        // elements: commas + 1 for a non-empty body; chunk c starts at
        // element (commas before its start)
        L.first.assign(nc, 0);
        int64_t tot = 0;
        for (int c = 0; c < nc; ++c) {
            L.first[c] = tot;
            tot += commas[c];
        }
        L.n = len == 0 ? 0 : tot + 1;
        counts[i] = L.n;
        h->lines.push_back(std::move(L));
        p = le + (nl ? 1 : 0);
    }
    *out = h;
    return SABLE_OK;
}

}  // extern "C"

namespace {

template <typename T>
int parse_line(SableText *h, int line, T *out)
{
    if (!h || line < 0 || line >= (int)h->lines.size() || (!out && h->lines[line].n))
        return text_fail(SABLE_ERR_INVALID, "bad arguments");
    const Line &L = h->lines[line];
    if (L.n == 0) return SABLE_OK;
    const int nc = (int)L.cut.size();
    std::vector<int64_t> bad(nc, -1);
    parallel_for(nc, [&](int c) {
        const char *p = L.cut[c], *t = c + 1 < nc ? L.cut[c + 1] : L.e;
        int64_t k = L.first[c];
        const int64_t k1 = c + 1 < nc ? L.first[c + 1] : L.n;
        for (;;) {
            while (p < t && *p == ' ') ++p;
            if (p < t && *p == '+') ++p;
            T v{};
            const auto r = std::from_chars(p, t, v);
            if (r.ec != std::errc() || k >= k1) {
                bad[c] = p - L.b;
                return;
            }
            out[k++] = v;
            p = r.ptr;
            while (p < t && *p == ' ') ++p;
            if (p >= t) break;
            if (*p != ',') {
                bad[c] = p - L.b;
                return;
            }
            ++p;
            // a chunk ends right after a comma: the next chunk holds the rest
            if (p == t && c + 1 < nc) break;
        }
        if (k != k1) bad[c] = p - L.b;
    });
    for (int c = 0; c < nc; ++c)
        if (bad[c] >= 0)
            return text_fail(SABLE_ERR_INVALID, "line " + std::to_string(line + 1) + ": bad token at body offset " +
                                                    std::to_string(bad[c]));
    return SABLE_OK;
}

// Python's str() of a double: shortest round-trip digits; fixed notation when
// the decimal exponent x of the first digit satisfies -4 <= x < 16 (with ".0"
// added when there is no fractional digit), else d[.ddd]e+XX / e-XX with at
// least two exponent digits; "inf", "-inf", "nan".
int format_double(double v, char *o)
{
    if (v != v) {
        memcpy(o, "nan", 3);
        return 3;
    }
    if (v == __builtin_inf() || v == -__builtin_inf()) {
        if (v < 0) {
            memcpy(o, "-inf", 4);
            return 4;
        }
        memcpy(o, "inf", 3);
        return 3;
    }
    char s[40];
    const auto r = std::to_chars(s, s + sizeof s, v, std::chars_format::scientific);
    // s = [-]d[.ddd]e(+|-)XX
    const char *p = s, *end = r.ptr;
    int n = 0;
    if (*p == '-') {
        o[n++] = '-';
        ++p;
    }
    char dig[24];
    int nd = 0;
    const char *q = p;
    for (; q < end && *q != 'e'; ++q)
        if (*q != '.') dig[nd++] = *q;
    int x = 0;
    std::from_chars(q + 1 + (q[1] == '+'), end, x);
    if (x >= -4 && x < 16) {
        if (x < 0) {
            o[n++] = '0';
            o[n++] = '.';
            for (int i = 0; i < -x - 1; ++i) o[n++] = '0';
            for (int i = 0; i < nd; ++i) o[n++] = dig[i];
        } else {
            for (int i = 0; i <= x; ++i) o[n++] = i < nd ? dig[i] : '0';
            o[n++] = '.';
            if (nd > x + 1)
                for (int i = x + 1; i < nd; ++i) o[n++] = dig[i];
            else
                o[n++] = '0';
        }
    } else {
        o[n++] = dig[0];
        if (nd > 1) {
            o[n++] = '.';
            for (int i = 1; i < nd; ++i) o[n++] = dig[i];
        }
        o[n++] = 'e';
        o[n++] = x < 0 ? '-' : '+';
        const int ax = x < 0 ? -x : x;
        if (ax < 10) o[n++] = '0';
        const auto rr = std::to_chars(o + n, o + n + 8, ax);
        n = (int)(rr.ptr - o);
    }
    return n;
}

template <typename T>
int format_value(T v, char *o)
{
    if constexpr (std::is_floating_point_v<T>) {
        return format_double((double)v, o);
    } else {
        const auto r = std::to_chars(o, o + 24, v);
        return (int)(r.ptr - o);
    }
}

template <typename T>
int write_line(FILE *f, const char *name, const T *a, int64_t n, int nthreads)
{
    if (fprintf(f, "%s=[", name) < 0) return text_fail(SABLE_ERR_INVALID, "write failed");
    if (n > 0) {
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n / 65536));
        std::vector<std::string> part(nt);
        parallel_for(nt, [&](int c) {
            const int64_t i0 = n * c / nt, i1 = n * (c + 1) / nt;
            std::string &s = part[c];
            s.resize((size_t)(i1 - i0) * 26);
            size_t w = 0;
            for (int64_t i = i0; i < i1; ++i) {
                if (i) s[w++] = ',';
                w += (size_t)format_value(a[i], &s[w]);
            }
            s.resize(w);
        });
        for (auto &s : part)
            if (fwrite(s.data(), 1, s.size(), f) != s.size()) return text_fail(SABLE_ERR_INVALID, "write failed");
    }
    if (fputs("]\n", f) < 0) return text_fail(SABLE_ERR_INVALID, "write failed");
    return SABLE_OK;
}

}  // namespace

extern "C" {

int sable_text_read_f64(SableText *h, int line, double *out)
{
    return parse_line<double>(h, line, out);
}

int sable_text_read_i64(SableText *h, int line, int64_t *out)
{
    return parse_line<int64_t>(h, line, out);
}

void sable_text_close(SableText *h)
{
    if (!h) return;
    if (h->base) munmap((void *)h->base, h->size);
    if (h->fd >= 0) close(h->fd);
    delete h;
}

int sable_text_write(const char *path, int nlines, const char *const *names, const int32_t *kinds,
                     const void *const *arrays, const int64_t *counts, int append, int nthreads)
{
    if (!path || nlines <= 0 || !names || !kinds || !arrays || !counts) return text_fail(SABLE_ERR_INVALID, "bad arguments");
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) return text_fail(SABLE_ERR_INVALID, std::string("cannot open ") + path + " for writing");
    const int nt = default_threads(nthreads);
    int rc = SABLE_OK;
    for (int i = 0; i < nlines && rc == SABLE_OK; ++i) {
        if (counts[i] < 0 || (counts[i] > 0 && !arrays[i])) rc = text_fail(SABLE_ERR_INVALID, "bad array");
        else if (kinds[i] == SABLE_TEXT_F64) rc = write_line(f, names[i], (const double *)arrays[i], counts[i], nt);
        else if (kinds[i] == SABLE_TEXT_I64) rc = write_line(f, names[i], (const int64_t *)arrays[i], counts[i], nt);
        else rc = text_fail(SABLE_ERR_INVALID, "unknown element kind");
    }
    if (fclose(f) != 0 && rc == SABLE_OK) rc = text_fail(SABLE_ERR_INVALID, "write failed");
    return rc;
}

}  // extern "C"
