// Host I/O of the BatchLayout file formats: .mtx -> CSR, -initf coordinates, output files.
//
// .mtx rules follow ReadASCII (IO.h:60-155), CSC(Triple*) (CSC.h:147-188) and
// CSR(const CSC&) (CSR.h:155-186):
//   * every leading line that starts with '%' is a comment; "symmetric" anywhere in one
//     of them turns on mirroring (IO.h:67-75);
//   * the size line is "m n nnz"; exactly nnz entry lines are read (IO.h:98, with the
//     count adjusted for dropped diagonals, which amounts to reading nnz lines);
//   * entries are "row col [value]", 1-based; values are ignored by every layout;
//   * symmetric: (r,c) and (c,r) for r != c, diagonal entries dropped (IO.h:122-134);
//     general: entries as-is, self-loops kept; duplicates are never merged;
//   * the CSR has each row's column ids ascending (CSC sorts rows per column and CSR
//     walks the columns in order).
// Differences, all where the reference crashes or reads out of bounds: fields may be
// separated by any blanks (the reference needs single spaces, IO.h:103-106), blank lines
// are skipped, and a short file, an index out of range or m != n is an error.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "batchlayout.h"

namespace {

struct Reader {
    std::vector<char> buf;
    size_t pos = 0;
    bool load(const char *path) {
        FILE *f = fopen(path, "rb");
        if (!f) return false;
        if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return false; }
        long sz = ftell(f);
        if (sz < 0) { fclose(f); return false; }
        rewind(f);
        buf.resize((size_t)sz + 1);
        size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
        fclose(f);
        if (got != (size_t)sz) return false;
        buf[(size_t)sz] = '\0';
        buf.resize((size_t)sz);
        return true;
    }
    bool eof() const { return pos >= buf.size(); }
    // [b, e) of the next line without the newline
    void line(size_t &b, size_t &e) {
        b = pos;
        while (pos < buf.size() && buf[pos] != '\n') ++pos;
        e = pos;
        if (e > b && buf[e - 1] == '\r') --e;
        if (pos < buf.size()) ++pos;
    }
};

inline bool parse_long(const char *&p, const char *end, long &v) {
    while (p < end && (*p == ' ' || *p == '\t')) ++p;
    bool neg = false;
    if (p < end && (*p == '-' || *p == '+')) { neg = *p == '-'; ++p; }
    if (p >= end || *p < '0' || *p > '9') return false;
    long x = 0;
    while (p < end && *p >= '0' && *p <= '9') { x = x * 10 + (*p - '0'); ++p; }
    v = neg ? -x : x;
    return true;
}

// Entries of one chunk of the .mtx body: 0-based (row, col) pairs, interleaved.  `bad`
// marks a malformed or out-of-range line, after which the chunk stops.
struct Chunk {
    std::vector<int32_t> rc;
    bool bad = false;
};

void parse_chunk(const char *buf, size_t b, size_t e, long m, Chunk &out) {
    out.rc.reserve((e - b) / 6 * 2 + 2);
    size_t pos = b;
    while (pos < e) {
        const char *p = buf + pos;
        const char *nl = (const char *)memchr(p, '\n', e - pos);
        const char *pe = nl ? nl : buf + e;
        pos = (size_t)(pe - buf) + (nl ? 1 : 0);
        if (pe > p && pe[-1] == '\r') --pe;
        while (p < pe && (*p == ' ' || *p == '\t')) ++p;
        if (p == pe) continue;  // blank line
        long ri, ci;
        if (!parse_long(p, pe, ri) || !parse_long(p, pe, ci) || ri < 1 || ri > m || ci < 1 || ci > m) {
            out.bad = true;
            return;
        }
        out.rc.push_back((int32_t)(ri - 1));
        out.rc.push_back((int32_t)(ci - 1));
    }
}

// Threads for a body of `bytes` bytes: one per 4 MiB, at most the hardware threads and 16;
// BL_IO_THREADS overrides.
int io_threads(size_t bytes) {
    int hw = (int)std::thread::hardware_concurrency();
    if (hw <= 0) hw = 1;
    int t = (int)std::min<size_t>(bytes / (4u << 20) + 1, (size_t)std::min(hw, 16));
    if (const char *s = getenv("BL_IO_THREADS")) {
        const int v = atoi(s);
        if (v > 0) t = std::min(v, 64);
    }
    return std::max(t, 1);
}

template <class F>
void run_threads(int T, F &&f) {
    if (T == 1) { f(0); return; }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (int t = 1; t < T; ++t) th.emplace_back(f, t);
    f(0);
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" void bl_free(void *p) { free(p); }

extern "C" int bl_mtx_read(const char *path, int32_t *n_out, int64_t *nnz_out, int32_t **rowptr_out,
                           int32_t **colids_out) {
    if (!path || !n_out || !nnz_out || !rowptr_out || !colids_out) return BL_E_ARG;
    Reader r;
    if (!r.load(path)) return BL_E_IO;
    bool symmetric = false;
    long m = -1, ncol = -1, nnz = -1;
    size_t b, e;
    for (;;) {
        if (r.eof()) return BL_E_FORMAT;
        r.line(b, e);
        if (e > b && r.buf[b] == '%') {
            if (std::string(r.buf.data() + b, e - b).find("symmetric") != std::string::npos) symmetric = true;
            continue;
        }
        const char *p = r.buf.data() + b, *pe = r.buf.data() + e;
        if (!parse_long(p, pe, m) || !parse_long(p, pe, ncol) || !parse_long(p, pe, nnz)) return BL_E_FORMAT;
        break;
    }
    if (m <= 0 || ncol != m || nnz < 0 || m > INT32_MAX) return BL_E_FORMAT;
    if ((symmetric ? 2 * nnz : nnz) > INT32_MAX) return BL_E_FORMAT;  // CSR offsets are int32

    // Entry lines are parsed by T threads, each over a newline-aligned chunk of the body.
    // Only the first nnz non-blank lines count: a chunk records how many entries it parsed
    // before a bad line, and the bad line is an error only if it falls inside the first nnz.
    const size_t body = r.pos, total = r.buf.size();
    const int T = io_threads(total - body);
    std::vector<size_t> cut(T + 1, total);
    cut[0] = body;
    for (int t = 1; t < T; ++t) {
        size_t c = std::max(cut[t - 1], body + (total - body) * (size_t)t / (size_t)T);
        while (c < total && c > body && r.buf[c - 1] != '\n') ++c;
        cut[t] = c;
    }
    std::vector<Chunk> chunks(T);
    run_threads(T, [&](int t) { parse_chunk(r.buf.data(), cut[t], cut[t + 1], m, chunks[t]); });
    // take chunks in file order until nnz entries
    long need = nnz;
    std::vector<size_t> take(T, 0);
    for (int t = 0; t < T && need > 0; ++t) {
        const long have = (long)chunks[t].rc.size() / 2;
        if (have >= need) { take[t] = (size_t)need; need = 0; break; }
        if (chunks[t].bad) return BL_E_FORMAT;  // malformed or out-of-range line within nnz
        take[t] = (size_t)have;
        need -= have;
    }
    if (need > 0) return BL_E_FORMAT;  // fewer entries than declared (assert, IO.h:142)

    int32_t *rowptr = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    if (!rowptr) return BL_E_NOMEM;
    // row counts (rowptr[i+1]) with relaxed atomics; symmetric drops diagonals and mirrors
    run_threads(T, [&](int t) {
        const int32_t *rc = chunks[t].rc.data();
        for (size_t k = 0; k < take[t]; ++k) {
            const int32_t ri = rc[2 * k], ci = rc[2 * k + 1];
            if (symmetric) {
                if (ri == ci) continue;
                __atomic_fetch_add(&rowptr[ci + 1], 1, __ATOMIC_RELAXED);
            }
            __atomic_fetch_add(&rowptr[ri + 1], 1, __ATOMIC_RELAXED);
        }
    });
    for (long i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    const size_t cnt = (size_t)rowptr[m];
    int32_t *colids = (int32_t *)malloc(sizeof(int32_t) * (cnt > 0 ? cnt : 1));
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    if (!colids || !fill) { free(rowptr); free(colids); free(fill); return BL_E_NOMEM; }
    memcpy(fill, rowptr, sizeof(int32_t) * (size_t)m);
    // scatter in any order: each row is sorted afterwards, and equal column ids are
    // indistinguishable, so the result is the sequential one
    run_threads(T, [&](int t) {
        const int32_t *rc = chunks[t].rc.data();
        for (size_t k = 0; k < take[t]; ++k) {
            const int32_t ri = rc[2 * k], ci = rc[2 * k + 1];
            if (symmetric) {
                if (ri == ci) continue;
                colids[__atomic_fetch_add(&fill[ci], 1, __ATOMIC_RELAXED)] = ri;
            }
            colids[__atomic_fetch_add(&fill[ri], 1, __ATOMIC_RELAXED)] = ci;
        }
    });
    free(fill);
    chunks.clear();
    // sort rows; threads take contiguous row ranges of about cnt/T entries each
    std::vector<long> rcut(T + 1, m);
    rcut[0] = 0;
    for (int t = 1; t < T; ++t) {
        const int32_t target = (int32_t)(cnt * (size_t)t / (size_t)T);
        rcut[t] = std::max(rcut[t - 1], (long)(std::lower_bound(rowptr, rowptr + m, target) - rowptr));
    }
    run_threads(T, [&](int t) {
        for (long i = rcut[t]; i < rcut[t + 1]; ++i) std::sort(colids + rowptr[i], colids + rowptr[i + 1]);
    });
    *n_out = (int32_t)m;
    *nnz_out = (int64_t)cnt;
    *rowptr_out = rowptr;
    *colids_out = colids;
    return BL_OK;
}

// algorithms::fileInitialization (algorithms.cpp:98-119): fgets(line, 256) then
// sscanf("%lf %lf %d"); a line that does not parse leaves x, y at the previous line's
// values (sscanf assigns nothing), so it repeats the previous coordinate.  At most n
// lines are stored (the reference writes past its array); missing lines are zero.
extern "C" int bl_initf_read(const char *path, int32_t n, double *xy, int32_t *lines_read) {
    if (!path || n <= 0 || !xy) return BL_E_ARG;
    FILE *f = fopen(path, "r");
    if (!f) return BL_E_IO;
    char line[256];
    double x = 0.0, y = 0.0;
    int id = 0;
    int32_t k = 0;
    while (k < n && fgets(line, sizeof(line), f)) {
        (void)sscanf(line, "%lf %lf %d", &x, &y, &id);
        xy[2 * k] = x;
        xy[2 * k + 1] = y;
        ++k;
    }
    fclose(f);
    for (int32_t i = k; i < n; ++i) xy[2 * i] = xy[2 * i + 1] = 0.0;
    if (lines_read) *lines_read = k;
    return BL_OK;
}

// algorithms::writeToFile (algorithms.cpp:1244-1256): "x\ty\tid\n", default ostream
// precision (6 significant digits, %g form), ids 1-based.
extern "C" int bl_output_write(const char *path, int32_t n, const double *xy) {
    if (!path || n < 0 || (n > 0 && !xy)) return BL_E_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return BL_E_IO;
    std::vector<char> out;
    out.reserve((size_t)n * 32);
    char tmp[96];
    for (int32_t i = 0; i < n; ++i) {
        int len = snprintf(tmp, sizeof(tmp), "%g\t%g\t%d\n", xy[2 * i], xy[2 * i + 1], i + 1);
        out.insert(out.end(), tmp, tmp + len);
    }
    const bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    return (fclose(f) == 0 && ok) ? BL_OK : BL_E_IO;
}
