// Barnes-Hut path of BatchLayout (-algo 4) on CDNA4 (gfx950).
//
// Reference: sample/algorithms.cpp:1020-1083 (the layout loop), sample/MortonCode.h:42-97
// (bounding box, 32-bit scaling, y-over-x Morton order), sample/BarnesHut.h:114-220 (sort
// of the coordinates and recursive quadtree build, unweighted centroids) and :69-86
// (traversal: an accepted cell or a leaf contributes a unit vector).
//
// Per iteration, from the positions at the start of the iteration (:1041-1047):
//   1. k_frame_*   min x, min y, MAXMIN = max(dy, dx)                       (FP64, exact)
//   2. k_keys      64-bit key = interleave(yscale, xscale), y bit above x bit
//   3. radix sort  LSD, 6 passes of 8 bits over (key, vertex id), stable: per-tile digit
//                  counts, a scan per digit, a scatter staged through LDS; the keys' low 16
//                  bits order only runs of keys equal above them, sorted in place (k_sort_ties)
//   4. gather      positions in Morton order (the last scatter pass writes them)
//   5. tree        node = cell where two adjacent keys first differ (base-4 digit d of
//                  k_i ^ k_i+1); its range, quadrants and children come from galloping and
//                  binary searches on the sorted keys -- the reference's tree with single-child
//                  chains collapsed into their lowest node (a chain shares one centroid,
//                  and the smallest diameter in the chain decides acceptance, so the
//                  traversal result is the same)
//   6. k_centroid_* unweighted mean of the non-empty quadrants' centroids, bottom-up,
//                  summed in quadrant order as :197-217
//   7. k_rep*      repulsion of every vertex, one wave per 64 Morton-consecutive vertices
//                  with a shared stack of (node, lane mask): the vertex's own position is
//                  the iteration-start one (it only moves in its own batch step), so all
//                  n traversals run in one launch
//   8. k_attract_move  one workgroup walks the ceil(n/B) batches in order: attraction
//                  over every CSR entry with live positions (:1055-1058), minus the
//                  repulsion, then the unit-step move (:1062-1066)
//   9. energy      sum of |F_i|^2
//
// FP32 mode evaluates the opening test D/|c-v| < theta in FP32 with a relative guard
// band and settles the borderline cases in FP64, so the set of accepted cells is the
// FP64 one; the unit vectors are summed in FP32.  FP64 mode visits cells in the
// reference's depth-first quadrant order and uses its FP64 operations, so the force is
// the reference's bit for bit.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "bh_common.h"

namespace bl {

// out: 4 + 256 doubles, the sum in out[0] (minibatch.hip)
int launch_energy(const double *eterm, int n, double *out, int precision, hipStream_t s);

namespace {

constexpr int WG = 256;
// Keys per thread per sort tile (tile = WG x SORT_ROUNDS keys, staged in LDS by the scatter:
// 30 KB).  The per-digit row scan takes any number of tiles.  On 1.66M keys a pass took
// 14.6 us with 8 rounds and with 16, against 27.5 us unstaged with 16 rounds and 34 us
// unstaged with 4.
constexpr int SORT_ROUNDS = 8;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = WG * SCAN_ITEMS;     // 2048 entries per workgroup
constexpr int STACK = 128;                     // per-wave traversal stack (depth <= 32 -> <= 100 entries)



// ---------------------------------------------------------------------------------------
// 1. bounding box
__global__ __launch_bounds__(WG) void k_frame_partial(const double2 *__restrict__ pos, int n,
                                                      double4 *__restrict__ part) {
    __shared__ double4 s[WG];
    double4 a = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);  // minx maxx miny maxy
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const double2 p = pos[i];
        a.x = fmin(a.x, p.x);
        a.y = fmax(a.y, p.x);
        a.z = fmin(a.z, p.y);
        a.w = fmax(a.w, p.y);
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double4 b = s[threadIdx.x + w];
            double4 &c = s[threadIdx.x];
            c.x = fmin(c.x, b.x);
            c.y = fmax(c.y, b.y);
            c.z = fmin(c.z, b.z);
            c.w = fmax(c.w, b.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// frame[0] = minX, frame[1] = minY, frame[2] = MAXMIN = max(maxY-minY, maxX-minX) (:45-51)
__global__ __launch_bounds__(WG) void k_frame_final(const double4 *__restrict__ part, int np, double *frame) {
    __shared__ double4 s[WG];
    double4 a = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int i = threadIdx.x; i < np; i += WG) {
        const double4 b = part[i];
        a.x = fmin(a.x, b.x);
        a.y = fmax(a.y, b.y);
        a.z = fmin(a.z, b.z);
        a.w = fmax(a.w, b.w);
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double4 b = s[threadIdx.x + w];
            double4 &c = s[threadIdx.x];
            c.x = fmin(c.x, b.x);
            c.y = fmax(c.y, b.y);
            c.z = fmin(c.z, b.z);
            c.w = fmax(c.w, b.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dy = s[0].w - s[0].z, dx = s[0].y - s[0].x;
        frame[0] = s[0].x;
        frame[1] = s[0].z;
        frame[2] = (dy < dx) ? dx : dy;
    }
}

// ---------------------------------------------------------------------------------------
// 2. Morton keys.  xscale(x) = (uint32_t)(0xffffffff * (x - minX) / MAXMIN), MortonCode.h:62-67.
__device__ __forceinline__ uint32_t scale32(double v, double mn, double maxmin) {
    double t = 4294967295.0 * (v - mn);
    t = t / maxmin;
    return (uint32_t)t;
}

__device__ __forceinline__ uint64_t spread(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__global__ __launch_bounds__(WG) void k_keys(const double2 *__restrict__ pos, int n, const double *__restrict__ frame,
                                             uint64_t *__restrict__ keys, int *__restrict__ ids,
                                             uint32_t *__restrict__ xs, uint32_t *__restrict__ ys) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    const double mnx = frame[0], mny = frame[1], mm = frame[2];
    const double2 p = pos[i];
    const uint32_t sx = scale32(p.x, mnx, mm), sy = scale32(p.y, mny, mm);
    keys[i] = (spread(sy) << 1) | spread(sx);
    ids[i] = i;
    if (xs) {
        xs[i] = sx;
        ys[i] = sy;
    }
}

// ---------------------------------------------------------------------------------------
// generic exclusive scan of int32 (3 phases)
__device__ __forceinline__ int block_excl_scan(int v, int *lds, int &total) {
    // lds: WG ints
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int off = 1; off < WG; off <<= 1) {
        const int add = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    total = lds[WG - 1];
    const int incl = lds[tid];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(WG) void k_scan_sums(const int *__restrict__ in, int n, int *__restrict__ bsum) {
    __shared__ int lds[WG];
    const int base = blockIdx.x * SCAN_TILE;
    int a = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = base + k * WG + threadIdx.x;
        if (i < n) a += in[i];
    }
    int total;
    (void)block_excl_scan(a, lds, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single workgroup: exclusive scan of the block sums in place, grand total at bsum[nb]
__global__ __launch_bounds__(WG) void k_scan_top(int *__restrict__ bsum, int nb) {
    __shared__ int lds[WG];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += WG) {
        const int i = b0 + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int total;
        const int ex = block_excl_scan(v, lds, total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(WG) void k_scan_apply(const int *__restrict__ in, int n, const int *__restrict__ bsum,
                                                   int *__restrict__ out) {
    __shared__ int lds[WG];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;  // contiguous per thread
    int v[SCAN_ITEMS];
    int a = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        a += v[k];
    }
    int total;
    int run = block_excl_scan(a, lds, total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// Exclusive scan of up to SCAN_SINGLE_MAX ints by one workgroup, in 8 tiles of 1024 x 4:
// thread t holds items 4t..4t+3 of every tile.  All 8 tiles are loaded at once, coalesced
// and unconditionally (the input buffer holds at least SCAN_SINGLE_MAX ints; items at or past
// n are masked), the thread sums of every tile are scanned over the wave by shuffles, and
// one barrier later every thread adds the wave totals before it.  *total = the sum.
constexpr int SCAN_SINGLE_TILES = 8;
constexpr int SCAN_SINGLE_MAX = SCAN_SINGLE_TILES * 4096;

__device__ __forceinline__ int4 scan_item4(const int4 *src, int q, int i, int n) {
    int4 v = src[q * 1024];
    v.x = i < n ? v.x : 0;
    v.y = i + 1 < n ? v.y : 0;
    v.z = i + 2 < n ? v.z : 0;
    v.w = i + 3 < n ? v.w : 0;
    return v;
}

__global__ __launch_bounds__(1024) void k_scan_single(const int *__restrict__ in, int n, int *__restrict__ out,
                                                      int *__restrict__ total) {
    __shared__ int wsum[SCAN_SINGLE_TILES][16];  // [tile][wave]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int4 *src = reinterpret_cast<const int4 *>(in) + t;
    int4 v[SCAN_SINGLE_TILES];
#pragma unroll
    for (int q = 0; q < SCAN_SINGLE_TILES; ++q) v[q] = scan_item4(src, q, q * 4096 + 4 * t, n);
    int ex[SCAN_SINGLE_TILES];  // exclusive prefix of the thread's items within the tile, over the wave
#pragma unroll
    for (int q = 0; q < SCAN_SINGLE_TILES; ++q) {
        const int a = v[q].x + v[q].y + v[q].z + v[q].w;
        int x = a;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[q][w] = x;
        ex[q] = x - a;
    }
    __syncthreads();
    int carry = 0;
#pragma unroll
    for (int q = 0; q < SCAN_SINGLE_TILES; ++q) {
        int before = 0, tile = 0;
        for (int p = 0; p < 16; ++p) {
            const int sp = wsum[q][p];
            tile += sp;
            before += p < w ? sp : 0;
        }
        const int i = q * 4096 + 4 * t;
        const int r0 = carry + before + ex[q], r1 = r0 + v[q].x, r2 = r1 + v[q].y, r3 = r2 + v[q].z;
        if (i + 3 < n) {
            *reinterpret_cast<int4 *>(out + i) = make_int4(r0, r1, r2, r3);
        } else {
            if (i < n) out[i] = r0;
            if (i + 1 < n) out[i + 1] = r1;
            if (i + 2 < n) out[i + 2] = r2;
        }
        carry += tile;
    }
    if (t == 0) *total = carry;
}

// ---------------------------------------------------------------------------------------
// 3. LSD radix sort, 8-bit digits.  hist is digit-major: hist[d * nblk + blk].
__global__ __launch_bounds__(WG) void k_sort_hist(const uint64_t *__restrict__ keys, int n, int shift, int nblk,
                                                  int rounds, int *__restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * WG * rounds;
#pragma unroll 4
    for (int r = 0; r < rounds; ++r) {
        const int i = base + r * WG + threadIdx.x;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 255u)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Per-digit scan of the histogram: workgroup d turns row d (the blocks' counts of digit d) into
// its exclusive prefix in place and writes the row's total to dtot[d].  Thread t takes 8
// consecutive counts of each chunk of WG x 8.
__global__ __launch_bounds__(WG) void k_sort_rowscan(int *__restrict__ hist, int nblk, int *__restrict__ dtot) {
    __shared__ int wsum[WG / 64];
    int *row = hist + (size_t)blockIdx.x * nblk;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int c0 = 0; c0 < nblk; c0 += WG * 8) {
        const int i0 = c0 + 8 * t;
        int v[8], a = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = i0 + k < nblk ? row[i0 + k] : 0;
            a += v[k];
        }
        int x = a;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int run = carry + x - a, tot = 0;
#pragma unroll
        for (int p = 0; p < WG / 64; ++p) {
            const int sp = wsum[p];
            tot += sp;
            run += p < w ? sp : 0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (i0 + k < nblk) row[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (t == 0) dtot[blockIdx.x] = carry;
}

// Stable scatter of one tile of WG x R keys, staged through LDS so that the global writes of a
// digit's run are consecutive.  Wave w holds tile entries w*64R + 64r + lane (r < R), so the
// waves' entries, and each wave's rounds, follow the input order: a wave ranks its keys by
// ballots against its own running digit counts in LDS, with no barrier between rounds.  The
// tile's digit starts (the scan of the waves' counts) place every key in LDS in sorted order;
// thread t then writes entries t, t + WG, ... to the digit's global bucket, which starts at the
// keys of all smaller digits (the exclusive scan of dtot) plus the tile's row prefix.  GATHER
// (the last pass): also spos[g] = pos[id], the positions in sorted order.
template <int R, bool GATHER>
__global__ __launch_bounds__(WG) void k_sort_scatter(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                                     uint64_t *__restrict__ kout, int *__restrict__ vout, int n,
                                                     int shift, int nblk, const int *__restrict__ offs,
                                                     const int *__restrict__ dtot, const double2 *__restrict__ pos,
                                                     double2 *__restrict__ spos) {
    constexpr int NW = WG / 64, TILE = WG * R;
    __shared__ uint64_t sk[TILE];
    __shared__ int sv[TILE];
    __shared__ int wrun[NW][256];  // each wave's running digit counts, then its LDS base per digit
    __shared__ int lstart[256];    // start of digit d in the staged tile
    __shared__ int gstart[256];    // global start of the tile's run of digit d
    __shared__ int wtot[2][NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int base = blockIdx.x * TILE;
    const int cnt = min(TILE, n - base);
    const int cd = dtot[tid], row = offs[tid * nblk + blockIdx.x];
    const int wb = w * 64 * R + lane;
    uint64_t k[R];
    int v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = wb + 64 * r;
        k[r] = e < cnt ? kin[base + e] : 0ull;
        v[r] = e < cnt ? vin[base + e] : 0;
    }
#pragma unroll
    for (int q = 0; q < NW; ++q) wrun[q][tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int rk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool valid = wb + 64 * r < cnt;
        const int d = (int)((k[r] >> shift) & 255u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int below = __popcll(peers & lt);
        // every lane reads before the leaders write: one wave's LDS operations complete in order
        const int before = wrun[w][d];
        rk[r] = before + below;
        if (valid && below == 0) wrun[w][d] = before + __popcll(peers);
    }
    __syncthreads();
    {
        // thread tid = digit: the tile's count and its start in LDS; the digit's global start
        int c[NW], tot = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            c[q] = wrun[q][tid];
            tot += c[q];
        }
        int x = tot, y = cd;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int xs = __shfl_up(x, off, 64), ys = __shfl_up(y, off, 64);
            if (lane >= off) {
                x += xs;
                y += ys;
            }
        }
        if (lane == 63) {
            wtot[0][w] = x;
            wtot[1][w] = y;
        }
        __syncthreads();
        int ls = x - tot, gs = y - cd;
#pragma unroll
        for (int p = 0; p < NW; ++p) {
            ls += p < w ? wtot[0][p] : 0;
            gs += p < w ? wtot[1][p] : 0;
        }
        lstart[tid] = ls;
        gstart[tid] = gs + row;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            wrun[q][tid] = ls;
            ls += c[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (wb + 64 * r < cnt) {
            const int d = (int)((k[r] >> shift) & 255u);
            const int l = wrun[w][d] + rk[r];
            sk[l] = k[r];
            sv[l] = v[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = r * WG + tid;
        if (e < cnt) {
            const uint64_t key = sk[e];
            const int d = (int)((key >> shift) & 255u);
            const int g = gstart[d] + e - lstart[d];
            const int id = sv[e];
            kout[g] = key;
            vout[g] = id;
            if (GATHER) spos[g] = pos[id];
        }
    }
}

// The radix passes order the keys by their top 64 - SORT_SKIP bits only: keys equal in those
// bits belong to points within 2^-24 of the frame's side of each other on both axes, rare and
// in short runs.  The passes are stable, so a run is in vertex-id order; thread i, at the start
// of a run, puts it in full-key order by a stable insertion sort, moving the ids and positions
// with the keys: the result is the full 64-bit sort.  Other threads read the run's keys while
// it is being sorted only to compare their top bits, which every key of the run shares.
constexpr int SORT_SKIP = 16;
__global__ __launch_bounds__(WG) void k_sort_ties(uint64_t *__restrict__ keys, int *__restrict__ ids,
                                                  double2 *__restrict__ spos, int n) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i + 1 >= n) return;
    const uint64_t h = keys[i] >> SORT_SKIP;
    if ((keys[i + 1] >> SORT_SKIP) != h) return;
    if (i > 0 && (keys[i - 1] >> SORT_SKIP) == h) return;  // inside a run, not its first key
    int e = i + 2;
    while (e < n && (keys[e] >> SORT_SKIP) == h) ++e;
    for (int a = i + 1; a < e; ++a) {
        const uint64_t k = keys[a];
        const int id = ids[a];
        const double2 p = spos[a];
        int b = a - 1;
        for (; b >= i && keys[b] > k; --b) {
            keys[b + 1] = keys[b];
            ids[b + 1] = ids[b];
            spos[b + 1] = spos[b];
        }
        keys[b + 1] = k;
        ids[b + 1] = id;
        spos[b + 1] = p;
    }
}

// ---------------------------------------------------------------------------------------
// 5. tree
__device__ __forceinline__ int digit_at(uint64_t k, int d) { return (int)((k >> (62 - 2 * d)) & 3u); }
__device__ __forceinline__ uint64_t prefix_of(uint64_t k, int d) { return d == 0 ? 0ull : (k >> (64 - 2 * d)); }

// first index in [lo, hi) whose 2d-bit prefix equals pfx (keys sorted)
__device__ __forceinline__ int first_with_prefix(const uint64_t *keys, int lo, int hi, int d, uint64_t pfx) {
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (prefix_of(keys[m], d) < pfx) lo = m + 1; else hi = m;
    }
    return lo;
}
// last index in [lo, hi) whose prefix equals pfx, given keys[lo] has it
__device__ __forceinline__ int last_with_prefix(const uint64_t *keys, int lo, int hi, int d, uint64_t pfx) {
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (prefix_of(keys[m], d) <= pfx) lo = m; else hi = m;
    }
    return lo;
}
// first index in [0, i] whose 2d-bit prefix equals pfx, given keys[i] has it: probes leftwards
// at distances 1, 3, 7, ... from i (a node's points are contiguous and most nodes are small),
// then a binary search over the last interval
__device__ __forceinline__ int gallop_first(const uint64_t *keys, int i, int d, uint64_t pfx) {
    int hi = i, step = 1, lo = i - 1;
    while (lo >= 0 && prefix_of(keys[lo], d) == pfx) {
        hi = lo;
        step <<= 1;
        lo = hi - step;
    }
    return first_with_prefix(keys, lo < 0 ? 0 : lo + 1, hi, d, pfx);
}
// last index in [j, n) whose prefix equals pfx, given keys[j] has it (the mirror image)
__device__ __forceinline__ int gallop_last(const uint64_t *keys, int j, int n, int d, uint64_t pfx) {
    int lo = j, step = 1, hi = j + 1;
    while (hi < n && prefix_of(keys[hi], d) == pfx) {
        lo = hi;
        step <<= 1;
        hi = lo + step;
    }
    return last_with_prefix(keys, lo, hi < n ? hi : n, d, pfx);
}
// first index in [lo, hi) whose digit d is >= q (lower_bound of BarnesHut.h:143-146)
__device__ __forceinline__ int lower_digit(const uint64_t *keys, int lo, int hi, int d, int q) {
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (digit_at(keys[m], d) < q) lo = m + 1; else hi = m;
    }
    return lo;
}
// the node whose points are [s, e] (inclusive, e > s): the boundary where its first
// quadrant ends
__device__ __forceinline__ int rep_of_range(const uint64_t *keys, int s, int e) {
    const int d = __clzll((long long)(keys[s] ^ keys[e])) >> 1;
    const int q0 = digit_at(keys[s], d);
    return lower_digit(keys, s, e + 1, d, q0 + 1) - 1;
}

// boundary i (between Morton ranks i and i+1) is a node representative when it ends the
// first quadrant of the cell where k_i and k_i+1 split
__global__ __launch_bounds__(WG) void k_tree_bounds(const uint64_t *__restrict__ keys, int n, int *__restrict__ isrep,
                                                    int *__restrict__ status) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * WG + threadIdx.x;
    if (i >= n - 1) return;
    const uint64_t a = keys[i], b = keys[i + 1];
    if (a == b) {
        atomicOr(status, ST_DUPKEY);
        isrep[i] = 0;
        return;
    }
    const int d = __clzll((long long)(a ^ b)) >> 1;
    const uint64_t pfx = prefix_of(a, d);
    const int L = d == 0 ? 0 : gallop_first(keys, i, d, pfx);
    isrep[i] = digit_at(keys[L], d) == digit_at(a, d) ? 1 : 0;
}

// root = the node spanning all points; also the node count (scan total)
__global__ void k_tree_root(const uint64_t *__restrict__ keys, int n, const int *__restrict__ nodeid,
                            const int *__restrict__ scan_total, int *__restrict__ misc) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || misc[0]) return;
    misc[1] = nodeid[rep_of_range(keys, 0, n - 1)];
    misc[2] = *scan_total;
}

__global__ __launch_bounds__(WG) void k_tree_nodes(const uint64_t *__restrict__ keys, int n,
                                                   const int *__restrict__ isrep, const int *__restrict__ nodeid,
                                                   BHNode *__restrict__ nodes, int *__restrict__ npts,
                                                   int *__restrict__ node_lo, int *__restrict__ misc,
                                                   const int *__restrict__ status) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * WG + threadIdx.x;
    if (*status) return;
    if (i >= n - 1 || !isrep[i]) return;
    const int v = nodeid[i];
    const uint64_t a = keys[i];
    const int d = __clzll((long long)(a ^ keys[i + 1])) >> 1;
    const uint64_t pfx = prefix_of(a, d);
    const int L = d == 0 ? 0 : gallop_first(keys, i, d, pfx);
    const int R = d == 0 ? n - 1 : gallop_last(keys, i + 1, n, d, pfx);
    int bnd[5];
    bnd[0] = L;
    bnd[4] = R + 1;
    for (int q = 1; q < 4; ++q) bnd[q] = lower_digit(keys, L, R + 1, d, q);
    int ch[4];
    int nint = 0;
    for (int q = 0; q < 4; ++q) {
        const int s = bnd[q], e = bnd[q + 1];
        if (e - s == 0) {
            ch[q] = -1;
        } else if (e - s == 1) {
            ch[q] = -(s + 2);
        } else {
            const int c = nodeid[rep_of_range(keys, s, e - 1)];
            ch[q] = c;
            nodes[c].parent = v;
            ++nint;
        }
    }
    nodes[v].ch = make_int4(ch[0], ch[1], ch[2], ch[3]);
    nodes[v].depth = d;
    nodes[v].nint = nint;
    nodes[v].cnt = 0;
    npts[v] = R - L + 1;
    node_lo[v] = L;
    if (v == misc[1]) nodes[v].parent = -1;
}

// 6. centroids, bottom-up; the last child to finish computes its parent.  WT: stored with
// agent-scope stores, which the L2 writes through to memory
template <bool WT = false>
__device__ __forceinline__ void centroid_of(BHNode *nodes, const double2 *spos, int v) {
    const int4 ch = nodes[v].ch;
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    double sx = 0.0, sy = 0.0, count = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (c[q] >= 0) {
            const double2 cc = nodes[c[q]].c;
            sx += cc.x;
            sy += cc.y;
            count++;
        } else if (c[q] <= -2) {
            const double2 p = spos[-c[q] - 2];
            sx += p.x;
            sy += p.y;
            count++;
        }
    }
    if (count > 0) {
        sx /= count;
        sy /= count;
    }
    if (WT) {
        __hip_atomic_store(&nodes[v].c.x, sx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&nodes[v].c.y, sy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        nodes[v].c = make_double2(sx, sy);
    }
}

// Centroids bottom-up in two launches that keep the agent-scope handoffs to the top of the
// tree.  A "small" node holds at most SMALL_CELL points.  k_centroid_small: one thread per
// small subtree (a small node whose parent is not small) computes its nodes in post-order,
// with no synchronisation.  k_centroid_top: the large nodes, from those without large
// children upwards; the last large child to arrive at a parent computes it; small children
// were done by the previous launch.  Every centroid is formed by centroid_of from its
// children, so the order of the operations is the reference's.
constexpr int SMALL_CELL = 32;

// The first internal child of node u in quadrant order (-1: none).
__device__ __forceinline__ int first_internal(const BHNode *nodes, int u) {
    const int4 ch = nodes[u].ch;
    return ch.x >= 0 ? ch.x : ch.y >= 0 ? ch.y : ch.z >= 0 ? ch.z : ch.w;  // ch.w: itself, or < 0
}

// The internal sibling after child u of node p in quadrant order (-1: none).
__device__ __forceinline__ int next_internal(const BHNode *nodes, int p, int u) {
    const int4 ch = nodes[p].ch;
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    int nx = -1;
    bool past = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (past && nx < 0 && c[q] >= 0) nx = c[q];
        past = past || c[q] == u;
    }
    return nx;
}

// One thread per small subtree: its internal nodes in post-order (children in quadrant order)
// without a stack, by the parent links (no LDS, so the mostly idle workgroups of the node-id
// grid do not hold back occupancy).
__global__ __launch_bounds__(WG) void k_centroid_small(BHNode *__restrict__ nodes, const double2 *__restrict__ spos,
                                                       const int *__restrict__ npts, const int *__restrict__ misc) {
    const int r = blockIdx.x * WG + threadIdx.x;
    if (misc[0] || r >= misc[2]) return;
    const int par = nodes[r].parent;
    if (npts[r] > SMALL_CELL || (par >= 0 && npts[par] <= SMALL_CELL)) return;
    int u = r;
    for (int d; (d = first_internal(nodes, u)) >= 0;) u = d;
    for (;;) {
        centroid_of(nodes, spos, u);  // its internal children are done
        if (u == r) break;
        const int p = nodes[u].parent;
        const int nx = next_internal(nodes, p, u);
        if (nx < 0) {
            u = p;
            continue;
        }
        u = nx;
        for (int d; (d = first_internal(nodes, u)) >= 0;) u = d;
    }
}

__device__ __forceinline__ int large_children(const BHNode *nodes, const int *npts, int v) {
    const int4 ch = nodes[v].ch;
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    int k = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) k += (c[q] >= 0 && npts[c[q]] > SMALL_CELL) ? 1 : 0;
    return k;
}

// The large nodes climb: a node's centroid is stored with agent-scope stores (written through
// the XCD's L2 to memory), and once they have completed a relaxed counter at the parent counts
// the arrival; the last child to arrive acquires (invalidating its own L2's stale lines) and
// computes the parent from memory.  An acq_rel counter would write the whole L2 back at every
// step.
__global__ __launch_bounds__(WG) void k_centroid_top(BHNode *__restrict__ nodes, const double2 *__restrict__ spos,
                                                     const int *__restrict__ npts, const int *__restrict__ misc) {
    int v = blockIdx.x * WG + threadIdx.x;
    if (misc[0] || v >= misc[2]) return;
    if (npts[v] <= SMALL_CELL || large_children(nodes, npts, v) != 0) return;
    for (;;) {
        centroid_of<true>(nodes, spos, v);
        const int u = nodes[v].parent;
        if (u < 0) break;
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const int old = __hip_atomic_fetch_add(&nodes[u].cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old != large_children(nodes, npts, u) - 1) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        v = u;
    }
}

// ---------------------------------------------------------------------------------------
// 7. repulsion by traversal
struct Frame {
    double maxmin;
    double theta;
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Unit vector d/|d| added to (fx, fy).  REFINE (the default precision): FP64 difference, the
// FP32 reciprocal square root refined by one FP64 Newton step (relative error ~1e-14) and an
// FP64 sum.  Otherwise (BL_PREC_FP32) all in FP32.  Over hundreds of iterations the FP32
// terms bias the layout: the 600-iteration energy on the 64x64 grid ends several percent
// below the reference's, while the refined terms agree with it.
template <bool REFINE>
struct UnitSum {
    typename std::conditional<REFINE, double, float>::type x = 0, y = 0;
    __device__ __forceinline__ void add(double dx, double dy) {
        if constexpr (REFINE) {
            const double r2 = fma(dy, dy, dx * dx);
            double q = (double)__builtin_amdgcn_rsqf((float)r2);
            q = q * fma(-0.5 * r2 * q, q, 1.5);
            x = fma(dx, q, x);
            y = fma(dy, q, y);
        } else {
            const float fx = (float)dx, fy = (float)dy;
            const float inv = __builtin_amdgcn_rsqf(fmaf(fy, fy, fx * fx));
            x = fmaf(fx, inv, x);
            y = fmaf(fy, inv, y);
        }
    }
};

// FP32 / mixed modes.  One wave per 64 Morton-consecutive vertices; shared stack of (node, mask).
template <bool REFINE>
__global__ __launch_bounds__(WG) void k_rep32(const double2 *__restrict__ qpos, const int *__restrict__ qid, int n,
                                              const double2 *__restrict__ spos,
                                              const BHNode *__restrict__ nodes, const int *__restrict__ misc,
                                              const double *__restrict__ frame, double theta,
                                              double2 *__restrict__ rep, const int *__restrict__ status) {
    __shared__ int stk_node[WG / 64][STACK];
    __shared__ uint64_t stk_mask[WG / 64][STACK];
    if (*status) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (xcd_tile(blockIdx.x, gridDim.x) * (WG / 64) + w) * 64 + lane;
    const bool valid = k < n;
    const double2 p = qpos[valid ? k : n - 1];
    const double maxmin = frame[2];
    const float th2 = (float)(theta * theta);
    const float lo_band = 1.0f - 1.0f / 4096.0f, hi_band = 1.0f + 1.0f / 4096.0f;
    UnitSum<REFINE> f;
    const int root = misc[1];
    int sp = 0;
    uint64_t m0 = __ballot(valid);
    if (m0 == 0) return;
    if (root < 0) {  // single vertex: no other point, no force
        if (valid) rep[qid[k]] = make_double2(0.0, 0.0);
        return;
    }
    if (lane == 0) {
        stk_node[w][0] = root;
        stk_mask[w][0] = m0;
    }
    sp = 1;
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        --sp;
        const int v = uni(stk_node[w][sp]);
        const uint64_t m = stk_mask[w][sp];
        __builtin_amdgcn_wave_barrier();
        const bool act = (m >> lane) & 1ull;
        const BHNode nd = nodes[v];
        const double dx64 = nd.c.x - p.x, dy64 = nd.c.y - p.y;
        const float dx = (float)dx64, dy = (float)dy64;
        const float r2 = fmaf(dy, dy, dx * dx);
        const double D = ldexp(maxmin, -nd.depth);
        const float D32 = (float)D;
        const float a = D32 * D32, b = th2 * r2;
        bool acc;
        if (a < b * lo_band) acc = true;
        else if (a > b * hi_band) acc = false;
        else acc = D / sqrt(dx64 * dx64 + dy64 * dy64) < theta;  // exact FP64 test, :71-72
        if (act && acc) f.add(dx64, dy64);
        const uint64_t rej = __ballot(act && !acc);
        if (rej) {
            const int c[4] = {nd.ch.x, nd.ch.y, nd.ch.z, nd.ch.w};
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                if (c[q] >= 0) {
                    if (lane == 0) {
                        stk_node[w][sp] = c[q];
                        stk_mask[w][sp] = rej;
                    }
                    ++sp;
                } else if (c[q] <= -2) {
                    if ((rej >> lane) & 1ull) {  // leaf, :78-83
                        const double2 l = spos[-c[q] - 2];
                        const double lx = l.x - p.x, ly = l.y - p.y;
                        if (lx * lx + ly * ly > 0.0) f.add(lx, ly);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (valid) rep[qid[k]] = make_double2((double)f.x, (double)f.y);
}

// FP64 mode: the reference's depth-first order (quadrants 0..3, leaves in place) and its
// FP64 operations; leaves are stack entries too so that every lane sums in its own order.
__global__ __launch_bounds__(WG) void k_rep64(const double2 *__restrict__ qpos, const int *__restrict__ qid, int n,
                                              const double2 *__restrict__ spos,
                                              const BHNode *__restrict__ nodes, const int *__restrict__ misc,
                                              const double *__restrict__ frame, double theta,
                                              double2 *__restrict__ rep, const int *__restrict__ status) {
    __shared__ int stk_node[WG / 64][STACK];
    __shared__ uint64_t stk_mask[WG / 64][STACK];
    if (*status) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (blockIdx.x * (WG / 64) + w) * 64 + lane;
    const bool valid = k < n;
    const double2 p = qpos[valid ? k : n - 1];
    const double maxmin = frame[2];
    double fx = 0.0, fy = 0.0;
    const int root = misc[1];
    uint64_t m0 = __ballot(valid);
    if (m0 == 0) return;
    if (root < 0) {
        if (valid) rep[qid[k]] = make_double2(0.0, 0.0);
        return;
    }
    if (lane == 0) {
        stk_node[w][0] = root;
        stk_mask[w][0] = m0;
    }
    int sp = 1;
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        --sp;
        const int v = uni(stk_node[w][sp]);
        const uint64_t m = stk_mask[w][sp];
        __builtin_amdgcn_wave_barrier();
        const bool act = (m >> lane) & 1ull;
        if (v <= -2) {  // leaf entry
            if (act) {
                const double2 l = spos[-v - 2];
                const double lx = l.x - p.x, ly = l.y - p.y;
                const double den = sqrt(lx * lx + ly * ly);
                if (den > 0) {
                    const double inv = 1.0 / den;
                    fx = fx + lx * inv;
                    fy = fy + ly * inv;
                }
            }
            continue;
        }
        const BHNode nd = nodes[v];
        const double dx = nd.c.x - p.x, dy = nd.c.y - p.y;
        const double dist = sqrt(dx * dx + dy * dy);
        const double D = ldexp(maxmin, -nd.depth);
        const bool acc = D / dist < theta;
        if (act && acc) {
            const double inv = 1.0 / dist;
            fx = fx + dx * inv;
            fy = fy + dy * inv;
        }
        const uint64_t rej = __ballot(act && !acc);
        if (rej) {
            const int c[4] = {nd.ch.x, nd.ch.y, nd.ch.z, nd.ch.w};
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                if (c[q] != -1) {
                    if (lane == 0) {
                        stk_node[w][sp] = c[q];
                        stk_mask[w][sp] = rej;
                    }
                    ++sp;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (valid) rep[qid[k]] = make_double2(fx, fy);
}

// ---------------------------------------------------------------------------------------
// 8. batch steps: attraction with live positions minus the precomputed repulsion, move.
// One workgroup walks all batches of the iteration in order (Gauss-Seidel across batches,
// Jacobi inside a batch: every force of the batch is formed before any move).
__global__ __launch_bounds__(1024) void k_attract_move(const int *__restrict__ rowptr, const int *__restrict__ colids,
                                                       const float *__restrict__ w_edge, double W,
                                                       double2 *__restrict__ pos, const double2 *__restrict__ rep,
                                                       double2 *__restrict__ fbuf, int n, int B, double step,
                                                       double *__restrict__ eterm, const int *__restrict__ status) {
    if (*status) return;
    const int nb = (n + B - 1) / B;
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B;
        const int b1 = min(n, b0 + B);
        for (int i = b0 + (int)threadIdx.x; i < b1; i += 1024) {
            const double2 pi = pos[i];
            double fx = 0.0, fy = 0.0;
            const int e = rowptr[i + 1];
            for (int k = rowptr[i]; k < e; ++k) {  // every CSR entry, :1055-1058
                const double2 pj = pos[colids[k]];
                const double dx = pj.x - pi.x, dy = pj.y - pi.y;
                const double Wk = w_edge ? W * (double)w_edge[k] : W;
                const double s = Wk * sqrt(dx * dx + dy * dy);
                fx = fx + dx * s;
                fy = fy + dy * s;
            }
            const double2 r = rep[i];
            fbuf[i] = make_double2(fx - r.x, fy - r.y);  // f - calcRepForce(x_i), :1059
        }
        __syncthreads();
        for (int i = b0 + (int)threadIdx.x; i < b1; i += 1024) {
            const double2 f = fbuf[i];
            const double mag = sqrt(f.x * f.x + f.y * f.y);
            double2 q = pos[i];
            q.x = q.x + (f.x / mag) * step;
            q.y = q.y + (f.y / mag) * step;
            pos[i] = q;
            eterm[i] = f.x * f.x + f.y * f.y;
        }
        __syncthreads();
    }
}

// Same batch walk for the FP32/mixed modes, with the attraction work of each batch split
// evenly over the 1024 threads by CSR entry (a hub row of tens of thousands of entries is
// shared by many threads instead of one).  Thread t takes the entries [E0 + t*c, E0 + t*c + c)
// of the batch; a row lying inside one thread's range is summed by that thread; a row spread
// over several threads is the sum, in thread order, of those threads' partial sums (the
// range's last row of the first thread, then whole ranges, then the first row of the last
// thread), so the result does not depend on timing.
//
// Latency: the walk is one dependent step per batch, so every global load on a step's path
// costs a full memory round trip.  With STAGE (batches of at most WALK_MAXB vertices) the
// batch's rowptr slice, own positions and repulsion sit in LDS, loaded one batch ahead (they
// do not depend on the moves of the current batch: a batch moves only its own vertices);
// row lookups are binary searches in LDS; the entries are fetched WALK_G at a time so their
// colids and neighbour-position loads overlap; the force of each batch vertex is combined and
// applied by the same thread, without a round trip through global memory.
constexpr int WALK_T = 1024;
constexpr int WALK_G = 8;
constexpr int WALK_MAXB = 2048;

template <bool STAGE>
__global__ __launch_bounds__(WALK_T) void k_walk(const int *__restrict__ rowptr, const int *__restrict__ colids,
                                                  const float *__restrict__ w_edge, double W,
                                                  double2 *__restrict__ pos, const double2 *__restrict__ rep,
                                                  double2 *__restrict__ fbuf, int n, int B, double step,
                                                  double *__restrict__ eterm, const int *__restrict__ status) {
    constexpr int T = WALK_T;
    constexpr int SB = STAGE ? WALK_MAXB : 1;
    constexpr int PER = (WALK_MAXB + T - 1) / T;  // staged vertices per thread
    __shared__ double2 head[T], tail[T];
    __shared__ int hrow[T], trow[T];
    __shared__ int rps[SB + 1];
    __shared__ double2 bpos[SB], rsum[SB];
    if (*status) return;
    const int t = threadIdx.x;
    const int nb = (n + B - 1) / B;
    // staged data of the next batch, held in registers until the current batch is done
    int nrp[PER + 1];
    double2 npos[PER], nrep[PER];
    auto fetch = [&](int b) {
        const int b0 = b * B, bc = min(n, b0 + B) - b0;
#pragma unroll
        for (int j = 0; j <= PER; ++j) {
            const int k = t + j * T;
            nrp[j] = k <= bc ? rowptr[b0 + k] : 0;
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = t + j * T;
            npos[j] = k < bc ? pos[b0 + k] : make_double2(0.0, 0.0);
            nrep[j] = k < bc ? rep[b0 + k] : make_double2(0.0, 0.0);
        }
    };
    if (STAGE) fetch(0);
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, b1 = min(n, b0 + B), bc = b1 - b0;
        if (STAGE) {
#pragma unroll
            for (int j = 0; j <= PER; ++j) {
                const int k = t + j * T;
                if (k <= bc) rps[k] = nrp[j];
            }
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int k = t + j * T;
                if (k < bc) bpos[k] = npos[j];
            }
            __syncthreads();
        }
        auto RP = [&](int i) { return STAGE ? rps[i - b0] : rowptr[i]; };  // i in [b0, b1]
        auto OWN = [&](int i) { return STAGE ? bpos[i - b0] : pos[i]; };   // i in [b0, b1)
        const int E0 = RP(b0), E1 = RP(b1);
        const int total = E1 - E0;
        const int c = max(1, (total + T - 1) / T);
        const int s = E0 + t * c, e = min(E1, s + c);
        hrow[t] = -1;
        trow[t] = -1;
        if (s < e) {
            int lo = b0, hi = b1;  // row containing entry s: last r with rowptr[r] <= s
            while (hi - lo > 1) {
                const int m = (lo + hi) >> 1;
                if (RP(m) <= s) lo = m; else hi = m;
            }
            int r = lo;
            const int first = r;
            int rend = RP(r + 1);
            double2 pi = OWN(r);
            double ax = 0.0, ay = 0.0;
            for (int k0 = s; k0 < e; k0 += WALK_G) {
                int cid[WALK_G];
                double2 pj[WALK_G];
                double wk[WALK_G];
#pragma unroll
                for (int g = 0; g < WALK_G; ++g) cid[g] = k0 + g < e ? colids[k0 + g] : -1;
#pragma unroll
                for (int g = 0; g < WALK_G; ++g) {
                    pj[g] = cid[g] >= 0 ? pos[cid[g]] : make_double2(0.0, 0.0);
                    wk[g] = (w_edge && k0 + g < e) ? W * (double)w_edge[k0 + g] : W;
                }
#pragma unroll
                for (int g = 0; g < WALK_G; ++g) {
                    const int k = k0 + g;
                    if (k >= e) break;
                    while (k >= rend) {  // row r complete inside this range
                        if (r == first) {
                            hrow[t] = r;
                            head[t] = make_double2(ax, ay);
                        } else if (STAGE) {
                            rsum[r - b0] = make_double2(ax, ay);
                        } else {
                            fbuf[r] = make_double2(ax, ay);
                        }
                        ++r;
                        rend = RP(r + 1);
                        pi = OWN(r);
                        ax = ay = 0.0;
                    }
                    const double dx = pj[g].x - pi.x, dy = pj[g].y - pi.y;
                    const double sc = wk[g] * sqrt(dx * dx + dy * dy);
                    ax = fma(dx, sc, ax);
                    ay = fma(dy, sc, ay);
                }
            }
            if (r == first) {
                hrow[t] = r;
                head[t] = make_double2(ax, ay);
            } else {
                trow[t] = r;
                tail[t] = make_double2(ax, ay);
            }
        }
        __syncthreads();
        // the next batch's staged data: issued now so that it arrives while this batch is
        // combined and moved
        double2 cur_rep[PER];
        if (STAGE) {
#pragma unroll
            for (int j = 0; j < PER; ++j) cur_rep[j] = nrep[j];
            if (b + 1 < nb) fetch(b + 1);
        }
        for (int i = b0 + t, j = 0; i < b1; i += T, ++j) {
            const int rs = RP(i), re = RP(i + 1);
            double2 f = make_double2(0.0, 0.0);
            if (re > rs) {
                const int ta = (rs - E0) / c, tb = (re - 1 - E0) / c;
                if (ta == tb) {
                    f = hrow[ta] == i ? head[ta] : (trow[ta] == i ? tail[ta] : (STAGE ? rsum[i - b0] : fbuf[i]));
                } else {
                    f = trow[ta] == i ? tail[ta] : head[ta];
                    for (int q = ta + 1; q <= tb; ++q) {
                        f.x += head[q].x;
                        f.y += head[q].y;
                    }
                }
            }
            double2 rr;
            if (STAGE) {
                rr = cur_rep[0];
#pragma unroll
                for (int u = 1; u < PER; ++u)
                    if (j == u) rr = cur_rep[u];
            } else {
                rr = rep[i];
            }
            f.x -= rr.x;  // f - calcRepForce(x_i), :1059
            f.y -= rr.y;
            const double mag = sqrt(f.x * f.x + f.y * f.y);
            double2 q = OWN(i);
            q.x = q.x + (f.x / mag) * step;
            q.y = q.y + (f.y / mag) * step;
            pos[i] = q;
            eterm[i] = f.x * f.x + f.y * f.y;
        }
        __syncthreads();
    }
}

// The one-workgroup walk with the batch's first entries prefetched (batches of at most
// WALK_MAXB vertices).  Besides the staged rowptr slice, own positions and repulsion, each
// thread loads during batch b the column ids, weights and neighbour positions of its first
// WALK_PF entries of batch b+1.  A neighbour in batch b is moved after that load: at the
// start of batch b+1 such entries take the moved position from the LDS copy the move phase
// keeps (newpos), every other neighbour's position is final for the iteration.  A batch
// whose entries fit in WALK_PF per thread thus reads no memory on its critical path.
// Same sums, in the same order, as k_walk.
constexpr int WALK_PF = 4;

__global__ __launch_bounds__(WALK_T) void k_walk_pf(const int *__restrict__ rowptr, const int *__restrict__ colids,
                                                     const float *__restrict__ w_edge, double W,
                                                     double2 *__restrict__ pos, const double2 *__restrict__ rep,
                                                     int n, int B, double step, double *__restrict__ eterm,
                                                     const int *__restrict__ status) {
    constexpr int T = WALK_T;
    constexpr int PER = (WALK_MAXB + T - 1) / T;  // staged vertices per thread
    __shared__ double2 head[T], tail[T];
    __shared__ int hrow[T], trow[T];
    __shared__ int rps[WALK_MAXB + 1];
    __shared__ double2 bpos[WALK_MAXB], rsum[WALK_MAXB], newpos[WALK_MAXB];
    if (__builtin_amdgcn_readfirstlane(*status)) return;
    const int t = threadIdx.x;
    const int nb = (n + B - 1) / B;
    int nrp[PER + 1];
    double2 npos[PER], nrep[PER];
    int pcid[WALK_PF];
    double2 ppos[WALK_PF];
    double pw[WALK_PF];
    auto fetch = [&](int b) {  // staged data of batch b
        const int b0 = b * B, bc = min(n, b0 + B) - b0;
#pragma unroll
        for (int j = 0; j <= PER; ++j) {
            const int k = t + j * T;
            nrp[j] = k <= bc ? rowptr[b0 + k] : 0;
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = t + j * T;
            npos[j] = k < bc ? pos[b0 + k] : make_double2(0.0, 0.0);
            nrep[j] = k < bc ? rep[b0 + k] : make_double2(0.0, 0.0);
        }
    };
    auto prefetch = [&](int b) {  // first entries of this thread's share of batch b
        const int b0 = b * B, b1 = min(n, b0 + B);
        const int E0 = __builtin_amdgcn_readfirstlane(rowptr[b0]);
        const int E1 = __builtin_amdgcn_readfirstlane(rowptr[b1]);
        const int c = max(1, (E1 - E0 + T - 1) / T);
        const int s = E0 + t * c, e = min(E1, s + c);
#pragma unroll
        for (int g = 0; g < WALK_PF; ++g) pcid[g] = s + g < e ? colids[s + g] : -1;
#pragma unroll
        for (int g = 0; g < WALK_PF; ++g) {
            ppos[g] = pcid[g] >= 0 ? pos[pcid[g]] : make_double2(0.0, 0.0);
            pw[g] = (w_edge && s + g < e) ? W * (double)w_edge[s + g] : W;
        }
    };
    fetch(0);
    prefetch(0);
    int pb0 = 0, pb1 = 0;  // the previous batch, whose moved positions are in newpos
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, b1 = min(n, b0 + B), bc = b1 - b0;
#pragma unroll
        for (int j = 0; j <= PER; ++j) {
            const int k = t + j * T;
            if (k <= bc) rps[k] = nrp[j];
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = t + j * T;
            if (k < bc) bpos[k] = npos[j];
        }
        // prefetched neighbours that the previous batch moved
#pragma unroll
        for (int g = 0; g < WALK_PF; ++g)
            if (pcid[g] >= pb0 && pcid[g] < pb1) ppos[g] = newpos[pcid[g] - pb0];
        __syncthreads();
        const int E0 = rps[0], E1 = rps[bc];
        const int total = E1 - E0;
        const int c = max(1, (total + T - 1) / T);
        const int s = E0 + t * c, e = min(E1, s + c);
        hrow[t] = -1;
        trow[t] = -1;
        if (s < e) {
            int lo = b0, hi = b1;  // row containing entry s: last r with rowptr[r] <= s
            while (hi - lo > 1) {
                const int m = (lo + hi) >> 1;
                if (rps[m - b0] <= s) lo = m; else hi = m;
            }
            int r = lo;
            const int first = r;
            int rend = rps[r + 1 - b0];
            double2 pi = bpos[r - b0];
            double ax = 0.0, ay = 0.0;
            for (int k0 = s; k0 < e; k0 += WALK_PF) {
                int cid[WALK_PF];
                double2 pj[WALK_PF];
                double wk[WALK_PF];
                if (k0 == s) {
#pragma unroll
                    for (int g = 0; g < WALK_PF; ++g) {
                        cid[g] = pcid[g];
                        pj[g] = ppos[g];
                        wk[g] = pw[g];
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < WALK_PF; ++g) cid[g] = k0 + g < e ? colids[k0 + g] : -1;
#pragma unroll
                    for (int g = 0; g < WALK_PF; ++g) {
                        pj[g] = cid[g] >= 0 ? pos[cid[g]] : make_double2(0.0, 0.0);
                        wk[g] = (w_edge && k0 + g < e) ? W * (double)w_edge[k0 + g] : W;
                    }
                }
#pragma unroll
                for (int g = 0; g < WALK_PF; ++g) {
                    const int k = k0 + g;
                    if (k >= e) break;
                    while (k >= rend) {  // row r complete inside this range
                        if (r == first) {
                            hrow[t] = r;
                            head[t] = make_double2(ax, ay);
                        } else {
                            rsum[r - b0] = make_double2(ax, ay);
                        }
                        ++r;
                        rend = rps[r + 1 - b0];
                        pi = bpos[r - b0];
                        ax = ay = 0.0;
                    }
                    const double dx = pj[g].x - pi.x, dy = pj[g].y - pi.y;
                    const double sc = wk[g] * sqrt(dx * dx + dy * dy);
                    ax = fma(dx, sc, ax);
                    ay = fma(dy, sc, ay);
                }
            }
            if (r == first) {
                hrow[t] = r;
                head[t] = make_double2(ax, ay);
            } else {
                trow[t] = r;
                tail[t] = make_double2(ax, ay);
            }
        }
        __syncthreads();
        // the next batch's staged data and first entries, issued before the moves
        double2 cur_rep[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) cur_rep[j] = nrep[j];
        if (b + 1 < nb) {
            fetch(b + 1);
            prefetch(b + 1);
        }
        for (int i = b0 + t, j = 0; i < b1; i += T, ++j) {
            const int rs = rps[i - b0], re = rps[i + 1 - b0];
            double2 f = make_double2(0.0, 0.0);
            if (re > rs) {
                const int ta = (rs - E0) / c, tb = (re - 1 - E0) / c;
                if (ta == tb) {
                    f = hrow[ta] == i ? head[ta] : (trow[ta] == i ? tail[ta] : rsum[i - b0]);
                } else {
                    f = trow[ta] == i ? tail[ta] : head[ta];
                    for (int q = ta + 1; q <= tb; ++q) {
                        f.x += head[q].x;
                        f.y += head[q].y;
                    }
                }
            }
            double2 rr = cur_rep[0];
#pragma unroll
            for (int u = 1; u < PER; ++u)
                if (j == u) rr = cur_rep[u];
            f.x -= rr.x;  // f - calcRepForce(x_i), :1059
            f.y -= rr.y;
            const double mag = sqrt(f.x * f.x + f.y * f.y);
            double2 q = bpos[i - b0];
            q.x = q.x + (f.x / mag) * step;
            q.y = q.y + (f.y / mag) * step;
            pos[i] = q;
            newpos[i - b0] = q;
            eterm[i] = f.x * f.x + f.y * f.y;
        }
        __syncthreads();
        pb0 = b0;
        pb1 = b1;
    }
}

// The one-workgroup walk with one thread per batch vertex, for batches of at most WROW_T
// vertices and WROW_CAP CSR entries whose rows are all short (at most WROW_MAXDEG entries).
// The workgroup is split into WROW_T computing threads and WROW_T loading threads, so that
// the gathers of one batch overlap the arithmetic of the previous one (a wave that issues a
// batch's worth of random gathers stalls for the whole issue, so the two must not share a
// wave):
//   computing thread t owns vertex b0 + t of every batch, sums its row itself in CSR order
//   (the reference's order) from the LDS stage and moves it, into LDS rings;
//   the loading threads, during batch b, write the moves of batch b-1 to memory and fill the
//   other half of the stage with batch b+1: its CSR entries (column id, weight, neighbour
//   position, spread over the threads by entry) and its vertices' rows, positions and
//   repulsion.  Column ids run a batch ahead of the positions they address, CSR ranges two.
// One barrier ends each batch.
//
// Memory holds the iteration-start position of every vertex of the current batch (Jacobi
// within a batch): moves reach memory only in the next batch.  A gather issued during batch
// b for batch b+1 may see the old position of a vertex of batch b-1 (stored meanwhile) or b
// (moved meanwhile), so every neighbour in batch b-2 or b-1 of the batch being computed
// takes its position from the ring of moved positions.  Older batches were stored, and the
// stores completed, before a barrier that every gather has passed since.
constexpr int WROW_T = 256;
constexpr int WROW_CAP = 2560;   // staged CSR entries per batch
constexpr int WROW_S = WROW_CAP / WROW_T;
constexpr int WROW_MAXDEG = 64;  // longest row for which this walk is chosen

constexpr size_t wrow_lds(bool has_w) {
    return sizeof(double2) * (3 * WROW_T + 2 * WROW_CAP + 4 * WROW_T) + sizeof(double) * 2 * WROW_T +
           sizeof(int) * (2 * WROW_CAP + 4 * WROW_T) + (has_w ? sizeof(float) * 2 * WROW_CAP : 0);
}

// Whether k_walk_row can take batches of B vertices of graph g, and if so the early/late
// split of the CSR on the device (both cached per batch size).  Early entries of a batch
// must fit the walk's stage.
bool row_walk_fits(bl_graph *g, int B) {
    if (g->row_walk_batch == B) return g->row_walk_ok;
    g->row_walk_batch = B;
    g->row_walk_ok = false;
    const int n = g->n;
    if (!(B <= WROW_T && g->nnz > 0 && g->max_row <= WROW_MAXDEG && (int)g->h_rowptr.size() == n + 1 &&
          (int64_t)g->h_colids.size() == g->nnz))
        return false;
    const bool hw = !g->h_w.empty();
    std::vector<int32_t> erp(n + 1, 0), lrp(n + 1, 0), eci, lci;
    std::vector<float> ew, lw;
    eci.reserve(g->nnz / 2 + 1);
    lci.reserve(g->nnz / 2 + 1);
    for (int i = 0; i < n; ++i) {
        const int b0 = (i / B) * B;
        for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
            const bool early = g->h_colids[e] < b0;
            (early ? eci : lci).push_back(g->h_colids[e]);
            if (hw) (early ? ew : lw).push_back(g->h_w[e]);
        }
        erp[i + 1] = (int32_t)eci.size();
        lrp[i + 1] = (int32_t)lci.size();
    }
    for (int b0 = 0; b0 < n; b0 += B)
        if (erp[std::min(n, b0 + B)] - erp[b0] > WROW_CAP) return false;
    // the walk reads index 0 of an empty list (clamped loads): keep one valid entry
    if (eci.empty()) eci.push_back(0), ew.push_back(0.0f);
    if (lci.empty()) lci.push_back(0), lw.push_back(0.0f);
    if (g->early_rp.ensure(sizeof(int32_t) * (n + 1)) || g->late_rp.ensure(sizeof(int32_t) * (n + 1)) ||
        g->early_ci.ensure(sizeof(int32_t) * eci.size()) || g->late_ci.ensure(sizeof(int32_t) * lci.size()))
        return false;
    if (hw && (g->early_w.ensure(sizeof(float) * ew.size()) || g->late_w.ensure(sizeof(float) * lw.size())))
        return false;
    if (hipMemcpy(g->early_rp.p, erp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->late_rp.p, lrp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->early_ci.p, eci.data(), sizeof(int32_t) * eci.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->late_ci.p, lci.data(), sizeof(int32_t) * lci.size(), hipMemcpyHostToDevice) != hipSuccess)
        return false;
    if (hw && (hipMemcpy(g->early_w.p, ew.data(), sizeof(float) * ew.size(), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(g->late_w.p, lw.data(), sizeof(float) * lw.size(), hipMemcpyHostToDevice) != hipSuccess))
        return false;
    g->row_walk_ok = true;
    return true;
}

// Attraction of every vertex's late neighbours (same or later batch: iteration-start
// positions), in CSR order, folded into the repulsion the walk subtracts:
// rep_i <- rep_i - sum_late d * (W w |d|), so the walk's f = sum_early - rep_i is the whole
// force (algorithms.cpp:1055-1059).  One thread per vertex (rows are short here).
__global__ void k_late_attract(const int *__restrict__ lrp, const int *__restrict__ lci,
                               const float *__restrict__ lw, double W, const double2 *__restrict__ pos,
                               double2 *__restrict__ rep, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 pi = pos[i];
    double ax = 0.0, ay = 0.0;
    for (int e = lrp[i]; e < lrp[i + 1]; ++e) {
        const double2 pj = pos[lci[e]];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y;
        const double sc = (lw ? W * (double)lw[e] : W) * sqrt(dx * dx + dy * dy);
        ax = fma(dx, sc, ax);
        ay = fma(dy, sc, ay);
    }
    const double2 r = rep[i];
    rep[i] = make_double2(r.x - ax, r.y - ay);
}

template <bool HAS_W>
__global__ __launch_bounds__(2 * WROW_T) void k_walk_row(const int *__restrict__ rowptr, const int *__restrict__ colids,
                                                         const float *__restrict__ w_edge, double W,
                                                         double2 *__restrict__ pos, const double2 *__restrict__ rep,
                                                         int n, int B, double step, double *__restrict__ eterm,
                                                         const int *__restrict__ status) {
    constexpr int T = WROW_T, S = WROW_S, C = WROW_CAP;
    extern __shared__ double2 lds[];
    double2 *ring = lds;                    // [3][T] moved positions of batch c (slot c % 3)
    double2 *spos = ring + 3 * T;           // [2][C] staged neighbour positions (half c & 1)
    double2 *sv = spos + 2 * C;             // [2][T] staged own positions
    double2 *sr = sv + 2 * T;               // [2][T] staged repulsion
    double *ering = reinterpret_cast<double *>(sr + 2 * T);  // [2][T] energy terms of batch c
    int *scid = reinterpret_cast<int *>(ering + 2 * T);      // [2][C] staged column ids
    int *srs = scid + 2 * C;                                 // [2][T] staged row starts
    int *sre = srs + 2 * T;                                  // [2][T] staged row ends
    float *sw = reinterpret_cast<float *>(sre + 2 * T);      // [2][C] staged weights (HAS_W)
    if (__builtin_amdgcn_readfirstlane(*status)) return;
    const int nb = (n + B - 1) / B;
    if (threadIdx.x < T) {  // ---- computing threads ----
        const int t = threadIdx.x;
        __syncthreads();  // batch 0 staged
        for (int b = 0; b < nb; ++b) {
            const int b0 = b * B, b1 = min(n, b0 + B);
            const int m1 = b0 - B, m2 = b0 - 2 * B;  // first vertices of batches b-1 and b-2
            const double2 *ring1 = ring + ((b + 2) % 3) * T, *ring2 = ring + ((b + 1) % 3) * T;
            const int cur = b & 1;
            const int i = b0 + t;
            if (i < b1) {
                const int E0 = srs[cur * T];
                const int rs = srs[cur * T + t], re = sre[cur * T + t];
                const double2 v = sv[cur * T + t], r = sr[cur * T + t];
                double ax = 0.0, ay = 0.0;
                for (int e = rs; e < re; e += 4) {
                    int j[4];
                    double2 pj[4];
                    double wk[4];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int k = cur * C + min(e + g, re - 1) - E0;
                        j[g] = scid[k];
                        pj[g] = spos[k];
                        wk[g] = HAS_W ? W * (double)sw[k] : W;
                    }
                    double dx[4], dy[4], sc[4];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        if (j[g] >= m2 && j[g] < b0) pj[g] = j[g] >= m1 ? ring1[j[g] - m1] : ring2[j[g] - m2];
                        dx[g] = pj[g].x - v.x;
                        dy[g] = pj[g].y - v.y;
                        sc[g] = wk[g] * sqrt(dx[g] * dx[g] + dy[g] * dy[g]);
                    }
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        if (e + g >= re) break;
                        ax = fma(dx[g], sc[g], ax);
                        ay = fma(dy[g], sc[g], ay);
                    }
                }
                const double2 f = make_double2(ax - r.x, ay - r.y);  // f - calcRepForce(x_i), :1059
                const double mag = sqrt(f.x * f.x + f.y * f.y);
                double2 q = v;
                q.x = q.x + (f.x / mag) * step;
                q.y = q.y + (f.y / mag) * step;
                ring[(b % 3) * T + t] = q;
                ering[cur * T + t] = f.x * f.x + f.y * f.y;
            }
            __syncthreads();  // end of batch b
        }
        return;
    }
    // ---- loading threads ----
    const int t = threadIdx.x - T;
    // CSR range of batch b (past the last batch: the empty range at nnz), loaded as vector loads
    // into values equal in all lanes: a uniform load would become a scalar load, whose wait
    // counter LDS accesses share.
    auto bounds = [&](int b, int &a, int &e) {
        const int ia = min(b * B, n), ie = min(n, b * B + B);
        int va, ve;
        __asm__("v_mov_b32 %0, %1" : "=v"(va) : "s"(ia));
        __asm__("v_mov_b32 %0, %1" : "=v"(ve) : "s"(ie));
        a = rowptr[va];
        e = rowptr[ve];
    };
    // Every load below is unconditional, with its index clamped into the array: slots past a
    // batch's entries (or vertices past its end) hold some valid entry and are never read.
    auto load_cids = [&](int a, int e, int *c, float *w) {  // this thread's entries of a batch
        a = __builtin_amdgcn_readfirstlane(a);
        e = __builtin_amdgcn_readfirstlane(e);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int k = max(0, min(a + t + s * T, e - 1));
            c[s] = colids[k];
            if (HAS_W) w[s] = w_edge[k];
        }
    };
    int cA[S], cB[S];  // column ids of this thread's entries of batches b+1 and b+2
    float wA[S], wB[S];
    double pAx[S], pAy[S];  // neighbour positions of the entries of batch b+1
    int E1a, E1e, E2a, E2e;  // CSR ranges of batches b+1, b+2
    // stage batch `half`: entries (column ids c/w, positions pAx/pAy) and own rows of batch bb
    auto stage = [&](int bb, int half) {
        const int i = min(bb * B + t, n - 1);
        const int rs = rowptr[i], re = rowptr[i + 1];
        const double2 v = pos[i], r = rep[i];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int k = half * C + t + s * T;
            spos[k] = make_double2(pAx[s], pAy[s]);
            scid[k] = cA[s];
            if (HAS_W) sw[k] = wA[s];
        }
        srs[half * T + t] = rs;
        sre[half * T + t] = re;
        sv[half * T + t] = v;
        sr[half * T + t] = r;
    };
    {
        int E0a, E0e;
        bounds(0, E0a, E0e);
        bounds(1, E1a, E1e);
        bounds(2, E2a, E2e);
        load_cids(E0a, E0e, cA, wA);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double2 q = pos[cA[s]];
            pAx[s] = q.x;
            pAy[s] = q.y;
        }
        stage(0, 0);
        load_cids(E1a, E1e, cA, wA);
    }
    __syncthreads();  // batch 0 staged
    for (int b = 0; b < nb; ++b) {
        if (b > 0) {  // the moves of batch b-1 to memory
            const int c0 = (b - 1) * B;
            if (t < B) {
                pos[c0 + t] = ring[((b + 2) % 3) * T + t];
                eterm[c0 + t] = ering[((b + 1) & 1) * T + t];
            }
        }
        // batch b+1: positions of its entries; batch b+2: column ids; batch b+3: CSR range
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double2 q = pos[cA[s]];
            pAx[s] = q.x;
            pAy[s] = q.y;
        }
        load_cids(E2a, E2e, cB, wB);
        int E3a, E3e;
        bounds(b + 3, E3a, E3e);
        stage(b + 1, (b + 1) & 1);
        // the stores of batch b-1 are complete before the barrier (gathers after it read them)
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // end of batch b
#pragma unroll
        for (int s = 0; s < S; ++s) {
            cA[s] = cB[s];
            if (HAS_W) wA[s] = wB[s];
        }
        E1a = E2a; E1e = E2e;
        E2a = E3a; E2e = E3e;
    }
    // the moves of the last batch
    const int c0 = (nb - 1) * B;
    if (t < B && c0 + t < n) {
        pos[c0 + t] = ring[((nb - 1) % 3) * T + t];
        eterm[c0 + t] = ering[((nb - 1) & 1) * T + t];
    }
}

// ---------------------------------------------------------------------------------------
// 8b. the batch walk over GW_G workgroups, for batches with many CSR entries.  One
// workgroup keeps only a few hundred misses in flight, so a batch of tens of thousands of
// entries with random neighbour positions is bound by that; spreading the entries over
// GW_G workgroups multiplies it, at the price of one grid-wide barrier per batch:
//   A. every workgroup takes a contiguous share of the batch's entries (each thread a
//      contiguous run of them) and sums per row: a row inside one thread -> fbuf; rows
//      spread over threads are combined in the workgroup by a segmented scan over the
//      threads' first-row partials (a fixed tree order); a row spread over workgroups
//      leaves the workgroup's partial of its first or last row (whead / wtail)
//   barrier
//   B. every workgroup combines, in workgroup order, the partials of all rows of the batch,
//      subtracts the repulsion and applies the moves into its own LDS copy (identical in
//      all workgroups); workgroup 0 also writes them to memory.  The next batch reads the
//      positions of this batch from that LDS copy, and older positions from memory, which
//      were written before a barrier that every reader has passed since.
// Loads run a batch ahead: the next batch's rowptr slice, own positions and each thread's
// first GW_K column ids are loaded before the barrier, the neighbour positions they address
// right after it (older batches are then final in memory; this batch's moves are patched
// from the LDS copy), so that the next batch starts on data already in registers.
// The results differ from the one-workgroup walk only in the order of the partial sums.
constexpr int GW_T = 1024;
constexpr int GW_G = 24;  // 8..32 measured on the R-MAT walk: 24 is fastest
constexpr int GW_K = 4;  // entries per thread loaded a batch ahead
constexpr int GW_MIN_ENTRIES = 4096;  // mean CSR entries per batch from which it is used

struct GridWalk {
    const int *rowptr, *colids;
    const int *rows;  // [batch][GW_G][GW_T + 2]: each thread's first row, then the share's first and last
    const float *w_edge;
    double W;
    double2 *pos;
    const double2 *rep;
    double2 *fbuf;
    int n, B;
    double step;
    double *eterm;
    const int *status;
    double2 *whead, *wtail;  // [2][GW_G] by batch parity
    int *wfirst, *wlast;     // [2][GW_G]
    unsigned *bar;
    int *err;
    long long ticks;
};

__global__ __launch_bounds__(GW_T) void k_walk_grid(GridWalk a) {
    constexpr int T = GW_T;
    constexpr int PR = (WALK_MAXB + T) / T;      // rowptr values per thread (a batch has bc + 1)
    constexpr int PP = (WALK_MAXB + T - 1) / T;  // vertices per thread
    __shared__ double2 head[T], tail[T], S[T];
    __shared__ int hrow[T], trow[T];
    __shared__ int rps[WALK_MAXB + 1];
    __shared__ double2 newpos[WALK_MAXB];
    __shared__ double2 bpos[WALK_MAXB];  // the batch's own positions, read before anyone moves them
    __shared__ double2 wsum[T / 64];
    __shared__ int wkfirst[T / 64], wklast[T / 64];
    __shared__ int bad;
    if (__builtin_amdgcn_readfirstlane(*a.status)) return;
    const int t = threadIdx.x, w = blockIdx.x, G = gridDim.x;
    const int lane = t & 63, wv = t >> 6;
    const int n = a.n, B = a.B;
    const int nb = (n + B - 1) / B;
    if (t == 0) bad = 0;
    int pb0 = 0, pb1 = 0;  // the previous batch, whose moved positions are in newpos
    // Loaded a batch ahead, so that no load sits on a batch's critical path: the next batch's
    // rowptr slice and own positions (nrp, nbp), and the column ids, weights and neighbour
    // positions of this thread's first GW_K entries of it (cid, wf, pj).
    int nrp[PR];
    double2 nbp[PP];
    int cid[GW_K];
    float wf[GW_K];
    double2 pj[GW_K];
    // the row of this thread's first entry and the share's first and last rows (host table),
    // loaded a batch ahead as well; the share's rows as vector loads of values equal in all
    // lanes (a uniform load would become a scalar load, whose wait counter LDS shares)
    int fr, frf, frl;
    auto load_rows = [&](int bb) {
        const size_t rb = ((size_t)bb * G + w) * (T + 2);
        int vb;
        __asm__("v_mov_b32 %0, %1" : "=v"(vb) : "s"(0));
        fr = a.rows[rb + t];
        frf = a.rows[rb + T + vb];
        frl = a.rows[rb + T + 1 + vb];
    };
    load_rows(0);
    // prologue: batch 0
    {
        const int cc = min(n, B);
#pragma unroll
        for (int k = 0; k < PR; ++k) nrp[k] = t + k * T <= cc ? a.rowptr[t + k * T] : 0;
#pragma unroll
        for (int k = 0; k < PP; ++k) nbp[k] = t + k * T < cc ? a.pos[t + k * T] : make_double2(0.0, 0.0);
        const int E0 = a.rowptr[0], E1 = a.rowptr[cc];
        const int Cw = (E1 - E0 + G - 1) / G;
        const int Sw = min(E1, E0 + w * Cw), Ew = min(E1, Sw + Cw);
        const int c = max(1, (Cw + T - 1) / T);
        const int s = min(Ew, Sw + t * c), e = min(Ew, s + c);
#pragma unroll
        for (int g = 0; g < GW_K; ++g) {
            cid[g] = s + g < e ? a.colids[s + g] : -1;
            wf[g] = (a.w_edge && s + g < e) ? a.w_edge[s + g] : 1.0f;
        }
#pragma unroll
        for (int g = 0; g < GW_K; ++g) pj[g] = cid[g] < 0 ? make_double2(0.0, 0.0) : a.pos[cid[g]];
    }
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, b1 = min(n, b0 + B), bc = b1 - b0, par = b & 1;
        const bool more = b + 1 < nb;
#pragma unroll
        for (int k = 0; k < PR; ++k)
            if (t + k * T <= bc) rps[t + k * T] = nrp[k];
#pragma unroll
        for (int k = 0; k < PP; ++k)
            if (t + k * T < bc) bpos[t + k * T] = nbp[k];
        __syncthreads();
        // the next batch's entry range, for loading its column ids before the barrier; as
        // vector loads of values equal in all lanes (a uniform load would become a scalar
        // load, whose wait counter the LDS accesses below share)
        const int ncc = more ? min(n, b1 + B) - b1 : 0;
        int nE0, nE1;
        {
            int va, ve;
            __asm__("v_mov_b32 %0, %1" : "=v"(va) : "s"(b1));
            __asm__("v_mov_b32 %0, %1" : "=v"(ve) : "s"(b1 + ncc));
            nE0 = a.rowptr[va];
            nE1 = a.rowptr[ve];
        }
        const int E0 = rps[0], E1 = rps[bc], total = E1 - E0;
        const int Cw = (total + G - 1) / G;
        const int Sw = min(E1, E0 + w * Cw), Ew = min(E1, Sw + Cw);
        const int c = max(1, (Cw + T - 1) / T);
        const int s = min(Ew, Sw + t * c), e = min(Ew, s + c);
        // A1. thread partials
        int hr = -1, tr = -1;
        double2 hv = make_double2(0.0, 0.0), tv = make_double2(0.0, 0.0);
        if (s < e) {
            int r = fr;  // row_of(s)
            const int first = r;
            int rend = rps[r + 1 - b0];
            double2 pi = bpos[r - b0];
            double ax = 0.0, ay = 0.0;
            for (int k0 = s; k0 < e; k0 += GW_K) {
                if (k0 != s) {  // entries beyond the first GW_K: loaded here
#pragma unroll
                    for (int g = 0; g < GW_K; ++g) {
                        cid[g] = k0 + g < e ? a.colids[k0 + g] : -1;
                        wf[g] = (a.w_edge && k0 + g < e) ? a.w_edge[k0 + g] : 1.0f;
                    }
#pragma unroll
                    for (int g = 0; g < GW_K; ++g) pj[g] = cid[g] < 0 ? make_double2(0.0, 0.0) : a.pos[cid[g]];
                }
                // a neighbour in the previous batch: its move, which memory may not show yet
#pragma unroll
                for (int g = 0; g < GW_K; ++g)
                    if (cid[g] >= pb0 && cid[g] < pb1) pj[g] = newpos[cid[g] - pb0];
#pragma unroll
                for (int g = 0; g < GW_K; ++g) {
                    const int k = k0 + g;
                    if (k >= e) break;
                    while (k >= rend) {  // row r complete inside this thread's run
                        if (r == first) {
                            hr = r;
                            hv = make_double2(ax, ay);
                        } else {
                            store_through(a.fbuf + r, make_double2(ax, ay));
                        }
                        ++r;
                        rend = rps[r + 1 - b0];
                        pi = bpos[r - b0];
                        ax = ay = 0.0;
                    }
                    const double dx = pj[g].x - pi.x, dy = pj[g].y - pi.y;
                    const double wk = a.w_edge ? a.W * (double)wf[g] : a.W;
                    const double sc = wk * sqrt(dx * dx + dy * dy);
                    ax = fma(dx, sc, ax);
                    ay = fma(dy, sc, ay);
                }
            }
            if (r == first) {
                hr = r;
                hv = make_double2(ax, ay);
            } else {
                tr = r;
                tv = make_double2(ax, ay);
            }
        }
        hrow[t] = hr;
        trow[t] = tr;
        head[t] = hv;
        tail[t] = tv;
        // A2. segmented inclusive scan of the first-row partials over the threads (runs of
        // equal hrow are contiguous): in the wave by DPP moves, then across waves
        double2 v = hv;
        seg_step<0x111, 0xF>(v, hr);  // row_shr:1
        seg_step<0x112, 0xF>(v, hr);  // row_shr:2
        seg_step<0x114, 0xF>(v, hr);  // row_shr:4
        seg_step<0x118, 0xF>(v, hr);  // row_shr:8
        seg_step<0x142, 0xA>(v, hr);  // row_bcast:15 into rows 1 and 3
        seg_step<0x143, 0xC>(v, hr);  // row_bcast:31 into rows 2 and 3
        if (lane == 63) {
            wsum[wv] = v;
            wklast[wv] = hr;
        }
        if (lane == 0) wkfirst[wv] = hr;
        __syncthreads();
        double2 carry = make_double2(0.0, 0.0);
        for (int q = wv - 1; q >= 0; --q) {  // earlier waves of the same run, nearest first
            if (wklast[q] != hr) break;
            carry = add2(carry, wsum[q]);
            if (wkfirst[q] != hr) break;
        }
        S[t] = add2(carry, v);
        __syncthreads();
        // A3. rows of this workgroup's share: complete ones -> fbuf, the first and last
        // (when they continue in other workgroups) -> whead / wtail
        if (Sw < Ew) {
            const int rf = frf, rl = frl;  // row_of(Sw), row_of(Ew - 1)
            if (t == 0) {
                store_through(a.wfirst + par * GW_G + w, rf);
                store_through(a.wlast + par * GW_G + w, rl != rf ? rl : -1);
            }
            for (int i = rf + t; i <= rl; i += T) {
                const int r0 = rps[i - b0], r1 = rps[i + 1 - b0];
                const int rs = max(r0, Sw), re = min(r1, Ew);
                if (re <= rs) continue;
                const int ta = (rs - Sw) / c, tb = (re - 1 - Sw) / c;
                if (ta == tb && hrow[ta] != i && trow[ta] != i) continue;  // already in fbuf
                const double2 val = hrow[ta] == i ? S[tb] : (tb > ta ? add2(tail[ta], S[tb]) : tail[ta]);
                if (r0 >= Sw && r1 <= Ew) {
                    store_through(a.fbuf + i, val);
                } else if (i == rf) {
                    store_through(a.whead + par * GW_G + w, val);
                } else {
                    store_through(a.wtail + par * GW_G + w, val);
                }
            }
        }
        // the next batch's first entries of this thread: column ids and weights now, the
        // positions they address after the barrier.  These loads, and those after the
        // barrier, are unconditional with clamped indices (a slot past the thread's entries,
        // or past the last batch, loads some valid entry and is never used): loads under a
        // branch would make the compiler wait for every load in flight where it rejoins.
        {
            const int nCw = (nE1 - nE0 + G - 1) / G;
            const int nSw = min(nE1, nE0 + w * nCw), nEw = min(nE1, nSw + nCw);
            const int nc = max(1, (nCw + T - 1) / T);
            const int ns = min(nEw, nSw + t * nc), ne = min(nEw, ns + nc);
#pragma unroll
            for (int g = 0; g < GW_K; ++g) {
                const int kk = max(0, min(ns + g, ne - 1));
                const int cv = a.colids[kk];
                cid[g] = ns + g < ne ? cv : -1;
            }
            if (a.w_edge) {
#pragma unroll
                for (int g = 0; g < GW_K; ++g) wf[g] = a.w_edge[max(0, min(ns + g, ne - 1))];
            }
        }
        // grid barrier: every workgroup's share of the batch is summed.  What other workgroups
        // read (row sums, partials, moved positions) was stored through to memory; each wave
        // waits for its stores to complete, so thread 0's relaxed arrival publishes them with no
        // L2 write-back (a release here wrote the whole L2 back at every batch).
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) {
            __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = (unsigned)G * (unsigned)(b + 1);
            const long long t0 = wall_clock64();
            // polled relaxed (an acquiring load would invalidate the L2 at every poll), then
            // one acquire fence
            while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    bad = 1;
                    break;
                }
                if (wall_clock64() - t0 > a.ticks) {
                    atomicOr(a.err, 1);
                    bad = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        // the same value in every thread: a uniform exit, not a per-lane branch out of the loop
        if (__builtin_amdgcn_readfirstlane(bad)) return;
        // B. combine in workgroup order, subtract the repulsion, move (every workgroup).  The
        // partials and repulsions are loaded first: loads complete in issue order, so the
        // next batch's loads, issued after them, do not delay them.
        double2 fv[PP], rv[PP];
#pragma unroll
        for (int k = 0; k < PP; ++k) {
            const int ic = min(b0 + t + k * T, b1 - 1);
            fv[k] = a.fbuf[ic];  // the row's sum when it lies in one workgroup's share
            rv[k] = a.rep[ic];
        }
        // rows over several workgroups' shares (hubs): combined before the next batch's loads
        // are issued, so that their loads are older than those too
#pragma unroll
        for (int k = 0; k < PP; ++k) {
            const int i = b0 + t + k * T;
            if (i >= b1) break;
            const int r0 = rps[i - b0], r1 = rps[i + 1 - b0];
            if (r1 <= r0) continue;
            const int wa = (r0 - E0) / Cw, wb = (r1 - 1 - E0) / Cw;
            if (wa != wb) {
                double2 f = a.wfirst[par * GW_G + wa] == i ? a.whead[par * GW_G + wa] : a.wtail[par * GW_G + wa];
                for (int q = wa + 1; q <= wb; ++q) f = add2(f, a.whead[par * GW_G + q]);
                fv[k] = f;
            }
        }
        // The next batch's rows and neighbour positions, in flight during the moves.  Every
        // batch before this one is final in memory (written before this barrier); this
        // batch's moves, made below, are taken from newpos when the next batch is summed.
#pragma unroll
        for (int k = 0; k < PR; ++k) nrp[k] = a.rowptr[min(b1 + t + k * T, n)];
#pragma unroll
        for (int k = 0; k < PP; ++k) nbp[k] = a.pos[min(b1 + t + k * T, n - 1)];
#pragma unroll
        for (int g = 0; g < GW_K; ++g) pj[g] = a.pos[max(cid[g], 0)];
        load_rows(min(b + 1, nb - 1));
#pragma unroll
        for (int k = 0; k < PP; ++k) {
            const int i = b0 + t + k * T;
            if (i >= b1) break;
            double2 f = rps[i + 1 - b0] > rps[i - b0] ? fv[k] : make_double2(0.0, 0.0);
            f.x -= rv[k].x;  // f - calcRepForce(x_i), :1059
            f.y -= rv[k].y;
            // x_i += (F/|F|) * STEP (:1062-1066), the unit vector F * rsqrt2(|F|^2)
            const double m2 = f.x * f.x + f.y * f.y;
            const double rs = rsqrt2(m2) * a.step;
            double2 q = bpos[i - b0];
            q.x = fma(f.x, rs, q.x);
            q.y = fma(f.y, rs, q.y);
            newpos[i - b0] = q;
            if (w == 0) {
                store_through(a.pos + i, q);
                a.eterm[i] = m2;
            }
        }
        __syncthreads();
        pb0 = b0;
        pb1 = b1;
    }
}

// ---------------------------------------------------------------------------------------
// host orchestration
struct TreeBufs {
    uint64_t *keys;  // sorted keys
    int *perm;       // sorted vertex ids
    double2 *spos;
    BHNode *nodes;
    int *misc;       // misc[0] status, misc[1] root, misc[2] node count
    double *frame;
};

// Exclusive scan of in[0, n) into out, the total at bsum[nb].  `in` must hold at least
// SCAN_SINGLE_MAX ints whatever n is (k_scan_single reads all its tiles; ensure_bh sizes the
// boundary-flag buffer so).
int scan_excl(const int *in, int n, int *out, int *bsum, hipStream_t s) {
    const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (n <= SCAN_SINGLE_MAX) {  // one launch instead of three
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, in, n, out, bsum + nb);
        BL_HIP(hipGetLastError());
        return BL_OK;
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(WG), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(WG), 0, s, bsum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(WG), 0, s, in, n, bsum, out);
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int ensure_bh(BHWork &w, int n) {
    if (w.n_alloc >= n) return BL_OK;
    const size_t N = (size_t)n;
    // every size up to n: the tiles are never smaller than WG x SORT_ROUNDS keys
    const int nblk = (n + WG * SORT_ROUNDS - 1) / (WG * SORT_ROUNDS);
    const int nsb = (n + SCAN_TILE - 1) / SCAN_TILE + 1;
    if (w.keys_a.ensure(8 * N) || w.keys_b.ensure(8 * N) || w.idx_a.ensure(4 * N) || w.idx_b.ensure(4 * N) ||
        w.hist.ensure(4 * (size_t)256 * nblk) || w.scan_tmp.ensure(4 * 256) || w.spos.ensure(16 * N) ||
        w.frame.ensure(sizeof(double4) * 1024 + 64) || w.isrep.ensure(4 * std::max(N, (size_t)SCAN_SINGLE_MAX)) ||
        w.nodeid.ensure(4 * N) || w.node_child.ensure(sizeof(BHNode) * N) || w.rep.ensure(16 * N) ||
        w.misc.ensure(4 * (size_t)(64 + nsb)) || w.node_lo.ensure(4 * N) || w.node_cnt.ensure(4 * N))
        return BL_E_NOMEM;
    w.n_alloc = n;
    return BL_OK;
}

// Builds the tree of the positions d_pos (device, n points) into w; returns BL_OK and
// leaves the status word in misc[0] (checked by the caller when it needs it).  centroids: the
// unweighted centroids of algo 4's traversal (the far-field modes summarise cells themselves).
int build_tree(BHWork &w, const double2 *d_pos, int n, hipStream_t s, uint32_t *d_xs, uint32_t *d_ys,
               bool centroids = true) {
    int rc = ensure_bh(w, n);
    if (rc) return rc;
    int *misc = w.misc.as<int>();
    double *frame = w.frame.as<double>();
    double4 *fpart = reinterpret_cast<double4 *>(frame + 8);
    BL_HIP(hipMemsetAsync(misc, 0, 16, s));
    const int gf = std::min(1024, (n + WG - 1) / WG);
    hipLaunchKernelGGL(k_frame_partial, dim3(gf), dim3(WG), 0, s, d_pos, n, fpart);
    hipLaunchKernelGGL(k_frame_final, dim3(1), dim3(WG), 0, s, fpart, gf, frame);
    uint64_t *ka = w.keys_a.as<uint64_t>(), *kb = w.keys_b.as<uint64_t>();
    int *ia = w.idx_a.as<int>(), *ib = w.idx_b.as<int>();
    const int g1 = (n + WG - 1) / WG;
    hipLaunchKernelGGL(k_keys, dim3(g1), dim3(WG), 0, s, d_pos, n, frame, ka, ia, d_xs, d_ys);
    const int rounds = SORT_ROUNDS;
    const int nblk = (n + WG * rounds - 1) / (WG * rounds);
    double2 *spos = w.spos.as<double2>();
    int *hist = w.hist.as<int>();
    int *dtot = w.scan_tmp.as<int>();
    int *bsum = misc + 64;
    for (int pass = SORT_SKIP / 8; pass < 8; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(WG), 0, s, ka, n, shift, nblk, rounds, hist);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(256), dim3(WG), 0, s, hist, nblk, dtot);
        if (pass < 7)
            hipLaunchKernelGGL((k_sort_scatter<SORT_ROUNDS, false>), dim3(nblk), dim3(WG), 0, s, ka, ia, kb, ib, n,
                               shift, nblk, hist, dtot, d_pos, spos);
        else
            hipLaunchKernelGGL((k_sort_scatter<SORT_ROUNDS, true>), dim3(nblk), dim3(WG), 0, s, ka, ia, kb, ib, n,
                               shift, nblk, hist, dtot, d_pos, spos);
        std::swap(ka, kb);
        std::swap(ia, ib);
    }
    // after 6 passes the data is back in keys_a / idx_a, the positions in spos, in the order of
    // the keys' top 48 bits; runs of keys equal in those bits are put in full-key order here
    static_assert((8 - SORT_SKIP / 8) % 2 == 0, "an even number of passes ends in keys_a");
    if (n >= 2) hipLaunchKernelGGL(k_sort_ties, dim3(g1), dim3(WG), 0, s, ka, ia, spos, n);
    if (n >= 2) {
        int *isrep = w.isrep.as<int>(), *nodeid = w.nodeid.as<int>();
        BHNode *nodes = w.node_child.as<BHNode>();
        hipLaunchKernelGGL(k_tree_bounds, dim3(g1), dim3(WG), 0, s, ka, n, isrep, misc);
        rc = scan_excl(isrep, n - 1, nodeid, bsum, s);
        if (rc) return rc;
        const int nsb = (n - 1 + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_tree_root, dim3(1), dim3(64), 0, s, ka, n, nodeid, bsum + nsb, misc);
        int *npts = w.node_cnt.as<int>();
        hipLaunchKernelGGL(k_tree_nodes, dim3(g1), dim3(WG), 0, s, ka, n, isrep, nodeid, nodes, npts,
                           w.node_lo.as<int>(), misc, misc);
        // n-1 boundaries give at most n-1 nodes; threads beyond the node count (misc[2]) exit
        if (centroids) {
            hipLaunchKernelGGL(k_centroid_small, dim3(g1), dim3(WG), 0, s, nodes, spos, npts, misc);
            hipLaunchKernelGGL(k_centroid_top, dim3(g1), dim3(WG), 0, s, nodes, spos, npts, misc);
        }
    } else {
        const int neg = -1;
        BL_HIP(hipMemcpyAsync(misc + 1, &neg, sizeof(int), hipMemcpyHostToDevice, s));
    }
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int launch_rep(BHWork &w, int n, double theta, int precision, hipStream_t s) {
    const int waves = (n + 63) / 64;
    const int g = (waves + WG / 64 - 1) / (WG / 64);
    if (precision == BL_PREC_FP64)
        hipLaunchKernelGGL(k_rep64, dim3(g), dim3(WG), 0, s, w.spos.as<double2>(), w.idx_a.as<int>(), n,
                           w.spos.as<double2>(), w.node_child.as<BHNode>(), w.misc.as<int>(), w.frame.as<double>(), theta,
                           w.rep.as<double2>(), w.misc.as<int>());
    else
        hipLaunchKernelGGL(precision == BL_PREC_FP32 ? k_rep32<false> : k_rep32<true>, dim3(g), dim3(WG), 0, s,
                           w.spos.as<double2>(), w.idx_a.as<int>(), n, w.spos.as<double2>(), w.node_child.as<BHNode>(),
                           w.misc.as<int>(), w.frame.as<double>(), theta, w.rep.as<double2>(), w.misc.as<int>());
    BL_HIP(hipGetLastError());
    return BL_OK;
}

}  // namespace

namespace {
// layout of BHWork::gwalk: whead[2G] wtail[2G] (double2), wfirst[2G] wlast[2G] (int), bar, err
constexpr size_t GW_BYTES = 2 * (2 * GW_G * sizeof(double2)) + 2 * (2 * GW_G * sizeof(int)) + 2 * sizeof(int);
int *gwalk_err(BHWork &w) {
    return reinterpret_cast<int *>(w.gwalk.as<char>() + GW_BYTES - sizeof(int));
}

// Launches one iteration's walk on GW_G co-resident workgroups; anything but BL_OK means
// nothing was launched (the caller uses the one-workgroup walk).
// The multi-workgroup walk's row lookup for batch size B (GridWalk::rows), built on the host
// once per graph and batch size: the same shares as the kernel computes; for each thread the
// last row r of the batch with rowptr[r] <= its first entry (-1 without entries), and the
// rows of the share's first and last entries.
// Past GW_ROWS_MAX_BYTES (about 11,000 batches) it is not built and the run takes the
// one-workgroup walk instead.
constexpr size_t GW_ROWS_MAX_BYTES = size_t(1) << 30;

int grid_rows_build(bl_graph *g, int B) {
    if (g->grid_rows_batch == B) return BL_OK;
    const int n = g->n, G = GW_G, T = GW_T, nb = (n + B - 1) / B;
    if ((size_t)nb * G * (T + 2) * sizeof(int32_t) > GW_ROWS_MAX_BYTES) return BL_E_NOMEM;
    const int32_t *rp = g->h_rowptr.data();
    std::vector<int32_t> tab((size_t)nb * G * (T + 2));
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, b1 = std::min(n, b0 + B);
        const int E0 = rp[b0], E1 = rp[b1];
        const int Cw = (E1 - E0 + G - 1) / G, c = std::max(1, (Cw + T - 1) / T);
        auto row_of = [&](int entry) { return (int)(std::upper_bound(rp + b0, rp + b1, entry) - rp) - 1; };
        for (int w = 0; w < G; ++w) {
            int32_t *o = &tab[((size_t)b * G + w) * (T + 2)];
            const int Sw = std::min(E1, E0 + w * Cw), Ew = std::min(E1, Sw + Cw);
            int r = b0;  // first entries rise with t: the row advances monotonically
            for (int t = 0; t < T; ++t) {
                const int s0 = std::min(Ew, Sw + t * c), e0 = std::min(Ew, s0 + c);
                if (s0 >= e0) {
                    o[t] = -1;
                    continue;
                }
                while (r + 1 < b1 && rp[r + 1] <= s0) ++r;
                o[t] = r;
            }
            o[T] = Sw < Ew ? row_of(Sw) : -1;
            o[T + 1] = Sw < Ew ? row_of(Ew - 1) : -1;
        }
    }
    if (g->grid_rows.ensure(tab.size() * sizeof(int32_t))) return BL_E_NOMEM;
    BL_HIP(hipMemcpy(g->grid_rows.p, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    g->grid_rows_batch = B;
    return BL_OK;
}

int launch_walk_grid(bl_graph *g, double2 *pos, const bl_params *p, double step, hipStream_t s) {
    BHWork &w = g->bh;
    if (w.gwalk.ensure(GW_BYTES)) return BL_E_NOMEM;
    if (int rc = grid_rows_build(g, p->batch)) return rc;
    char *base = w.gwalk.as<char>();
    GridWalk a;
    a.rowptr = g->rowptr.as<int>();
    a.colids = g->colids.as<int>();
    a.rows = g->grid_rows.as<int>();
    a.w_edge = g->has_w ? g->w_edge.as<float>() : nullptr;
    a.W = p->weight;
    a.pos = pos;
    a.rep = w.rep.as<double2>();
    a.fbuf = g->scratch.as<double2>();
    a.n = g->n;
    a.B = p->batch;
    a.step = step;
    a.eterm = g->eterm.as<double>();
    a.status = w.misc.as<int>();
    a.whead = reinterpret_cast<double2 *>(base);
    a.wtail = a.whead + 2 * GW_G;
    a.wfirst = reinterpret_cast<int *>(a.wtail + 2 * GW_G);
    a.wlast = a.wfirst + 2 * GW_G;
    a.bar = reinterpret_cast<unsigned *>(a.wlast + 2 * GW_G);
    a.err = reinterpret_cast<int *>(a.bar + 1);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0)
        khz = 100000;
    a.ticks = (long long)khz * 2000;  // 2 s per barrier wait
    // clears the barrier counter, and the error flag unless an earlier iteration of this run set it
    if (hipMemsetAsync(a.bar, 0, sizeof(unsigned), s) != hipSuccess) return BL_E_DEVICE;
    if (!w.gwalk_used && hipMemsetAsync(a.err, 0, sizeof(int), s) != hipSuccess) return BL_E_DEVICE;
    void *args[] = {&a};
    const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_walk_grid), dim3(GW_G),
                                                    dim3(GW_T), args, 0, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BL_E_DEVICE;
    }
    w.gwalk_used = 1;
    return BL_OK;
}
}  // namespace

int bh_build_tree(BHWork &w, const double2 *d_pos, int n, hipStream_t s) {
    return build_tree(w, d_pos, n, s, nullptr, nullptr, false);
}

int run_bh(bl_graph *g, double *d_xy, const bl_params *p, double step0, int iters, double theta, bl_stats *st,
           int far_mode) {
    const int n = g->n;
    hipStream_t s = g->stream;
    double2 *pos = reinterpret_cast<double2 *>(d_xy);
    if (g->eterm.ensure(sizeof(double) * (size_t)n) || g->energy.ensure(sizeof(double) * (4 + 256)) ||
        g->scratch.ensure(sizeof(double2) * (size_t)n))
        return BL_E_NOMEM;
    int rc = ensure_bh(g->bh, n);
    if (rc) return rc;
    Prof prof;
    prof.on = p->profile != 0;
    prof.s = s;
    double STEP = step0;
    // the batch walk over several workgroups pays off for batches with many entries;
    // BL_WALK=grid|single|pf|row overrides the choice (single: the best one-workgroup walk)
    const int nbt = (n + p->batch - 1) / p->batch;
    bool use_grid = p->precision != BL_PREC_FP64 && p->batch <= WALK_MAXB &&
                    (double)g->nnz / (double)nbt >= (double)GW_MIN_ENTRIES;
    bool use_row = p->precision != BL_PREC_FP64 && row_walk_fits(g, p->batch);
    // where the multi-workgroup walk would run, the wide lagged walk (helper workgroups) runs
    // instead when its plan fits; BL_WALK=grid keeps the multi-workgroup walk, BL_WALK=wide
    // takes the wide walk wherever it fits
    bool want_wide = use_grid;
    if (const char *wenv = getenv("BL_WALK")) {
        if (!strcmp(wenv, "grid")) use_grid = p->precision != BL_PREC_FP64 && p->batch <= WALK_MAXB, want_wide = false;
        else if (!strcmp(wenv, "single")) use_grid = false, want_wide = false;
        else if (!strcmp(wenv, "pf")) use_grid = false, use_row = false, want_wide = false;
        else if (!strcmp(wenv, "row")) use_grid = false, want_wide = false;
        else if (!strcmp(wenv, "wide")) want_wide = p->precision != BL_PREC_FP64;
    }
    const bool use_wide = want_wide && wide_walk_fits(g, p->batch);
    int wide_h = 255;
    if (const char *v = getenv("BL_WIDE_H")) wide_h = std::max(1, atoi(v));
    if (use_wide) {
        rc = wide_walk_prepare(g->has_w);
        if (rc) return rc;
        use_grid = false;
    }
    if (use_grid || use_wide) use_row = false;
    // the lagged walk (helper workgroups) where the row walk fits and the graph has enough batches;
    // BL_LAG_D / BL_LAG_H set its lag and helper count
    int lag_d = 0, lag_h = 32;
    if (const char *v = getenv("BL_LAG_D")) lag_d = atoi(v);
    if (const char *v = getenv("BL_LAG_H")) lag_h = std::max(0, atoi(v));
    if (use_row && lag_d == 0) lag_d = lag_walk_choose(g, p->batch);
    bool use_lag = use_row && lag_d > 0 && lag_walk_fits(g, p->batch, lag_d);
    if (const char *wenv = getenv("BL_WALK"))  // row and single: one workgroup, no helpers
        if (!strcmp(wenv, "row") || !strcmp(wenv, "single")) use_lag = false;
    if (use_lag) {
        rc = lag_walk_prepare(g->has_w);
        if (rc) return rc;
    }
    if (use_row) {
        const size_t lb = wrow_lds(g->has_w);
        BL_HIP(hipFuncSetAttribute(g->has_w ? reinterpret_cast<const void *>(k_walk_row<true>)
                                            : reinterpret_cast<const void *>(k_walk_row<false>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb));
    }
    g->bh.gwalk_used = 0;
    int walk = BL_WALK_NONE;
    // cells accepted by the far-field traversal, counted per iteration in misc[8..9]
    unsigned long long *ncells = reinterpret_cast<unsigned long long *>(g->bh.misc.as<int>() + 8);
    // the learned far field is judged every judge_every iterations from the first (BL_FAR_JUDGE=k
    // sets the period; 0 keeps the model on without judging)
    int judge_every = 10;
    if (const char *je = getenv("BL_FAR_JUDGE")) judge_every = atoi(je);
    const bool judged = far_mode == FAR_LEARNED && judge_every > 0 && n >= 2;
    int verdict = 1;  // the judge's last verdict (1: the model)
    if (judged) {
        rc = far_judge_reset(g->bh, s);
        if (rc) return rc;
    }
    // the exact mode builds no tree, so nothing else clears the status word the walk reads
    if (far_mode == FAR_EXACT) BL_HIP(hipMemsetAsync(g->bh.misc.p, 0, 16, s));
    for (int it = 0; it < iters; ++it) {
        if (far_mode != FAR_EXACT) {
            prof.begin(3);
            // tree of the iteration-start positions (centroids for algo 4's traversal only)
            rc = build_tree(g->bh, pos, n, s, nullptr, nullptr, far_mode == 0);
            prof.end();
            if (rc) return rc;
        }
        const float *model = g->has_far_model ? g->far_model.as<float>() : nullptr;
        if (far_mode && far_mode != FAR_EXACT) {
            prof.begin(4);
            rc = far_cells_of(g->bh, n, s);
            prof.end();
            if (rc) return rc;
            BL_HIP(hipMemsetAsync(ncells, 0, sizeof(unsigned long long), s));
        }
        if (judged && it % judge_every == 0) {
            prof.begin(6);
            rc = far_judge(g->bh, pos, n, theta, model, it, s, &verdict);
            prof.end();
            if (rc) return rc;
        }
        // the learned mode takes the analytic second-order kernel while the model's last verdict
        // was against it
        const int mode_now = (far_mode == FAR_LEARNED && judged && !verdict) ? (int)FAR_QUAD : far_mode;
        prof.begin(0);
        rc = far_mode ? far_repulsion(g->bh, pos, n, theta, mode_now, model, s, ncells, judged)
                      : launch_rep(g->bh, n, theta, p->precision, s);
        prof.end();
        if (rc) return rc;
        if (p->precision != BL_PREC_FP64 && !use_grid && use_row) {
            // the row walk's pre-pass: attraction of the neighbours in the same or a later batch
            prof.begin(5);
            hipLaunchKernelGGL(k_late_attract, dim3((n + 255) / 256), dim3(256), 0, s, g->late_rp.as<int>(),
                               g->late_ci.as<int>(), g->has_w ? g->late_w.as<float>() : nullptr, p->weight, pos,
                               g->bh.rep.as<double2>(), n);
            prof.end();
        }
        prof.begin(1);
        if (p->precision == BL_PREC_FP64) {
            walk = BL_WALK_ROW_ORDER;
            hipLaunchKernelGGL(k_attract_move, dim3(1), dim3(1024), 0, s, g->rowptr.as<int>(), g->colids.as<int>(),
                               g->has_w ? g->w_edge.as<float>() : nullptr, p->weight, pos, g->bh.rep.as<double2>(),
                               g->scratch.as<double2>(), n, p->batch, STEP, g->eterm.as<double>(),
                               g->bh.misc.as<int>());
        } else if (use_wide) {
            walk = BL_WALK_WIDE;
            rc = launch_walk_wide(g, pos, p, STEP, wide_h, s);
            if (rc) return rc;
        } else if (use_grid && launch_walk_grid(g, pos, p, STEP, s) == BL_OK) {
            walk = BL_WALK_GRID;
        } else if (use_lag) {
            walk = BL_WALK_LAG;
            rc = launch_walk_lag(g, pos, p, STEP, lag_h, s);
            if (rc) return rc;
        } else if (use_row) {
            walk = BL_WALK_ROW;
            hipLaunchKernelGGL(g->has_w ? k_walk_row<true> : k_walk_row<false>, dim3(1), dim3(2 * WROW_T),
                               wrow_lds(g->has_w), s, g->early_rp.as<int>(), g->early_ci.as<int>(),
                               g->has_w ? g->early_w.as<float>() : nullptr, p->weight, pos, g->bh.rep.as<double2>(),
                               n, p->batch, STEP, g->eterm.as<double>(), g->bh.misc.as<int>());
        } else if (p->batch <= WALK_MAXB) {
            walk = BL_WALK_PF;
            hipLaunchKernelGGL(k_walk_pf, dim3(1), dim3(WALK_T), 0, s, g->rowptr.as<int>(), g->colids.as<int>(),
                               g->has_w ? g->w_edge.as<float>() : nullptr, p->weight, pos, g->bh.rep.as<double2>(), n,
                               p->batch, STEP, g->eterm.as<double>(), g->bh.misc.as<int>());
        } else {
            walk = BL_WALK_SPLIT;
            hipLaunchKernelGGL(k_walk<false>, dim3(1), dim3(WALK_T), 0, s, g->rowptr.as<int>(), g->colids.as<int>(),
                               g->has_w ? g->w_edge.as<float>() : nullptr, p->weight, pos, g->bh.rep.as<double2>(),
                               g->scratch.as<double2>(), n, p->batch, STEP, g->eterm.as<double>(),
                               g->bh.misc.as<int>());
        }
        prof.end();
        prof.begin(2);
        rc = launch_energy(g->eterm.as<double>(), n, g->energy.as<double>(), p->precision, s);
        prof.end();
        if (rc) return rc;
        prof.flush();
        STEP = STEP * 0.999;  // :1070
    }
    prof.report(st);
    if (st) st->walk = walk;
    unsigned long long cells = 0;
    if (far_mode && far_mode != FAR_EXACT && iters > 0)
        BL_HIP(hipMemcpyAsync(&cells, ncells, sizeof(cells), hipMemcpyDeviceToHost, s));
    int helped = 0;
    if (walk == BL_WALK_LAG) BL_HIP(hipMemcpyAsync(&helped, lag_walk_helped(g), sizeof(int), hipMemcpyDeviceToHost, s));
    int model_iters = far_mode == FAR_LEARNED ? iters : 0;
    float judge_err[2] = {0.0f, 0.0f};
    if (judged && iters > 0) {
        rc = far_judge_read(g->bh, &model_iters, judge_err, s);
        if (rc) return rc;
    }
    int status = 0;
    double E = 1.79769313486231570815e+308;  // ENERGY starts at DBL_MAX (:1024)
    if (iters > 0) {
        BL_HIP(hipMemcpyAsync(&status, g->bh.misc.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
        BL_HIP(hipMemcpyAsync(&E, g->energy.as<double>(), sizeof(double), hipMemcpyDeviceToHost, s));
    }
    int gerr = 0, werr = 0, waited = 0;
    if (g->bh.gwalk_used) BL_HIP(hipMemcpyAsync(&gerr, gwalk_err(g->bh), sizeof(int), hipMemcpyDeviceToHost, s));
    if (walk == BL_WALK_WIDE && (rc = wide_walk_read(g, &werr, &waited, s))) return rc;
    BL_HIP(hipStreamSynchronize(s));
    if (status & ST_DUPKEY) return BL_E_DUPKEY;
    if (gerr) return BL_E_DEVICE;  // a grid barrier timed out (the walk stopped early)
    if (werr) {  // a wait of the wide walk timed out: its chunk counts are incomplete
        (void)wide_walk_reset(g);
        return BL_E_DEVICE;
    }
    if (walk == BL_WALK_WIDE) helped = nbt - waited;
    if (st) {
        st->far_cells = (int64_t)cells;
        st->far_model_iters = model_iters;
        st->walk_helped = helped;
        st->walk_lag = walk == BL_WALK_LAG ? lag_d : walk == BL_WALK_WIDE ? g->wide_dmax : 0;
        st->far_judge_err[0] = judge_err[0];
        st->far_judge_err[1] = judge_err[1];
        st->energy = E;
        st->iters_done = iters;
        st->step_final = STEP;
    }
    return BL_OK;
}

// ---- host entry points for the pieces ----------------------------------------------------
namespace {
struct Scoped {
    hipStream_t s = nullptr;
    BHWork w;
    void *buf = nullptr;
    ~Scoped() {
        if (s) (void)hipStreamSynchronize(s);
        w.release();
        if (buf) (void)hipFree(buf);
        if (s) (void)hipStreamDestroy(s);
    }
};
}  // namespace

int bh_morton_host(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin) {
    Scoped sc;
    BL_HIP(hipStreamCreateWithFlags(&sc.s, hipStreamNonBlocking));
    const size_t pb = 16 * (size_t)n;
    BL_HIP(hipMalloc(&sc.buf, pb + 8 * (size_t)n));
    double2 *dpos = static_cast<double2 *>(sc.buf);
    uint32_t *dxs = reinterpret_cast<uint32_t *>(static_cast<char *>(sc.buf) + pb);
    uint32_t *dys = dxs + n;
    BL_HIP(hipMemcpyAsync(dpos, xy, pb, hipMemcpyHostToDevice, sc.s));
    int rc = build_tree(sc.w, dpos, n, sc.s, dxs, dys);
    if (rc) return rc;
    BL_HIP(hipMemcpyAsync(xs, dxs, 4 * (size_t)n, hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipMemcpyAsync(ys, dys, 4 * (size_t)n, hipMemcpyDeviceToHost, sc.s));
    double fr[3];
    BL_HIP(hipMemcpyAsync(fr, sc.w.frame.p, sizeof(fr), hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipStreamSynchronize(sc.s));
    if (maxmin) *maxmin = fr[2];
    return BL_OK;
}

int bh_tree_host(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes, int32_t *root,
                 double *diam, double *cx, double *cy, int32_t *child) {
    Scoped sc;
    BL_HIP(hipStreamCreateWithFlags(&sc.s, hipStreamNonBlocking));
    const size_t pb = 16 * (size_t)n;
    BL_HIP(hipMalloc(&sc.buf, pb));
    double2 *dpos = static_cast<double2 *>(sc.buf);
    BL_HIP(hipMemcpyAsync(dpos, xy, pb, hipMemcpyHostToDevice, sc.s));
    int rc = build_tree(sc.w, dpos, n, sc.s, nullptr, nullptr);
    if (rc) return rc;
    int misc[3];
    double fr[3];
    BL_HIP(hipMemcpyAsync(misc, sc.w.misc.p, sizeof(misc), hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipMemcpyAsync(fr, sc.w.frame.p, sizeof(fr), hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipMemcpyAsync(perm, sc.w.idx_a.p, 4 * (size_t)n, hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipStreamSynchronize(sc.s));
    const int m = n >= 2 ? misc[2] : 0;
    if (misc[0] & ST_DUPKEY) return BL_E_DUPKEY;
    *nnodes = m;
    *root = n >= 2 ? misc[1] : -1;
    if (m > max_nodes) return BL_E_ARG;
    if (m > 0) {
        std::vector<BHNode> h(m);
        BL_HIP(hipMemcpyAsync(h.data(), sc.w.node_child.p, sizeof(BHNode) * m, hipMemcpyDeviceToHost, sc.s));
        BL_HIP(hipStreamSynchronize(sc.s));
        for (int v = 0; v < m; ++v) {
            if (diam) diam[v] = ldexp(fr[2], -h[v].depth);
            if (cx) cx[v] = h[v].c.x;
            if (cy) cy[v] = h[v].c.y;
            const int c[4] = {h[v].ch.x, h[v].ch.y, h[v].ch.z, h[v].ch.w};
            for (int q = 0; q < 4 && child; ++q)
                child[4 * v + q] = c[q] <= -2 ? -(perm[-c[q] - 2] + 2) : c[q];
        }
    }
    return BL_OK;
}

int bh_forces_host(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                   int32_t precision) {
    // Builds the tree of xy, then traverses it for the nq query positions.
    Scoped sc;
    BL_HIP(hipStreamCreateWithFlags(&sc.s, hipStreamNonBlocking));
    const size_t pb = 16 * (size_t)n, qb = 16 * (size_t)nq;
    BL_HIP(hipMalloc(&sc.buf, pb + qb + 4 * (size_t)nq + 16 * (size_t)nq));
    double2 *dpos = static_cast<double2 *>(sc.buf);
    double2 *dq = reinterpret_cast<double2 *>(static_cast<char *>(sc.buf) + pb);
    int *qid = reinterpret_cast<int *>(static_cast<char *>(sc.buf) + pb + qb);
    double2 *dout = reinterpret_cast<double2 *>(static_cast<char *>(sc.buf) + pb + qb + 4 * (size_t)nq);
    BL_HIP(hipMemcpyAsync(dpos, xy, pb, hipMemcpyHostToDevice, sc.s));
    if (nq > 0) BL_HIP(hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, sc.s));
    int rc = build_tree(sc.w, dpos, n, sc.s, nullptr, nullptr);
    if (rc) return rc;
    if (nq == 0) return BL_OK;
    std::vector<int> ids(nq);
    for (int k = 0; k < nq; ++k) ids[k] = k;
    BL_HIP(hipMemcpyAsync(qid, ids.data(), 4 * (size_t)nq, hipMemcpyHostToDevice, sc.s));
    const int waves = (nq + 63) / 64;
    const int gq = (waves + WG / 64 - 1) / (WG / 64);
    BHWork &w = sc.w;
    if (precision == BL_PREC_FP64)
        hipLaunchKernelGGL(k_rep64, dim3(gq), dim3(WG), 0, sc.s, dq, qid, nq, w.spos.as<double2>(), w.node_child.as<BHNode>(),
                           w.misc.as<int>(), w.frame.as<double>(), theta, dout, w.misc.as<int>());
    else
        hipLaunchKernelGGL(precision == BL_PREC_FP32 ? k_rep32<false> : k_rep32<true>, dim3(gq), dim3(WG), 0, sc.s, dq,
                           qid, nq, w.spos.as<double2>(), w.node_child.as<BHNode>(), w.misc.as<int>(),
                           w.frame.as<double>(), theta, dout, w.misc.as<int>());
    BL_HIP(hipGetLastError());
    int status = 0;
    BL_HIP(hipMemcpyAsync(&status, w.misc.p, sizeof(int), hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipMemcpyAsync(out, dout, 16 * (size_t)nq, hipMemcpyDeviceToHost, sc.s));
    BL_HIP(hipStreamSynchronize(sc.s));
    if (status & ST_DUPKEY) return BL_E_DUPKEY;
    return BL_OK;
}

}  // namespace bl
