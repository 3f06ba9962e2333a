// eq_cells.hip -- cell <-> mesh coupling of eQ's updateCells on gfx950.
//
// Reference semantics (src/abm/eQabm.cpp:234-449): for every cell in list
// order, find the covered nodes S_c (findInteriorPoints, 268-305), read the
// mean of every species field over S_c (readHSL, 326-337), then add
// q_c/|S_c| to every covered node of every species (writeHSL, 338-359).  A
// cell therefore samples the deposits of all earlier cells.
//
// GPU plan (deterministic: every node sees its cells in list order, every
// sum is formed in a fixed order), batched over a range of traps:
//   1. k_cell_geo     one wavefront per cell: pole window, |S_c| (ballot
//                     counts), and the 64x16-node tiles the window touches
//   2. k_pair_block_sums, k_pair_scan, k_pair_offsets
//                     exclusive scan of the per-cell tile counts
//   3. k_pairs        one (tile, cell) pair per touched tile, cell-major
//   4. radix sort     pairs by global tile id, stable, so every tile's cells
//                     come out in list order
//   5. k_tile_bounds  first/last pair of every tile
//   6. k_tile_couple  one block per tile, four rows of one column per
//                     thread, all species in registers: for each of the
//                     tile's cells in order, the covered nodes' sum (one
//                     wave sum per wavefront, summed over the wavefronts
//                     once per chunk of cells: one partial per pair) and
//                     then the deposit
//   7. k_cell_sample  per cell: its pair partials in tile order / |S_c|
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "eq_internal.h"

namespace eqk {

constexpr int kCTileW = 64;                 // tile columns (one wavefront)
constexpr int kCRows = 4;                   // rows per thread
constexpr int kCTileH = 4 * kCRows;         // tile rows (4 wavefronts)
constexpr int kCThreads = 256;
constexpr int kSpeciesPass = 4;             // species held in registers per pass

struct CellGeo {
    double cx, cy, cs, sn, o, rad; // centre, cos/sin of the angle, box half-length/half-width
    int32_t i1, i2, j1, j2;        // search window (inclusive); the centre node for fallback cells
    int32_t cnt;                   // nodes found by the scan (0: centre-node fallback)
    int32_t trap;                  // trap index (in the batch)
    int32_t tx1, ty1, ntx;         // first tile column/row of the window, tile columns
    int32_t npairs;                // tiles the window touches
    int64_t poff;                  // first pair of this cell
};

// per trap of the batch (device table)
struct TrapGrid {
    int32_t nodes_w, nodes_h;
    int32_t tiles_x, tiles_y;
    int32_t tile_base;             // first global tile id
    int32_t prob0;                 // problem of species 0
    int32_t pitch;
    int32_t tens;                  // 1: the coupling sets this trap's diffusion tensor from the cells
    double trap_w, trap_h;
    double Dx, Dy;                 // axial and transverse diffusion factors (tens = 1)
};

// size_t(round(v)) of eQ::data::ij_from_xy (src/eQ.h:119-124); negative values clamp to 0
__device__ __forceinline__ int32_t to_index(double v, double npm)
{
    const double r = round(v * npm);
    return r < 0.0 ? 0 : (int32_t)r;
}

#pragma clang fp contract(off)
// cpmEcoli::pointIsInCell (src/abm/cpmEcoli.cpp:313-328): world->local with
// the inverse rigid transform of body A (Chipmunk cpBodyWorldToLocal,
// centre of gravity 0), then the four strict edge tests of the body-A box
// (-o,r),(o,r),(o,-r),(-o,-r).  With the box's axis-aligned edges the four
// cpvcross(edge, p - vertex) < 0 tests are exactly -o < lx < o and
// -r < ly < r (products with a zero edge component are +-0, and the sign of
// a rounded difference is the sign of the exact difference), that is
// |lx| < o and |ly| < r: one compare each (false for NaN, and for o <= 0 or
// r <= 0 as the two-sided tests are).
__device__ __forceinline__ double inv_tx(double cx, double cy, double cs, double sn) { return (-sn) * cy - cs * cx; }
__device__ __forceinline__ double inv_ty(double cx, double cy, double cs, double sn) { return sn * cx - cs * cy; }

// the test with the products of the node coordinates formed by the caller:
// csX = cs*X, snX = (-sn)*X, snY = sn*Y, csY = cs*Y
__device__ __forceinline__ bool inside_box(double csX, double snX, double snY, double csY, double itx, double ity,
                                           double o, double rad)
{
    const double lx = csX + snY + itx;
    const double ly = snX + csY + ity;
    return (fabs(lx) < o) && (fabs(ly) < rad);
}

__device__ __forceinline__ bool inside_cell(double cx, double cy, double cs, double sn, double o, double rad,
                                            double X, double Y)
{
    return inside_box(cs * X, (-sn) * X, sn * Y, cs * Y, inv_tx(cx, cy, cs, sn), inv_ty(cx, cy, cs, sn), o, rad);
}

// Ecoli::updatePoleCenters (src/abm/Ecoli.cpp:36-63)
__device__ __forceinline__ void pole(const TrapGrid& G, double cx, double cy, double cs, double sn, double L,
                                     double w, int r, double& px, double& py)
{
    const double s = ((double)r * 0.5) * L - w / 2.0;
    px = cx + cs * s;
    py = cy + sn * s;
    if (px < 0.0) px = 0.0;
    if (py < 0.0) py = 0.0;
    if (px > G.trap_w) px = G.trap_w;
    if (py > G.trap_h) py = G.trap_h;
}
#pragma clang fp contract(on)

// row_shr:n of both halves of a double, zero where the source lane is
// outside the 16-lane row
template <int N>
__device__ __forceinline__ double dpp_shr(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x110 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x110 + N, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int k)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}

// trap holding cell c: offsets[t] <= c < offsets[t+1] (offsets of the traps
// of this call, ntr + 1 entries)
__device__ __forceinline__ int trap_of(const int64_t* __restrict__ off, int ntr, int64_t c)
{
    int lo = 0, hi = ntr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_cell_geo(const eq_cell* __restrict__ cells, int64_t nc,
                                                  const int64_t* __restrict__ off, int ntr, int t0,
                                                  const TrapGrid* __restrict__ traps, double npm,
                                                  CellGeo* __restrict__ geo)
{
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nc) return;
    const int t = t0 + trap_of(off, ntr, c);
    const TrapGrid G = traps[t];
    const eq_cell cell = cells[c];
    CellGeo g;
    g.cx = cell.x;
    g.cy = cell.y;
    g.cs = cos(cell.angle);
    g.sn = sin(cell.angle);
    // box of body A for the cell's length (src/abm/cpmEcoli.cpp:102-124)
    g.o = (cell.length - cell.width) * 0.5;
    g.rad = cell.width * 0.5;
    // search window (src/abm/eQabm.cpp:268-287); nodesToEdge from the default 1 um width (eQabm.cpp:75)
    const int32_t nte = (int32_t)round(npm * 1.0 / 2.0);
    double ax, ay, bx, by;
    pole(G, g.cx, g.cy, g.cs, g.sn, cell.length, cell.width, 1, ax, ay);
    pole(G, g.cx, g.cy, g.cs, g.sn, cell.length, cell.width, -1, bx, by);
    const int32_t Ai = to_index(ay, npm), Aj = to_index(ax, npm);
    const int32_t Bi = to_index(by, npm), Bj = to_index(bx, npm);
    int32_t i1 = min(Ai, Bi), i2 = max(Ai, Bi), j1 = min(Aj, Bj), j2 = max(Aj, Bj);
    i1 = (i1 >= nte) ? i1 - nte : 0;
    j1 = (j1 >= nte) ? j1 - nte : 0;
    i2 = (i2 + nte >= G.nodes_h - 1) ? G.nodes_h - 1 : i2 + nte;
    j2 = (j2 + nte >= G.nodes_w - 1) ? G.nodes_w - 1 : j2 + nte;
    // Only nodes of the window inside the box's axis-aligned bounding box
    // (one node of margin; the rounded transform is off by ~1e-15) can pass
    // the strict inside test, so the scan, and everything after it, uses
    // the window clipped to that box: the covered set is unchanged.
    {
        const double ex = fabs(g.o * g.cs) + fabs(g.rad * g.sn), ey = fabs(g.o * g.sn) + fabs(g.rad * g.cs);
        const double bj1 = floor((g.cx - ex) * npm) - 1.0, bj2 = ceil((g.cx + ex) * npm) + 1.0;
        const double bi1 = floor((g.cy - ey) * npm) - 1.0, bi2 = ceil((g.cy + ey) * npm) + 1.0;
        if (bj1 > (double)j1) j1 = bj1 > (double)j2 ? j2 + 1 : (int32_t)bj1;
        if (bj2 < (double)j2) j2 = bj2 < (double)j1 ? j1 - 1 : (int32_t)bj2;
        if (bi1 > (double)i1) i1 = bi1 > (double)i2 ? i2 + 1 : (int32_t)bi1;
        if (bi2 < (double)i2) i2 = bi2 < (double)i1 ? i1 - 1 : (int32_t)bi2;
    }
    int32_t cnt = 0;
    {
        const double itx = inv_tx(g.cx, g.cy, g.cs, g.sn), ity = inv_ty(g.cx, g.cy, g.cs, g.sn);
        const double msn = -g.sn;
        for (int32_t jb = j1; jb <= j2; jb += 64) {
            const int32_t j = jb + lane;
            const double X = (double)j / npm;
            const double csX = g.cs * X, snX = msn * X;
            for (int32_t ib = i1; ib <= i2; ib += 64) {
                // lane k holds Y of row ib + k; rows are taken one by one
                const double Yl = (double)(ib + lane) / npm;
                const int nr = min(64, i2 - ib + 1);
                for (int k = 0; k < nr; ++k) {
                    const double Y = readlane_d(Yl, k);
                    const bool in = (j <= j2) && inside_box(csX, snX, g.sn * Y, g.cs * Y, itx, ity, g.o, g.rad);
                    cnt += __popcll(__ballot(in));
                }
            }
        }
    }
    g.cnt = cnt;
    if (cnt == 0) {
        // fallback: the centre node (src/abm/eQabm.cpp:299-303)
        i1 = i2 = min(to_index(g.cy, npm), G.nodes_h - 1);
        j1 = j2 = min(to_index(g.cx, npm), G.nodes_w - 1);
    }
    g.i1 = i1; g.i2 = i2; g.j1 = j1; g.j2 = j2;
    g.trap = t;
    g.tx1 = j1 / kCTileW;
    g.ty1 = i1 / kCTileH;
    g.ntx = j2 / kCTileW - g.tx1 + 1;
    g.npairs = (i2 / kCTileH - g.ty1 + 1) * g.ntx;
    g.poff = 0;
    if (lane == 0) geo[c] = g;
}

// Exclusive scan of the per-cell pair counts into geo[c].poff, in three
// launches (integer sums: exact, so the order does not matter):
//   k_pair_block_sums  one block per kScanBlock cells: the block's total
//   k_pair_scan        one block: exclusive scan of those totals, and the sum
//   k_pair_offsets     one block per kScanBlock cells: the block's own scan
//                      plus its offset
constexpr int kScanBlock = 256;

__device__ __forceinline__ int64_t block_inclusive_scan(int64_t v, int64_t* sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const int64_t u = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    return sh[t];
}

__global__ __launch_bounds__(kScanBlock) void k_pair_block_sums(const CellGeo* __restrict__ geo, int64_t nc,
                                                                int64_t* __restrict__ bsum)
{
    __shared__ int64_t sh[kScanBlock];
    const int64_t c = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int64_t v = block_inclusive_scan(c < nc ? (int64_t)geo[c].npairs : 0, sh);
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = v;
}

__global__ __launch_bounds__(1024) void k_pair_scan(int64_t* bsum, int64_t nb, int64_t* total)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, nt = blockDim.x;
    const int64_t per = (nb + nt - 1) / nt;
    const int64_t lo = min(nb, (int64_t)t * per), hi = min(nb, lo + per);
    int64_t s = 0;
    for (int64_t k = lo; k < hi; ++k) s += bsum[k];
    part[t] = s;
    __syncthreads();
    // inclusive Hillis-Steele scan of the thread sums
    for (int off = 1; off < nt; off <<= 1) {
        const int64_t v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    if (t == nt - 1) *total = part[t];
    int64_t acc = part[t] - s;
    for (int64_t k = lo; k < hi; ++k) {
        const int64_t v = bsum[k];
        bsum[k] = acc;
        acc += v;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_pair_offsets(CellGeo* __restrict__ geo, int64_t nc,
                                                             const int64_t* __restrict__ boff)
{
    __shared__ int64_t sh[kScanBlock];
    const int64_t c = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int64_t v = c < nc ? (int64_t)geo[c].npairs : 0;
    const int64_t incl = block_inclusive_scan(v, sh);
    if (c < nc) geo[c].poff = boff[blockIdx.x] + (incl - v);
}

// pair p of cell c: key = global tile id, value = p; the cell's tiles row by row
// (pairs at or past cap are not written: the call's pair buffers hold cap)
__global__ __launch_bounds__(256) void k_pairs(const CellGeo* __restrict__ geo, int64_t nc,
                                               const TrapGrid* __restrict__ traps, uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ vals, uint32_t* __restrict__ pair_cell,
                                               int64_t cap)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const CellGeo& g = geo[c];
    const TrapGrid& G = traps[g.trap];
    for (int k = 0; k < g.npairs; ++k) {
        const int ty = g.ty1 + k / g.ntx, tx = g.tx1 + k % g.ntx;
        const int64_t p = g.poff + k;
        if (p >= cap) break;
        keys[p] = (uint32_t)(G.tile_base + ty * G.tiles_x + tx);
        vals[p] = (uint32_t)p;
        pair_cell[p] = (uint32_t)c;
    }
}

// A call sized before its pair count is known sorts cap slots: the slots past
// the count take a key above every tile id (the sort is over the low bits of
// the keys, and every tile id is below 2^bits - 1), so they sort last; the
// count is also written to host memory (mapped) for the next call's sizing.
__global__ __launch_bounds__(256) void k_pad_pairs(uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   const int64_t* __restrict__ total, int64_t cap,
                                                   int64_t* __restrict__ host_total)
{
    const int64_t n = *total;
    if (blockIdx.x == 0 && threadIdx.x == 0 && host_total) *host_total = n;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + n; k < cap; k += (int64_t)gridDim.x * blockDim.x) {
        keys[k] = 0xffffffffu;
        vals[k] = (uint32_t)k;
    }
}

// [start, end) of every tile's run in the sorted pairs (tiles without pairs
// keep the zeros they were cleared to); total (optional): the pairs are the
// first min(*total, n) of the n sorted slots
__global__ __launch_bounds__(256) void k_tile_bounds(const uint32_t* __restrict__ keys, int64_t n, uint32_t tile0,
                                                     uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                     const int64_t* __restrict__ total)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (total) n = min(n, *total);
    if (k >= n) return;
    const uint32_t key = keys[k];
    if (k == 0 || keys[k - 1] != key) start[key - tile0] = (uint32_t)k;
    if (k == n - 1 || keys[k + 1] != key) end[key - tile0] = (uint32_t)(k + 1);
}

// sum over the wavefront in a fixed order: prefix sums within each 16-lane
// row by DPP shifts, then the four row totals in lane order (no LDS round
// trips, unlike a shuffle tree)
__device__ __forceinline__ double wave_sum64(double v)
{
    v += dpp_shr<1>(v);
    v += dpp_shr<2>(v);
    v += dpp_shr<4>(v);
    v += dpp_shr<8>(v);
    return ((readlane_d(v, 15) + readlane_d(v, 31)) + readlane_d(v, 47)) + readlane_d(v, 63);
}

// one block per tile of the call's trap range
constexpr int kChunk = 64; // cells staged in LDS at a time

// P species held in registers per pass (2 covers one or two species, 4 more);
// TENS: a trap of the call sets its diffusion tensor from the cells
template <int P, bool TENS>
__global__ __launch_bounds__(kCThreads) void k_tile_couple(
    const TrapGrid* __restrict__ traps, int t0, int ntr, uint32_t tile0, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ end, const uint32_t* __restrict__ sorted_pairs, const uint32_t* __restrict__ pair_cell,
    const CellGeo* __restrict__ geo, double* const* __restrict__ fields, int ns, const double* __restrict__ q,
    double* __restrict__ partial, double npm, unsigned flags, double* const* __restrict__ tens64,
    float* const* __restrict__ tens32, int64_t tstride, const int64_t* __restrict__ total, int64_t cap)
{
    __shared__ double red[kChunk][kCThreads / 64][P]; // per cell of the chunk: each wave's sums
    __shared__ CellGeo sg[kChunk];      // cx, cy replaced by the inverse translation
    __shared__ uint32_t sp[kChunk];
    __shared__ double sq[kChunk][P]; // q_c / |S_c|
    // more pairs than the call's buffers hold: k_couple_affine does the call
    if (total && *total > cap) return;
    const uint32_t tl = blockIdx.x;
    const uint32_t e0 = start[tl], e1 = end[tl];
    const bool want_tens = TENS && (flags & EQ_COUPLE_TENSOR) != 0;
    if (e0 >= e1 && !want_tens) return;
    // the tile's trap: traps are numbered in tile order
    int lo = t0, hi = t0 + ntr - 1;
    const uint32_t tg = tile0 + tl;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((uint32_t)traps[mid].tile_base <= tg) lo = mid; else hi = mid - 1;
    }
    const TrapGrid G = traps[lo];
    const int lt = (int)(tg - (uint32_t)G.tile_base);
    const int ty = lt / G.tiles_x, tx = lt - ty * G.tiles_x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = tx * kCTileW + lane;
    const int ib = ty * kCTileH + w * kCRows;
    const bool dep = (flags & EQ_COUPLE_DEPOSIT) != 0;
    const double X = (double)j / npm;
    double Y[kCRows];
#pragma unroll
    for (int r = 0; r < kCRows; ++r) Y[r] = (double)(ib + r) / npm;
    // setDiffusionTensor (src/abm/eQabm.cpp:241-248, 306-325): the nodes start
    // isotropic and each covering cell, in list order, overwrites them
    const bool tens = want_tens && G.tens;
    double t11[kCRows], t22[kCRows], t12[kCRows];
#pragma unroll
    for (int r = 0; r < kCRows; ++r) {
        t11[r] = 1.0;
        t22[r] = 1.0;
        t12[r] = 0.0;
    }
    for (int s0 = 0; s0 < ns && e0 < e1; s0 += P) {
        const int nsp = min(P, ns - s0);
        double u[P][kCRows];
        bool in[kCRows];
#pragma unroll
        for (int r = 0; r < kCRows; ++r) {
            in[r] = (ib + r < G.nodes_h) && (j < G.nodes_w);
#pragma unroll
            for (int s = 0; s < P; ++s)
                u[s][r] = (s < nsp && in[r]) ? fields[G.prob0 + s0 + s][(int64_t)(ib + r) * G.pitch + j] : 0.0;
        }
        for (uint32_t cb = e0; cb < e1; cb += kChunk) {
            const int n = (int)min((uint32_t)kChunk, e1 - cb);
            __syncthreads(); // the previous chunk's entries and wave sums are no longer read
            if ((int)threadIdx.x < n) {
                const uint32_t p = sorted_pairs[cb + threadIdx.x];
                const uint32_t c = pair_cell[p];
                CellGeo g = geo[c];
                const double itx = inv_tx(g.cx, g.cy, g.cs, g.sn), ity = inv_ty(g.cx, g.cy, g.cs, g.sn);
                g.cx = itx;
                g.cy = ity;
                sg[threadIdx.x] = g;
                sp[threadIdx.x] = p;
                const double cntd = (double)(g.cnt > 0 ? g.cnt : 1);
#pragma unroll
                for (int s = 0; s < P; ++s)
                    sq[threadIdx.x][s] = (s < nsp && dep) ? q[(int64_t)c * ns + s0 + s] / cntd : 0.0;
            }
            __syncthreads();
            for (int k = 0; k < n; ++k) {
                const CellGeo& g = sg[k];
                const double cs = g.cs, sn = g.sn, itx = g.cx, ity = g.cy, o = g.o, rad = g.rad;
                const bool fallback = g.cnt == 0;
                const bool colin = j >= g.j1 && j <= g.j2;
                // a wavefront whose four rows miss the cell's window adds nothing
                // (uniform over the wavefront)
                if (__ballot(colin && ib <= g.i2 && ib + kCRows - 1 >= g.i1) == 0ull) {
                    if (lane == 0)
#pragma unroll
                        for (int s = 0; s < P; ++s) red[k][w][s] = 0.0;
                    continue;
                }
                const double csX = cs * X, snX = (-sn) * X;
                bool cov[kCRows];
                bool anyc = false;
#pragma unroll
                for (int r = 0; r < kCRows; ++r) {
                    const int i = ib + r;
                    bool v = in[r] && colin && i >= g.i1 && i <= g.i2;
                    if (v && !fallback) v = inside_box(csX, snX, sn * Y[r], cs * Y[r], itx, ity, o, rad);
                    cov[r] = v;
                    anyc = anyc || v;
                }
                if (tens && s0 == 0 && anyc) {
                    // cos^2, sin^2 and sin cos as the reference forms them, each product rounded
                    const double c2 = __dmul_rn(cs, cs), s2 = __dmul_rn(sn, sn), sc = __dmul_rn(sn, cs);
                    const double d11 = __dadd_rn(__dmul_rn(G.Dx, c2), __dmul_rn(G.Dy, s2));
                    const double d22 = __dadd_rn(__dmul_rn(G.Dx, s2), __dmul_rn(G.Dy, c2));
                    const double d12 = __dmul_rn(__dsub_rn(G.Dx, G.Dy), sc);
#pragma unroll
                    for (int r = 0; r < kCRows; ++r)
                        if (cov[r]) {
                            t11[r] = d11;
                            t22[r] = d22;
                            t12[r] = d12;
                        }
                }
                const bool wave_any = __ballot(anyc) != 0ull; // uniform over the wavefront
#pragma unroll
                for (int s = 0; s < P; ++s) {
                    if (s >= nsp) break;
                    double sum = 0.0;
                    const double d = sq[k][s];
#pragma unroll
                    for (int r = 0; r < kCRows; ++r) {
                        if (cov[r]) {
                            sum += u[s][r];
                            if (dep) u[s][r] += d;
                        }
                    }
                    if (wave_any) sum = wave_sum64(sum);
                    if (lane == 0) red[k][w][s] = wave_any ? sum : 0.0;
                }
            }
            // the waves walk the chunk's cells independently (each owns its
            // own nodes); the per-pair partial sums, waves in order, follow
            // once every wave is through the chunk
            __syncthreads();
            for (int t = threadIdx.x; t < n * nsp; t += kCThreads) {
                const int k = t / nsp, s = t - k * nsp;
                double acc = 0.0;
                for (int q4 = 0; q4 < kCThreads / 64; ++q4) acc += red[k][q4][s];
                partial[(int64_t)sp[k] * ns + s0 + s] = acc;
            }
        }
        if (dep) {
#pragma unroll
            for (int r = 0; r < kCRows; ++r)
#pragma unroll
                for (int s = 0; s < P; ++s)
                    if (s < nsp && in[r]) fields[G.prob0 + s0 + s][(int64_t)(ib + r) * G.pitch + j] = u[s][r];
        }
    }
    if (!tens) return;
    // every species of the trap carries the same tensor, in FP64 (the Krylov
    // operator) and, for an FP32 V-cycle, in FP32
    for (int s = 0; s < ns; ++s) {
        double* d = tens64[G.prob0 + s];
        float* f = tens32 ? tens32[G.prob0 + s] : nullptr;
#pragma unroll
        for (int r = 0; r < kCRows; ++r) {
            if (!((ib + r < G.nodes_h) && (j < G.nodes_w))) continue;
            const int64_t k = (int64_t)(ib + r) * G.pitch + j;
            d[k] = t11[r];
            d[k + tstride] = t22[r];
            d[k + 2 * tstride] = t12[r];
            if (f) {
                f[k] = (float)t11[r];
                f[k + tstride] = (float)t22[r];
                f[k + 2 * tstride] = (float)t12[r];
            }
        }
    }
}

// sampled[c,s] = (sum of the cell's pair partials in tile order) / |S_c|
// (more pairs than the call's buffers hold: k_couple_affine does the call)
__global__ __launch_bounds__(256) void k_cell_sample(const CellGeo* __restrict__ geo, int64_t nc, int ns,
                                                     const double* __restrict__ partial, double* __restrict__ sampled,
                                                     int32_t* __restrict__ counts, const int64_t* __restrict__ total,
                                                     int64_t cap)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc || (total && *total > cap)) return;
    const CellGeo& g = geo[c];
    const int cnt = g.cnt > 0 ? g.cnt : 1;
    if (counts) counts[c] = cnt;
    if (!sampled) return;
    for (int s = 0; s < ns; ++s) {
        double acc = 0.0;
        for (int k = 0; k < g.npairs; ++k) acc += partial[(g.poff + k) * ns + s];
        sampled[c * ns + s] = acc / (double)cnt;
    }
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// The coupling of a cell whose amount depends on its own sample, q = a + b e
// (a strain's membrane exchange, src/Strain.h:312-313): per trap, one
// workgroup walks the cells in list order; for each cell every thread scans
// its share of the cell's window, the workgroup sums the covered nodes of
// every species (read before the cell's own write, after every earlier
// cell's), forms e = sum/|S_c| and q, and the covered nodes take q/|S_c|
// (src/abm/eQabm.cpp:326-359, 384-405).  A barrier separates each cell's
// reads from its writes and its writes from the next cell's reads.
// With qb null the amount is qa itself (qa null: nothing deposited): the
// tile path's result, by the same node tests and the same per-node order of
// additions.  total (optional): run only when *total > cap, as the tile
// path's stand-in for a call whose pairs outgrew its buffers.
constexpr int kAThreads = 1024;
__global__ __launch_bounds__(kAThreads) void k_couple_affine(const TrapGrid* __restrict__ traps, int t0,
                                                             const int64_t* __restrict__ off,
                                                             const CellGeo* __restrict__ geo,
                                                             double* const* __restrict__ fields, int ns,
                                                             const double* __restrict__ qa, const double* __restrict__ qb,
                                                             double* __restrict__ sampled, int32_t* __restrict__ counts,
                                                             double npm, const int64_t* __restrict__ total, int64_t cap)
{
    __shared__ double red[kAThreads / 64][kSpeciesPass];
    __shared__ double dep[kSpeciesPass];
    if (total && *total <= cap) return;
    const int t = blockIdx.x;
    const TrapGrid G = traps[t0 + t];
    const int64_t c0 = off[t], c1 = off[t + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t c = c0; c < c1; ++c) {
        const CellGeo g = geo[c];
        const int cnt = g.cnt > 0 ? g.cnt : 1;
        const int wj = g.j2 - g.j1 + 1, wi = g.i2 - g.i1 + 1;
        const int nwin = (wi > 0 && wj > 0) ? wi * wj : 0;
        if (counts && threadIdx.x == 0) counts[c] = cnt;
        for (int s0 = 0; s0 < ns; s0 += kSpeciesPass) {
            const int nsp = min(kSpeciesPass, ns - s0);
            double sum[kSpeciesPass];
#pragma unroll
            for (int s = 0; s < kSpeciesPass; ++s) sum[s] = 0.0;
            for (int k = threadIdx.x; k < nwin; k += kAThreads) {
                const int i = g.i1 + k / wj, j = g.j1 + k % wj;
                const bool in = g.cnt == 0 ||
                                inside_cell(g.cx, g.cy, g.cs, g.sn, g.o, g.rad, (double)j / npm, (double)i / npm);
                if (!in) continue;
                const int64_t q = (int64_t)i * G.pitch + j;
#pragma unroll
                for (int s = 0; s < kSpeciesPass; ++s)
                    if (s < nsp) sum[s] += fields[G.prob0 + s0 + s][q];
            }
#pragma unroll
            for (int s = 0; s < kSpeciesPass; ++s) {
                const double v = wave_sum64(sum[s]);
                if (lane == 0) red[w][s] = v;
            }
            __syncthreads();
            if (threadIdx.x < (unsigned)nsp) {
                const int s = threadIdx.x;
                double tot = 0.0;
                for (int q4 = 0; q4 < kAThreads / 64; ++q4) tot += red[q4][s];
                const double e = tot / (double)cnt;
                if (sampled) sampled[c * ns + s0 + s] = e;
                if (qa) {
                    const double amount = qb ? qa[c * ns + s0 + s] + qb[c * ns + s0 + s] * e : qa[c * ns + s0 + s];
                    dep[s] = amount / (double)cnt;
                }
            }
            __syncthreads();
            if (!qa) continue; // nothing deposited (no barrier skipped: qa is uniform)
            for (int k = threadIdx.x; k < nwin; k += kAThreads) {
                const int i = g.i1 + k / wj, j = g.j1 + k % wj;
                const bool in = g.cnt == 0 ||
                                inside_cell(g.cx, g.cy, g.cs, g.sn, g.o, g.rad, (double)j / npm, (double)i / npm);
                if (!in) continue;
                const int64_t q = (int64_t)i * G.pitch + j;
#pragma unroll
                for (int s = 0; s < kSpeciesPass; ++s)
                    if (s < nsp) fields[G.prob0 + s0 + s][q] += dep[s];
            }
            __syncthreads(); // this cell's writes before the next cell's reads
        }
    }
}

// zero n 32-bit words (tile bounds); a kernel rather than a memset, so the
// same sequence can sit in a captured graph
__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}

// The allocation-free coupling's scratch: one fixed buffer per device, held
// from the first use of the library's code on that device until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host it
// names a stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once with that device current.
constexpr size_t kFixedScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// a call's scratch: the batch's growable buffer, or the fixed one
struct CoupleScratch {
    char* base = nullptr;
    size_t cap = 0;
};

static int ensure_scratch(eq_batch* b, size_t bytes)
{
    if (b->cell_scratch_bytes >= bytes) return EQ_OK;
    if (b->cell_scratch) {
        hipStreamSynchronize(b->stream);
        hipFree(b->cell_scratch);
        b->cell_scratch = nullptr;
        b->cell_scratch_bytes = 0;
    }
    const size_t want = bytes + bytes / 4 + 4096;
    if (hipMalloc(&b->cell_scratch, want) != hipSuccess) return set_error(EQ_ERR_NOMEM, "cell scratch allocation");
    b->cell_scratch_bytes = want;
    return EQ_OK;
}

// trap grids (first use): the tile tables of every trap, in trap order
static int ensure_traps(eq_batch* b)
{
    if (b->d_cell_traps) return EQ_OK;
    const int ns = b->nspecies;
    std::vector<TrapGrid> tg(b->ntraps);
    int64_t tiles = 0;
    for (int t = 0; t < b->ntraps; ++t) {
        const LevelProb& P = b->levels[0].probs[t * ns];
        TrapGrid& G = tg[t];
        G.nodes_w = P.nx + 1;
        G.nodes_h = P.ny + 1;
        G.tiles_x = (G.nodes_w + kCTileW - 1) / kCTileW;
        G.tiles_y = (G.nodes_h + kCTileH - 1) / kCTileH;
        G.tile_base = (int32_t)tiles;
        G.prob0 = t * ns;
        G.pitch = P.pitch;
        G.trap_w = b->descs[t * ns].width;
        G.trap_h = b->descs[t * ns].height;
        G.tens = 0;
        G.Dx = G.Dy = 1.0;
        tiles += (int64_t)G.tiles_x * G.tiles_y;
    }
    if (tiles >= ((int64_t)1 << 31)) return set_error(EQ_ERR_INVALID, "too many coupling tiles");
    b->cell_tile_base.resize(b->ntraps + 1);
    for (int t = 0; t < b->ntraps; ++t) b->cell_tile_base[t] = tg[t].tile_base;
    b->cell_tile_base[b->ntraps] = (int32_t)tiles;
    if (hipMalloc((void**)&b->d_cell_traps, sizeof(TrapGrid) * b->ntraps) != hipSuccess ||
        hipHostMalloc((void**)&b->h_cell_off, sizeof(int64_t) * (b->ntraps + 1), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&b->d_cell_off, sizeof(int64_t) * (b->ntraps + 1)) != hipSuccess)
        return set_error(EQ_ERR_NOMEM, "coupling tables");
    EQ_HIP(hipMemcpy(b->d_cell_traps, tg.data(), sizeof(TrapGrid) * b->ntraps, hipMemcpyHostToDevice));
    return EQ_OK;
}

// a call's cell offsets (relative to its first cell) to the device, through
// the pinned staging once the previous upload from it has run
static int upload_cell_off(eq_batch* b, const int64_t* cell_off, int ntr)
{
    if (b->cell_off_ev) EQ_HIP(hipEventSynchronize(b->cell_off_ev));
    else EQ_HIP(hipEventCreateWithFlags(&b->cell_off_ev, hipEventDisableTiming));
    for (int k = 0; k <= ntr; ++k) b->h_cell_off[k] = cell_off[k] - cell_off[0];
    EQ_HIP(hipMemcpyAsync(b->d_cell_off, b->h_cell_off, sizeof(int64_t) * (ntr + 1), hipMemcpyHostToDevice, b->stream));
    EQ_HIP(hipEventRecord(b->cell_off_ev, b->stream));
    return EQ_OK;
}

// cells of traps t0 .. t0+ntr-1, concatenated; cell_off[k] (host) = first cell
// of trap t0+k, cell_off[ntr] = total
static int couple_impl(eq_batch* b, int32_t t0, int32_t ntr, const int64_t* cell_off, const eq_cell* d_cells,
                       double npm, const double* d_q, double* d_sampled, int32_t* d_counts, uint32_t flags,
                       const CoupleScratch* fixed = nullptr)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (ntr < 1 || t0 < 0 || t0 + ntr > b->ntraps) return set_error(EQ_ERR_INVALID, "trap index out of range");
    if (!(npm > 0)) return set_error(EQ_ERR_INVALID, "npm must be > 0");
    for (int k = 0; k < ntr; ++k)
        if (cell_off[k + 1] < cell_off[k]) return set_error(EQ_ERR_INVALID, "cell offsets must not decrease");
    const int64_t nc = cell_off[ntr] - cell_off[0];
    if (nc == 0) return EQ_OK;
    if (nc > ((int64_t)1 << 30)) return set_error(EQ_ERR_INVALID, "too many cells");
    if (!d_cells || (!d_q && (flags & EQ_COUPLE_DEPOSIT))) return set_error(EQ_ERR_INVALID, "NULL cells or q");
    int rc = ensure_traps(b);
    if (rc) return rc;
    const int ns = b->nspecies;
    hipStream_t s = b->stream;
    const uint32_t tile0 = (uint32_t)b->cell_tile_base[t0];
    const uint32_t ntiles = (uint32_t)(b->cell_tile_base[t0 + ntr] - b->cell_tile_base[t0]);
    // cell offsets relative to the first cell of the call
    if ((rc = upload_cell_off(b, cell_off, ntr))) return rc;
    // scratch: geo | total | tile start | tile end | (pairs: keys, vals, keys', vals', pair_cell, partial, sort temp)
    const size_t geo_bytes = align256(sizeof(CellGeo) * nc);
    const size_t tb_bytes = align256(sizeof(uint32_t) * ntiles);
    const int64_t nblk = (nc + kScanBlock - 1) / kScanBlock;
    const size_t blk_bytes = align256(sizeof(int64_t) * (size_t)nblk);
    const size_t head = geo_bytes + 256 + 2 * tb_bytes + blk_bytes;
    auto too_big = [&](size_t need) {
        char msg[160];
        std::snprintf(msg, sizeof(msg),
                      "the coupling needs %zu bytes of scratch and the fixed buffer holds %zu: couple fewer traps "
                      "per call", need, fixed->cap);
        return set_error(EQ_ERR_INVALID, msg);
    };
    if (fixed) {
        if (head + 256 > fixed->cap) return too_big(head + 256);
    } else if ((rc = ensure_scratch(b, head + 256))) {
        return rc;
    }
    char* base = fixed ? fixed->base : (char*)b->cell_scratch;
    CellGeo* geo = (CellGeo*)base;
    int64_t* d_total = (int64_t*)(base + geo_bytes);
    int64_t* d_blk = (int64_t*)(base + geo_bytes + 256 + 2 * tb_bytes);
    const TrapGrid* traps = (const TrapGrid*)b->d_cell_traps;
    const int kid = EQ_KERNEL_COUPLE_SETUP;
    const double gb = (double)sizeof(CellGeo) * (double)nc;
    auto geometry = [&]() {
        launch(b, kid, (double)sizeof(eq_cell) * nc + gb, k_cell_geo, dim3((unsigned)((nc + 3) / 4)), dim3(256), 0, s,
               d_cells, (int64_t)nc, (const int64_t*)b->d_cell_off, ntr, t0, traps, npm, geo);
        launch(b, kid, gb + 8.0 * nblk, k_pair_block_sums, dim3((unsigned)nblk), dim3(kScanBlock), 0, s,
               (const CellGeo*)geo, (int64_t)nc, d_blk);
        launch(b, kid, 16.0 * nblk, k_pair_scan, dim3(1), dim3(1024), 0, s, d_blk, (int64_t)nblk, d_total);
        launch(b, kid, 2.0 * gb + 8.0 * nblk, k_pair_offsets, dim3((unsigned)nblk), dim3(kScanBlock), 0, s, geo,
               (int64_t)nc, (const int64_t*)d_blk);
    };
    // the tensor is written only where a trap of the call has one set from its cells
    bool tens = false;
    if (flags & EQ_COUPLE_TENSOR)
        for (int t : b->cell_tensor_traps) tens = tens || (t >= t0 && t < t0 + ntr);
    // The pair buffers are sized before the call's pair count is known (no
    // host wait inside the call) from the counts of earlier calls: the last
    // one read back once its call has run, the buffers kept at 1.5x the
    // largest.  A call with more pairs than that is done by k_couple_affine
    // in list order (same deposits), and the next call gets room for it.  A
    // first call, a tensor call and the fixed scratch read the count first.
    if (b->pair_pending && hipEventQuery(b->pair_ev) == hipSuccess) {
        b->pair_pending = false;
        const int64_t last = *b->h_pair_total;
        if (last > b->pair_cap - b->pair_cap / 4) b->pair_cap = last + last / 2 + 4096;
    }
    const bool sized = !fixed && !tens && b->pair_cap > 0;
    geometry();
    int64_t total = 0;
    if (sized) {
        total = b->pair_cap;
    } else {
        EQ_HIP(hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        EQ_HIP(hipStreamSynchronize(s));
    }
    if (total >= ((int64_t)1 << 32)) return set_error(EQ_ERR_INVALID, "too many cell/tile pairs");
    if (!fixed && !sized) b->pair_cap = std::max(b->pair_cap, total + total / 2 + 4096);
    // slots the pair buffers hold (sized: the kept capacity; else the count)
    const size_t pb = align256(sizeof(uint32_t) * (size_t)total);
    size_t sort_bytes = 0;
    EQ_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)total, 0,
                                     32, s));
    const size_t need = head + 5 * pb + align256(sizeof(double) * (size_t)total * ns) + align256(sort_bytes);
    if (fixed) {
        if (need > fixed->cap) return too_big(need); // nothing written outside the scratch yet
    } else if (need > b->cell_scratch_bytes) {
        // grow (keeps nothing: recompute the geometry in the new buffer)
        if ((rc = ensure_scratch(b, need))) return rc;
        base = (char*)b->cell_scratch;
        geo = (CellGeo*)base;
        d_total = (int64_t*)(base + geo_bytes);
        d_blk = (int64_t*)(base + geo_bytes + 256 + 2 * tb_bytes);
        geometry();
    }
    uint32_t* t_start = (uint32_t*)(base + geo_bytes + 256);
    uint32_t* t_end = (uint32_t*)(base + geo_bytes + 256 + tb_bytes);
    char* pbase = base + head;
    uint32_t* keys = (uint32_t*)pbase;
    uint32_t* vals = (uint32_t*)(pbase + pb);
    uint32_t* keys2 = (uint32_t*)(pbase + 2 * pb);
    uint32_t* vals2 = (uint32_t*)(pbase + 3 * pb);
    uint32_t* pair_cell = (uint32_t*)(pbase + 4 * pb);
    double* partial = (double*)(pbase + 5 * pb);
    void* sort_tmp = pbase + 5 * pb + align256(sizeof(double) * (size_t)total * ns);
    launch(b, kid, gb + 12.0 * total, k_pairs, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, (const CellGeo*)geo,
           (int64_t)nc, traps, keys, vals, pair_cell, (int64_t)total);
    // sized: the unused slots sort last; the count goes to the host for later calls
    const int64_t* gate = sized ? (const int64_t*)d_total : nullptr;
    if (sized) {
        if (!b->h_pair_total &&
            hipHostMalloc((void**)&b->h_pair_total, sizeof(int64_t), hipHostMallocMapped) != hipSuccess)
            return set_error(EQ_ERR_NOMEM, "pair count staging");
        if (!b->pair_ev) EQ_HIP(hipEventCreateWithFlags(&b->pair_ev, hipEventDisableTiming));
        int64_t* host_total = nullptr;
        EQ_HIP(hipHostGetDevicePointer((void**)&host_total, b->h_pair_total, 0));
        launch(b, kid, 8.0 * total, k_pad_pairs, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 1024)), dim3(256),
               0, s, keys, vals, (const int64_t*)d_total, (int64_t)total, host_total);
    }
    // key bits: global tile ids up to the last tile of the batch
    unsigned bits = 1;
    while (bits < 32 && ((uint64_t)b->cell_tile_base[b->ntraps] >> bits) != 0) ++bits;
    EQ_HIP(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, (const uint32_t*)keys, keys2, (const uint32_t*)vals, vals2,
                                     (size_t)total, 0, bits, s));
    b->acct_to->bytes += 16.0 * total; // the sort's keys and values, read and written once
    b->acct_to->launches += 1;
    launch(b, kid, (double)(2 * tb_bytes), k_zero_u32, dim3((unsigned)((2 * tb_bytes / 4 + 255) / 256)), dim3(256), 0,
           s, t_start, (int64_t)(2 * tb_bytes / 4));
    launch(b, kid, 4.0 * total + 8.0 * ntiles, k_tile_bounds, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
           (const uint32_t*)keys2, (int64_t)total, tile0, t_start, t_end, gate);
    // algorithmic bytes: every node of the call's traps read (and written when
    // depositing) once per species, the pairs, the cell records and amounts,
    // and one partial per pair and species
    double nodes = 0.0;
    for (int t = t0; t < t0 + ntr; ++t) {
        const LevelProb& P = b->levels[0].probs[t * ns];
        nodes += (double)(P.nx + 1) * (P.ny + 1);
    }
    const double cb = nodes * ns * sizeof(double) * ((flags & EQ_COUPLE_DEPOSIT) ? 2 : 1) + 8.0 * total + gb +
                      8.0 * ns * nc + 8.0 * ns * total + 8.0 * ntiles;
    const unsigned kflags = tens ? flags : (flags & ~EQ_COUPLE_TENSOR);
    const double tb = tens ? nodes * ns * 3.0 * (sizeof(double) + (b->mixed ? sizeof(float) : 0)) : 0.0;
    auto kern = ns <= 2 ? (tens ? k_tile_couple<2, true> : k_tile_couple<2, false>)
                        : (tens ? k_tile_couple<kSpeciesPass, true> : k_tile_couple<kSpeciesPass, false>);
    launch(b, EQ_KERNEL_COUPLE, cb + tb, kern, dim3(ntiles), dim3(kCThreads), 0, s, traps, t0, ntr, tile0,
           (const uint32_t*)t_start, (const uint32_t*)t_end, (const uint32_t*)vals2, (const uint32_t*)pair_cell,
           (const CellGeo*)geo, (double* const*)b->d_field_ptrs, ns, d_q, partial, npm, kflags,
           (double* const*)b->d_tens64, (float* const*)b->d_tens32, (int64_t)b->levels[0].pool, gate, (int64_t)total);
    if (d_sampled || d_counts)
        launch(b, kid, gb + 8.0 * ns * total + 12.0 * ns * nc, k_cell_sample, dim3((unsigned)((nc + 255) / 256)),
               dim3(256), 0, s, (const CellGeo*)geo, (int64_t)nc, ns, (const double*)partial,
               (flags & EQ_COUPLE_SAMPLE) ? d_sampled : nullptr, d_counts, gate, (int64_t)total);
    if (sized) {
        // the call's stand-in when its pairs outgrew the buffers (exits at once otherwise)
        launch(b, EQ_KERNEL_COUPLE, 0.0, k_couple_affine, dim3((unsigned)ntr), dim3(kAThreads), 0, s, traps, (int)t0,
               (const int64_t*)b->d_cell_off, (const CellGeo*)geo, (double* const*)b->d_field_ptrs, ns,
               (flags & EQ_COUPLE_DEPOSIT) ? d_q : (const double*)nullptr, (const double*)nullptr,
               (flags & EQ_COUPLE_SAMPLE) ? d_sampled : nullptr, d_counts, npm, (const int64_t*)d_total, (int64_t)total);
        EQ_HIP(hipEventRecord(b->pair_ev, s));
        b->pair_pending = true;
    }
    EQ_HIP(hipGetLastError());
    if (tens) return refresh_cell_tensors(b);
    return EQ_OK;
}

// set-up of the coupling when the batch is created: the trap tables, and the
// fixed scratch's address on the batch's device (current here)
int init_coupling(eq_batch* b)
{
    int rc = ensure_traps(b);
    if (rc) return rc;
    if (b->device < 0 || b->device >= kMaxDevices)
        return set_error(EQ_ERR_INVALID, "device index beyond the fixed scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[b->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[b->device] = p;
    }
    return EQ_OK;
}

int set_trap_tensor(eq_batch* b, int trap, double Dx, double Dy, int on)
{
    int rc = ensure_traps(b);
    if (rc) return rc;
    TrapGrid G;
    TrapGrid* d = (TrapGrid*)b->d_cell_traps + trap;
    EQ_HIP(hipStreamSynchronize(b->stream));
    EQ_HIP(hipMemcpy(&G, d, sizeof(TrapGrid), hipMemcpyDeviceToHost));
    G.tens = on ? 1 : 0;
    G.Dx = on ? Dx : 1.0;
    G.Dy = on ? Dy : 1.0;
    EQ_HIP(hipMemcpy(d, &G, sizeof(TrapGrid), hipMemcpyHostToDevice));
    return EQ_OK;
}

} // namespace eqk

using namespace eqk;

extern "C" {

int eq_batch_couple_dev(eq_batch* b, int32_t trap, const eq_cell* d_cells, size_t nc, double npm, const double* d_q,
                        double* d_sampled, int32_t* d_counts, uint32_t flags)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    hipSetDevice(b->device);
    const int64_t off[2] = {0, (int64_t)nc};
    return couple_impl(b, trap, 1, off, d_cells, npm, d_q, d_sampled, d_counts, flags);
}

int eq_batch_couple_traps_dev(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                              const eq_cell* d_cells, double npm, const double* d_q, double* d_sampled,
                              int32_t* d_counts, uint32_t flags)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (!cell_offsets) return set_error(EQ_ERR_INVALID, "NULL cell offsets");
    hipSetDevice(b->device);
    return couple_impl(b, trap0, ntraps, cell_offsets, d_cells, npm, d_q, d_sampled, d_counts, flags);
}

int eq_batch_couple_affine_dev(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                               const eq_cell* d_cells, double npm, const double* d_a, const double* d_b,
                               double* d_sampled, int32_t* d_counts)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (!cell_offsets) return set_error(EQ_ERR_INVALID, "NULL cell offsets");
    if (ntraps < 1 || trap0 < 0 || trap0 + ntraps > b->ntraps) return set_error(EQ_ERR_INVALID, "trap index out of range");
    if (!(npm > 0)) return set_error(EQ_ERR_INVALID, "npm must be > 0");
    for (int k = 0; k < ntraps; ++k)
        if (cell_offsets[k + 1] < cell_offsets[k]) return set_error(EQ_ERR_INVALID, "cell offsets must not decrease");
    const int64_t nc = cell_offsets[ntraps] - cell_offsets[0];
    if (nc == 0) return EQ_OK;
    if (nc > ((int64_t)1 << 30)) return set_error(EQ_ERR_INVALID, "too many cells");
    if (!d_cells || !d_a || !d_b) return set_error(EQ_ERR_INVALID, "NULL cells, a or b");
    hipSetDevice(b->device);
    int rc = ensure_traps(b);
    if (rc) return rc;
    hipStream_t s = b->stream;
    if ((rc = upload_cell_off(b, cell_offsets, ntraps))) return rc;
    const size_t geo_bytes = align256(sizeof(CellGeo) * nc);
    if ((rc = ensure_scratch(b, geo_bytes))) return rc;
    CellGeo* geo = (CellGeo*)b->cell_scratch;
    const TrapGrid* traps = (const TrapGrid*)b->d_cell_traps;
    const int ns = b->nspecies;
    launch(b, EQ_KERNEL_COUPLE_SETUP, (double)sizeof(eq_cell) * nc + (double)sizeof(CellGeo) * nc, k_cell_geo,
           dim3((unsigned)((nc + 3) / 4)), dim3(256), 0, s, d_cells, (int64_t)nc, (const int64_t*)b->d_cell_off,
           (int)ntraps, (int)trap0, traps, npm, geo);
    // every covered node read and written once per species, the cells' records and amounts
    launch(b, EQ_KERNEL_COUPLE, (double)sizeof(CellGeo) * nc + 24.0 * ns * nc, k_couple_affine, dim3((unsigned)ntraps),
           dim3(kAThreads), 0, s, traps, (int)trap0, (const int64_t*)b->d_cell_off, (const CellGeo*)geo,
           (double* const*)b->d_field_ptrs, ns, d_a, d_b, d_sampled, d_counts, npm, (const int64_t*)nullptr,
           (int64_t)0);
    EQ_HIP(hipGetLastError());
    return EQ_OK;
}

int eq_batch_couple_traps_dev_noalloc(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                                      const eq_cell* d_cells, double npm, const double* d_q, double* d_sampled,
                                      int32_t* d_counts, uint32_t flags)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (!cell_offsets) return set_error(EQ_ERR_INVALID, "NULL cell offsets");
    if (b->device < 0 || b->device >= kMaxDevices || !g_scratch_ptr[b->device])
        return set_error(EQ_ERR_INVALID, "no fixed scratch on the batch's device");
    hipSetDevice(b->device);
    CoupleScratch fx;
    fx.base = (char*)g_scratch_ptr[b->device];
    fx.cap = kFixedScratchBytes;
    return couple_impl(b, trap0, ntraps, cell_offsets, d_cells, npm, d_q, d_sampled, d_counts, flags, &fx);
}

int eq_batch_couple(eq_batch* b, int32_t trap, const eq_cell* cells, size_t nc, double npm, const double* q,
                    double* sampled, int32_t* counts, uint32_t flags)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (nc == 0) return EQ_OK;
    if (!cells || (!q && (flags & EQ_COUPLE_DEPOSIT))) return set_error(EQ_ERR_INVALID, "NULL cells or q");
    hipSetDevice(b->device);
    const int ns = b->nspecies;
    const size_t cb = sizeof(eq_cell) * nc, qb = sizeof(double) * nc * ns, sb = sizeof(double) * nc * ns,
                 kb = sizeof(int32_t) * nc;
    char* buf = nullptr;
    const size_t tot = align256(cb) + align256(qb) + align256(sb) + align256(kb);
    if (hipMalloc((void**)&buf, tot) != hipSuccess) return set_error(EQ_ERR_NOMEM, "couple staging");
    eq_cell* dc = (eq_cell*)buf;
    double* dq = (double*)(buf + align256(cb));
    double* ds = (double*)(buf + align256(cb) + align256(qb));
    int32_t* dk = (int32_t*)(buf + align256(cb) + align256(qb) + align256(sb));
    int rc = EQ_OK;
    hipStream_t s = b->stream;
    if (hipMemcpyAsync(dc, cells, cb, hipMemcpyHostToDevice, s) != hipSuccess ||
        (q && hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, s) != hipSuccess)) {
        rc = set_error(EQ_ERR_HIP, "couple upload");
    } else if (!q) {
        hipMemsetAsync(dq, 0, qb, s);
    }
    const int64_t off[2] = {0, (int64_t)nc};
    if (rc == EQ_OK)
        rc = couple_impl(b, trap, 1, off, dc, npm, dq, sampled ? ds : nullptr, counts ? dk : nullptr, flags);
    if (rc == EQ_OK && sampled && hipMemcpyAsync(sampled, ds, sb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    if (rc == EQ_OK && counts && hipMemcpyAsync(counts, dk, kb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    hipStreamSynchronize(s);
    hipFree(buf);
    return rc;
}

} // extern "C"
