// eq_cells.hip -- cell <-> mesh coupling of eQ's updateCells on gfx950.
//
// Reference semantics (src/abm/eQabm.cpp:234-449): for every cell in list
// order, find the covered nodes S_c (findInteriorPoints, 268-305), read the
// mean of every species field over S_c (readHSL, 326-337), then add
// q_c/|S_c| to every covered node of every species (writeHSL, 338-359).  A
// cell therefore samples the deposits of all earlier cells.
//
// GPU plan (no atomics, deterministic):
//   1. k_cell_geo    one wavefront per cell: pole window, |S_c| (ballot counts)
//   2. k_scan        exclusive scan of (window rows x row segments) per cell
//   3. k_cell_rows   one wavefront per (row, 4096-column segment, species):
//                    the row segment lives in LDS; the wave walks the cells
//                    in list order (64 at a time by ballot), for each cell
//                    covering the row it sums the covered values (the read)
//                    and then adds q_c/|S_c| (the write).  Per node the
//                    additions happen in list order, so the deposited field
//                    is bit-identical to the sequential loop.
//   4. k_cell_sample one thread per cell: mean over its row partials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "eq_internal.h"

namespace eqk {

constexpr int kSeg = 4096; // columns per LDS row segment (32 KB of FP64)

struct CellGeo {
    double cx, cy, cs, sn, o, rad; // centre, cos/sin of the angle, box half-length/half-width
    int32_t i1, i2, j1, j2;        // search window (inclusive)
    int32_t cnt;                   // nodes found by the scan (0: centre-node fallback)
    int32_t ci, cj;                // centre node (fallback)
    int32_t nseg0;                 // first row segment touched
    int64_t poff;                  // offset of this cell's row partials
};

struct GridDesc {
    int32_t nodes_w, nodes_h;
    double npm, trap_w, trap_h;
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
// a rounded difference is the sign of the exact difference).
__device__ __forceinline__ bool inside_cell(const CellGeo& g, double X, double Y)
{
    const double itx = (-g.sn) * g.cy - g.cs * g.cx;
    const double ity = g.sn * g.cx - g.cs * g.cy;
    const double lx = g.cs * X + g.sn * Y + itx;
    const double ly = (-g.sn) * X + g.cs * Y + ity;
    return (lx - g.o < 0.0) && (ly - g.rad < 0.0) && (ly + g.rad > 0.0) && (lx + g.o > 0.0);
}

// Ecoli::updatePoleCenters (src/abm/Ecoli.cpp:36-63)
__device__ __forceinline__ void pole(const GridDesc& G, double cx, double cy, double cs, double sn, double L, double w,
                                     int r, double& px, double& py)
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

__global__ __launch_bounds__(256) void k_cell_geo(const eq_cell* __restrict__ cells, int nc, GridDesc G,
                                                  CellGeo* __restrict__ geo)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nc) return;
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
    const int32_t nte = (int32_t)round(G.npm * 1.0 / 2.0);
    double ax, ay, bx, by;
    pole(G, g.cx, g.cy, g.cs, g.sn, cell.length, cell.width, 1, ax, ay);
    pole(G, g.cx, g.cy, g.cs, g.sn, cell.length, cell.width, -1, bx, by);
    const int32_t Ai = to_index(ay, G.npm), Aj = to_index(ax, G.npm);
    const int32_t Bi = to_index(by, G.npm), Bj = to_index(bx, G.npm);
    int32_t i1 = min(Ai, Bi), i2 = max(Ai, Bi), j1 = min(Aj, Bj), j2 = max(Aj, Bj);
    i1 = (i1 >= nte) ? i1 - nte : 0;
    j1 = (j1 >= nte) ? j1 - nte : 0;
    i2 = (i2 + nte >= G.nodes_h - 1) ? G.nodes_h - 1 : i2 + nte;
    j2 = (j2 + nte >= G.nodes_w - 1) ? G.nodes_w - 1 : j2 + nte;
    g.i1 = i1; g.i2 = i2; g.j1 = j1; g.j2 = j2;
    const int32_t ww = j2 - j1 + 1;
    const int64_t area = (i2 >= i1 && j2 >= j1) ? (int64_t)(i2 - i1 + 1) * ww : 0;
    int32_t cnt = 0;
    for (int64_t t0 = 0; t0 < area; t0 += 64) {
        const int64_t t = t0 + lane;
        bool in = false;
        if (t < area) {
            const int32_t i = i1 + (int32_t)(t / ww), j = j1 + (int32_t)(t % ww);
            in = inside_cell(g, (double)j / G.npm, (double)i / G.npm);
        }
        cnt += __popcll(__ballot(in));
    }
    g.cnt = cnt;
    // fallback: the centre node (src/abm/eQabm.cpp:299-303)
    g.ci = min(to_index(g.cy, G.npm), G.nodes_h - 1);
    g.cj = min(to_index(g.cx, G.npm), G.nodes_w - 1);
    g.nseg0 = (cnt > 0) ? j1 / kSeg : g.cj / kSeg;
    g.poff = 0;
    if (lane == 0) geo[c] = g;
}

__device__ __forceinline__ int64_t cell_slots(const CellGeo& g)
{
    if (g.cnt == 0) return 1;
    return (int64_t)(g.i2 - g.i1 + 1) * (g.j2 / kSeg - g.j1 / kSeg + 1);
}

// exclusive scan of the per-cell partial slots (single block, deterministic)
__global__ __launch_bounds__(1024) void k_scan(CellGeo* geo, int nc, int64_t* total)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (nc + nt - 1) / nt;
    const int lo = min(nc, t * per), hi = min(nc, lo + per);
    int64_t s = 0;
    for (int c = lo; c < hi; ++c) s += cell_slots(geo[c]);
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t acc = 0;
        for (int k = 0; k < nt; ++k) {
            const int64_t v = part[k];
            part[k] = acc;
            acc += v;
        }
        *total = acc;
    }
    __syncthreads();
    int64_t acc = part[t];
    for (int c = lo; c < hi; ++c) {
        geo[c].poff = acc;
        acc += cell_slots(geo[c]);
    }
}

__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// one wavefront per (row, segment); blockIdx.y = species
__global__ __launch_bounds__(64) void k_cell_rows(const CellGeo* __restrict__ geo, int nc, GridDesc G, int nseg,
                                                  double* const* __restrict__ fields, const double* __restrict__ q,
                                                  int ns, double* __restrict__ partial, int64_t total,
                                                  unsigned flags, int pitch)
{
    __shared__ double row[kSeg];
    const int lane = threadIdx.x;
    const int i = blockIdx.x / nseg, seg = blockIdx.x - (blockIdx.x / nseg) * nseg;
    const int s = blockIdx.y;
    const int jb = seg * kSeg, je = min(G.nodes_w, jb + kSeg) - 1;
    double* U = fields[s] + (int64_t)i * pitch;
    for (int j = jb + lane; j <= je; j += 64) row[j - jb] = U[j];
    __syncthreads();
    const bool dep = (flags & EQ_COUPLE_DEPOSIT) != 0;
    const double Y = (double)i / G.npm;
    double* part = partial + (int64_t)s * total;
    for (int cb = 0; cb < nc; cb += 64) {
        bool hit = false;
        const int c = cb + lane;
        if (c < nc) {
            const CellGeo& g = geo[c];
            if (g.cnt > 0)
                hit = (i >= g.i1 && i <= g.i2 && g.j1 <= je && g.j2 >= jb);
            else
                hit = (i == g.ci && g.cj >= jb && g.cj <= je);
        }
        uint64_t mask = __ballot(hit);
        while (mask) {
            const int k = __ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1;
            const CellGeo g = geo[cb + k];
            const double d = q[(int64_t)(cb + k) * ns + s] / (double)(g.cnt > 0 ? g.cnt : 1);
            double sum = 0.0;
            int64_t slot;
            if (g.cnt == 0) {
                if (lane == 0) {
                    const double v = row[g.cj - jb];
                    sum = v;
                    if (dep) row[g.cj - jb] = v + d;
                }
                slot = g.poff;
            } else {
                const int lo = max(g.j1, jb), hi = min(g.j2, je);
                for (int j = lo + lane; j <= hi; j += 64) {
                    if (inside_cell(g, (double)j / G.npm, Y)) {
                        const double v = row[j - jb];
                        sum += v;
                        if (dep) row[j - jb] = v + d;
                    }
                }
                const int nsegc = g.j2 / kSeg - g.j1 / kSeg + 1;
                slot = g.poff + (int64_t)(i - g.i1) * nsegc + (seg - g.nseg0);
            }
            sum = wave_sum64(sum);
            if (lane == 0) part[slot] = sum;
        }
    }
    __syncthreads();
    if (dep)
        for (int j = jb + lane; j <= je; j += 64) U[j] = row[j - jb];
}

// sampled[c,s] = (sum of the cell's row partials in row order) / |S_c|
__global__ __launch_bounds__(256) void k_cell_sample(const CellGeo* __restrict__ geo, int nc, int ns,
                                                     const double* __restrict__ partial, int64_t total,
                                                     double* __restrict__ sampled, int32_t* __restrict__ counts)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const CellGeo g = geo[c];
    const int64_t n = cell_slots(g);
    const int cnt = g.cnt > 0 ? g.cnt : 1;
    if (counts) counts[c] = cnt;
    if (!sampled) return;
    for (int s = 0; s < ns; ++s) {
        const double* part = partial + (int64_t)s * total + g.poff;
        double acc = 0.0;
        for (int64_t k = 0; k < n; ++k) acc += part[k];
        sampled[(int64_t)c * ns + s] = acc / (double)cnt;
    }
}

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

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static int couple_impl(eq_batch* b, int32_t trap, const eq_cell* d_cells, size_t nc, double npm, const double* d_q,
                       double* d_sampled, int32_t* d_counts, uint32_t flags)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (trap < 0 || trap >= b->ntraps) return set_error(EQ_ERR_INVALID, "trap index out of range");
    if (!(npm > 0)) return set_error(EQ_ERR_INVALID, "npm must be > 0");
    if (nc == 0) return EQ_OK;
    if (nc > (size_t)1 << 30) return set_error(EQ_ERR_INVALID, "too many cells");
    if (!d_cells || (!d_q && (flags & EQ_COUPLE_DEPOSIT))) return set_error(EQ_ERR_INVALID, "NULL cells or q");
    const int ns = b->nspecies;
    const int p0 = trap * ns;
    const LevelProb& P = b->levels[0].probs[p0];
    const eq_field_desc& d = b->descs[p0];
    GridDesc G;
    G.nodes_w = P.nx + 1;
    G.nodes_h = P.ny + 1;
    G.npm = npm;
    G.trap_w = d.width;
    G.trap_h = d.height;
    hipStream_t s = b->stream;
    // scratch: geo | total | field pointers | partials
    const size_t geo_bytes = align256(sizeof(CellGeo) * nc);
    const size_t head = geo_bytes + 256 + align256(sizeof(double*) * ns);
    int rc = ensure_scratch(b, head + 256);
    if (rc) return rc;
    char* base = (char*)b->cell_scratch;
    CellGeo* geo = (CellGeo*)base;
    int64_t* d_total = (int64_t*)(base + geo_bytes);
    k_cell_geo<<<(unsigned)((nc + 3) / 4), 256, 0, s>>>(d_cells, (int)nc, G, geo);
    k_scan<<<1, 1024, 0, s>>>(geo, (int)nc, d_total);
    int64_t total = 0;
    EQ_HIP(hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EQ_HIP(hipStreamSynchronize(s));
    const size_t need = head + align256(sizeof(double) * (size_t)total * ns);
    if (need > b->cell_scratch_bytes) {
        // grow (keeps nothing: recompute the geometry in the new buffer)
        if ((rc = ensure_scratch(b, need))) return rc;
        base = (char*)b->cell_scratch;
        geo = (CellGeo*)base;
        d_total = (int64_t*)(base + geo_bytes);
        k_cell_geo<<<(unsigned)((nc + 3) / 4), 256, 0, s>>>(d_cells, (int)nc, G, geo);
        k_scan<<<1, 1024, 0, s>>>(geo, (int)nc, d_total);
    }
    double** d_fields = (double**)(base + geo_bytes + 256);
    double* partial = (double*)(base + head);
    std::vector<double*> hf(ns);
    for (int k = 0; k < ns; ++k) hf[k] = b->u + b->levels[0].probs[p0 + k].origin;
    EQ_HIP(hipMemcpyAsync(d_fields, hf.data(), sizeof(double*) * ns, hipMemcpyHostToDevice, s));
    const int nseg = (G.nodes_w + kSeg - 1) / kSeg;
    prof_begin(b, EQ_KERNEL_COUPLE);
    k_cell_rows<<<dim3((unsigned)(G.nodes_h * nseg), (unsigned)ns), 64, 0, s>>>(
        geo, (int)nc, G, nseg, d_fields, d_q, ns, partial, total, flags, P.pitch);
    // algorithmic bytes: every row read (+written when depositing) once per species
    prof_end(b, EQ_KERNEL_COUPLE,
             (double)G.nodes_w * G.nodes_h * ns * sizeof(double) * ((flags & EQ_COUPLE_DEPOSIT) ? 2 : 1));
    if (d_sampled || d_counts)
        k_cell_sample<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(
            geo, (int)nc, ns, partial, total, (flags & EQ_COUPLE_SAMPLE) ? d_sampled : nullptr, d_counts);
    EQ_HIP(hipGetLastError());
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
    return couple_impl(b, trap, d_cells, nc, npm, d_q, d_sampled, d_counts, flags);
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
    if (rc == EQ_OK)
        rc = couple_impl(b, 0 + trap, dc, nc, npm, dq, sampled ? ds : nullptr, counts ? dk : nullptr, flags);
    if (rc == EQ_OK && sampled && hipMemcpyAsync(sampled, ds, sb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    if (rc == EQ_OK && counts && hipMemcpyAsync(counts, dk, kb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    hipStreamSynchronize(s);
    hipFree(buf);
    return rc;
}

} // extern "C"
