// eq_cells.hip -- cell <-> mesh coupling of eQ's updateCells on gfx950.
//
// Reference semantics (src/abm/eQabm.cpp:234-449): for every cell in list
// order, find the covered nodes S_c (findInteriorPoints, 268-305), read the
// mean of every species field over S_c (readHSL, 326-337), then add
// q_c/|S_c| to every covered node of every species (writeHSL, 338-359).  A
// cell therefore samples the deposits of all earlier cells.
//
// GPU plan (deterministic, no atomics anywhere):
//   1. k_cell_geo    one wavefront per cell: pole window, |S_c| (ballot counts)
//   2. k_scan        exclusive scan of (window rows x row segments) per cell
//   3. k_band_masks  one wavefront per band of kBand rows: bit c of word c/64
//                    says cell c's window meets the band (ballots)
//   4. k_cell_rows   one wavefront per (row, column segment, species): the
//                    segment lives in LDS; the wave visits the band's cells
//                    in list order, and for each cell covering the row sums
//                    the covered values (the read) and then adds q_c/|S_c|
//                    (the write).  Per node the additions happen in list
//                    order, so the deposited field is bit-identical to the
//                    sequential loop.
//   5. k_cell_sample one thread per cell: mean over its row partials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "eq_internal.h"

namespace eqk {

constexpr int kSegMax = 512;  // LDS row segment capacity (4 KB of FP64)
constexpr int kBand = 32;     // rows per band of the cell bitmasks

struct CellGeo {
    double cx, cy, cs, sn, o, rad; // centre, cos/sin of the angle, box half-length/half-width
    int32_t i1, i2, j1, j2;        // search window (inclusive); the centre node for fallback cells
    int32_t cnt;                   // nodes found by the scan (0: centre-node fallback)
    int32_t nseg0;                 // first row segment touched
    int64_t poff;                  // offset of this cell's row partials
};

struct GridDesc {
    int32_t nodes_w, nodes_h;
    double npm, trap_w, trap_h;
    int32_t segw, nseg;            // column segment width and count
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
__device__ __forceinline__ bool inside_cell(double cx, double cy, double cs, double sn, double o, double rad,
                                            double X, double Y)
{
    const double itx = (-sn) * cy - cs * cx;
    const double ity = sn * cx - cs * cy;
    const double lx = cs * X + sn * Y + itx;
    const double ly = (-sn) * X + cs * Y + ity;
    return (lx - o < 0.0) && (ly - rad < 0.0) && (ly + rad > 0.0) && (lx + o > 0.0);
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
    const int32_t ww = j2 - j1 + 1;
    const int64_t area = (i2 >= i1 && j2 >= j1) ? (int64_t)(i2 - i1 + 1) * ww : 0;
    int32_t cnt = 0;
    for (int64_t t0 = 0; t0 < area; t0 += 64) {
        const int64_t t = t0 + lane;
        bool in = false;
        if (t < area) {
            const int32_t i = i1 + (int32_t)(t / ww), j = j1 + (int32_t)(t % ww);
            in = inside_cell(g.cx, g.cy, g.cs, g.sn, g.o, g.rad, (double)j / G.npm, (double)i / G.npm);
        }
        cnt += __popcll(__ballot(in));
    }
    g.cnt = cnt;
    if (cnt == 0) {
        // fallback: the centre node (src/abm/eQabm.cpp:299-303)
        i1 = i2 = min(to_index(g.cy, G.npm), G.nodes_h - 1);
        j1 = j2 = min(to_index(g.cx, G.npm), G.nodes_w - 1);
    }
    g.i1 = i1; g.i2 = i2; g.j1 = j1; g.j2 = j2;
    g.nseg0 = j1 / G.segw;
    g.poff = 0;
    if (lane == 0) geo[c] = g;
}

__device__ __forceinline__ int64_t cell_slots(const CellGeo& g, int segw)
{
    return (int64_t)(g.i2 - g.i1 + 1) * (g.j2 / segw - g.j1 / segw + 1);
}

// exclusive scan of the per-cell partial slots (single block, deterministic)
__global__ __launch_bounds__(1024) void k_scan(CellGeo* geo, int nc, int segw, int64_t* total)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (nc + nt - 1) / nt;
    const int lo = min(nc, t * per), hi = min(nc, lo + per);
    int64_t s = 0;
    for (int c = lo; c < hi; ++c) s += cell_slots(geo[c], segw);
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
        acc += cell_slots(geo[c], segw);
    }
}

// one wavefront per band: bit (c % 64) of word c/64 = cell c's window meets the band
__global__ __launch_bounds__(64) void k_band_masks(const CellGeo* __restrict__ geo, int nc, int nwords,
                                                   uint64_t* __restrict__ masks)
{
    const int band = blockIdx.x, lane = threadIdx.x;
    const int r0 = band * kBand, r1 = r0 + kBand - 1;
    for (int w = 0; w < nwords; ++w) {
        const int c = w * 64 + lane;
        bool hit = false;
        if (c < nc) hit = (geo[c].i1 <= r1 && geo[c].i2 >= r0);
        const uint64_t m = __ballot(hit);
        if (lane == 0) masks[(int64_t)band * nwords + w] = m;
    }
}

__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double bcast(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int bcasti(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

__device__ __forceinline__ int64_t bcastl(int64_t v, int k)
{
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), k);
    return ((int64_t)hi << 32) | (unsigned int)lo;
}

// one wavefront per (row, segment); blockIdx.y = species
__global__ __launch_bounds__(64) void k_cell_rows(const CellGeo* __restrict__ geo, int nc, GridDesc G,
                                                  const uint64_t* __restrict__ masks, int nwords,
                                                  double* const* __restrict__ fields, const double* __restrict__ q,
                                                  int ns, double* __restrict__ partial, int64_t total,
                                                  unsigned flags, int pitch)
{
    __shared__ double row[kSegMax];
    const int lane = threadIdx.x;
    const int i = blockIdx.x / G.nseg, seg = blockIdx.x - i * G.nseg;
    const int s = blockIdx.y;
    const int jb = seg * G.segw, je = min(G.nodes_w, jb + G.segw) - 1;
    double* U = fields[s] + (int64_t)i * pitch;
    for (int j = jb + lane; j <= je; j += 64) row[j - jb] = U[j];
    __syncthreads();
    const bool dep = (flags & EQ_COUPLE_DEPOSIT) != 0;
    const double Y = (double)i / G.npm;
    double* part = partial + (int64_t)s * total;
    const uint64_t* bm = masks + (int64_t)(i / kBand) * nwords;
    for (int wb = 0; wb < nwords; wb += 64) {
        const uint64_t myword = (wb + lane < nwords) ? bm[wb + lane] : 0ull;
        uint64_t nz = __ballot(myword != 0ull);
        while (nz) {
            const int wl = __ffsll((unsigned long long)nz) - 1;
            nz &= nz - 1;
            const uint64_t word = (uint64_t)bcastl((int64_t)myword, wl);
            const int c = (wb + wl) * 64 + lane;
            bool hit = false;
            CellGeo g{};
            double qc = 0.0;
            if ((word >> lane) & 1ull) {
                g = geo[c];
                hit = (i >= g.i1 && i <= g.i2 && g.j1 <= je && g.j2 >= jb);
                if (hit) qc = q[(int64_t)c * ns + s];
            }
            uint64_t mask = __ballot(hit);
            while (mask) {
                const int k = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const double cx = bcast(g.cx, k), cy = bcast(g.cy, k), cs = bcast(g.cs, k), sn = bcast(g.sn, k);
                const double o = bcast(g.o, k), rad = bcast(g.rad, k);
                const int i1 = bcasti(g.i1, k), j1 = bcasti(g.j1, k), j2 = bcasti(g.j2, k);
                const int cnt = bcasti(g.cnt, k), nseg0 = bcasti(g.nseg0, k);
                const int64_t poff = bcastl(g.poff, k);
                const double d = bcast(qc, k) / (double)(cnt > 0 ? cnt : 1);
                const int lo = max(j1, jb), hi = min(j2, je);
                double sum = 0.0;
                for (int j = lo + lane; j <= hi; j += 64) {
                    if (cnt == 0 || inside_cell(cx, cy, cs, sn, o, rad, (double)j / G.npm, Y)) {
                        const double v = row[j - jb];
                        sum += v;
                        if (dep) row[j - jb] = v + d;
                    }
                }
                const int nsegc = j2 / G.segw - j1 / G.segw + 1;
                const int64_t slot = poff + (int64_t)(i - i1) * nsegc + (seg - nseg0);
                sum = wave_sum64(sum);
                if (lane == 0) part[slot] = sum;
            }
        }
    }
    __syncthreads();
    if (dep)
        for (int j = jb + lane; j <= je; j += 64) U[j] = row[j - jb];
}

// sampled[c,s] = (sum of the cell's row partials in row order) / |S_c|
__global__ __launch_bounds__(256) void k_cell_sample(const CellGeo* __restrict__ geo, int nc, int ns, int segw,
                                                     const double* __restrict__ partial, int64_t total,
                                                     double* __restrict__ sampled, int32_t* __restrict__ counts)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const CellGeo g = geo[c];
    const int64_t n = cell_slots(g, segw);
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
    // column segments of at most kSegMax nodes (more wavefronts in flight)
    G.nseg = (G.nodes_w + kSegMax - 1) / kSegMax;
    G.segw = (G.nodes_w + G.nseg - 1) / G.nseg;
    const int nwords = (int)((nc + 63) / 64);
    const int nbands = (G.nodes_h + kBand - 1) / kBand;
    hipStream_t s = b->stream;
    // scratch: geo | total | band masks | partials
    const size_t geo_bytes = align256(sizeof(CellGeo) * nc);
    const size_t mask_bytes = align256(sizeof(uint64_t) * (size_t)nwords * nbands);
    const size_t head = geo_bytes + 256 + mask_bytes;
    int rc = ensure_scratch(b, head + 256);
    if (rc) return rc;
    char* base = (char*)b->cell_scratch;
    CellGeo* geo = (CellGeo*)base;
    int64_t* d_total = (int64_t*)(base + geo_bytes);
    k_cell_geo<<<(unsigned)((nc + 3) / 4), 256, 0, s>>>(d_cells, (int)nc, G, geo);
    k_scan<<<1, 1024, 0, s>>>(geo, (int)nc, G.segw, d_total);
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
        k_scan<<<1, 1024, 0, s>>>(geo, (int)nc, G.segw, d_total);
    }
    uint64_t* masks = (uint64_t*)(base + geo_bytes + 256);
    double* const* d_fields = b->d_field_ptrs + p0; // the trap's species, in order
    double* partial = (double*)(base + head);
    k_band_masks<<<nbands, 64, 0, s>>>(geo, (int)nc, nwords, masks);
    prof_begin(b, EQ_KERNEL_COUPLE);
    k_cell_rows<<<dim3((unsigned)(G.nodes_h * G.nseg), (unsigned)ns), 64, 0, s>>>(
        geo, (int)nc, G, masks, nwords, d_fields, d_q, ns, partial, total, flags, P.pitch);
    // algorithmic bytes: every row read (and written when depositing) once per species
    prof_end(b, EQ_KERNEL_COUPLE,
             (double)G.nodes_w * G.nodes_h * ns * sizeof(double) * ((flags & EQ_COUPLE_DEPOSIT) ? 2 : 1));
    if (d_sampled || d_counts)
        k_cell_sample<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(
            geo, (int)nc, ns, G.segw, partial, total, (flags & EQ_COUPLE_SAMPLE) ? d_sampled : nullptr, d_counts);
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
        rc = couple_impl(b, trap, dc, nc, npm, dq, sampled ? ds : nullptr, counts ? dk : nullptr, flags);
    if (rc == EQ_OK && sampled && hipMemcpyAsync(sampled, ds, sb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    if (rc == EQ_OK && counts && hipMemcpyAsync(counts, dk, kb, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = set_error(EQ_ERR_HIP, "couple download");
    hipStreamSynchronize(s);
    hipFree(buf);
    return rc;
}

} // extern "C"
