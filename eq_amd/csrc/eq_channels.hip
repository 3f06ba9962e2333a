// eq_channels.hip -- flow channels coupled to the traps, and per-node wall
// values (fenicsInterface::stepDiffusion with MICROFLUIDIC_TRAP and a channel
// trap type, src/fHSL.cpp:110-152).
//
// A channel is a 1-D P1 interval mesh over the trap's columns (fenicsChannel,
// src/fHSL.h:240-332, with zero channel extensions: src/fHSL.cpp:262-275).
// Its values are the trap's Dirichlet data on that wall (DirichletBC bound to
// the channel Function, src/fHSL.cpp:514, 531).  After every trap solve:
//   flux_j = dt*D*(u_inner,j - u_wall,j)/wellScaling          (src/fHSL.cpp:54-96)
//   repeat substeps times:  c += flux/substeps;  Crank-Nicolson step of
//     dt/substeps of   w(c - c0) + k D w' (c0+c)'/2 + k v (c0+c)'/2 w
//                      + k r_{1,2} ((c0+c)/2 - s_{1,2}) w |_{x=0, W}
//   (fenics/AdvectionDiffusion.ufl:62-64), a tridiagonal system solved by the
//   Thomas algorithm (diagonally dominant for the reference's step control,
//   src/main.cpp:414-430).
// Channels are small (nx+1 unknowns) and independent: one thread per channel.
#include <hip/hip_runtime.h>

#include <array>
#include <vector>

#include "eq_internal.h"

namespace eqk {

struct ChannelDev {
    eq_channel_desc ch;
    double* c;         // channel values (nx+1), also the wall's Dirichlet data
    double* cp;        // Thomas scratch (nx+1 each)
    double* dp;
    const double* u;   // node (0,0) of the trap field
    int32_t nx, ny, pitch, top; // top: 1 = channel above the trap (wall row ny)
    double hc, D, dt;
};

__global__ __launch_bounds__(64) void k_channels(const ChannelDev* __restrict__ chans, int n)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= n) return;
    const ChannelDev C = chans[t];
    const int N = C.nx;
    const int nsub = C.ch.substeps > 0 ? C.ch.substeps : 1;
    const double k = C.dt / nsub, hc = C.hc, D = C.D, v = C.ch.v;
    const double m1 = hc / 6.0, m2 = 2.0 * hc / 6.0;
    // per-row T = D K + v C (+ Robin) entries; interior rows: lower, diag, upper
    const double Tl = -D / hc - 0.5 * v, Td = 2.0 * D / hc, Tu = -D / hc + 0.5 * v;
    // boundary rows: node 0 lies on one element (local 0), node N on one (local 1)
    const double Td0 = D / hc - 0.5 * v + C.ch.r_left, Tu0 = -D / hc + 0.5 * v;
    const double TlN = -D / hc - 0.5 * v, TdN = D / hc + 0.5 * v + C.ch.r_right;
    const int64_t rin = C.top ? (int64_t)(C.ny - 1) * C.pitch : C.pitch;   // inner row
    const int64_t rwall = C.top ? (int64_t)C.ny * C.pitch : 0;               // wall row
    const double fscale = C.dt * D / C.ch.well_scaling / (double)nsub;
    for (int s = 0; s < nsub; ++s) {
        // forward sweep; the right-hand side B c is formed from the values
        // before this substep's solve (kept in registers as the sweep moves)
        double cprev = 0.0;
        double ccur = C.c[0] + fscale * (C.u[rin] - C.u[rwall]);
        for (int i = 0; i <= N; ++i) {
            const double cnext = (i < N) ? C.c[i + 1] + fscale * (C.u[rin + i + 1] - C.u[rwall + i + 1]) : 0.0;
            double a, b, c, r;
            if (i == 0) {
                a = 0.0;
                b = m2 + 0.5 * k * Td0;
                c = m1 + 0.5 * k * Tu0;
                r = (m2 - 0.5 * k * Td0) * ccur + (m1 - 0.5 * k * Tu0) * cnext + k * C.ch.r_left * C.ch.s_left;
            } else if (i == N) {
                a = m1 + 0.5 * k * TlN;
                b = m2 + 0.5 * k * TdN;
                c = 0.0;
                r = (m1 - 0.5 * k * TlN) * cprev + (m2 - 0.5 * k * TdN) * ccur + k * C.ch.r_right * C.ch.s_right;
            } else {
                a = m1 + 0.5 * k * Tl;
                b = 2.0 * m2 + 0.5 * k * Td;
                c = m1 + 0.5 * k * Tu;
                r = (m1 - 0.5 * k * Tl) * cprev + (2.0 * m2 - 0.5 * k * Td) * ccur + (m1 - 0.5 * k * Tu) * cnext;
            }
            if (i == 0) {
                C.cp[0] = c / b;
                C.dp[0] = r / b;
            } else {
                const double m = b - a * C.cp[i - 1];
                C.cp[i] = c / m;
                C.dp[i] = (r - a * C.dp[i - 1]) / m;
            }
            cprev = ccur;
            ccur = cnext;
        }
        double x = C.dp[N];
        C.c[N] = x;
        for (int i = N - 1; i >= 0; --i) {
            x = C.dp[i] - C.cp[i] * x;
            C.c[i] = x;
        }
    }
}

} // namespace eqk

using namespace eqk;

namespace {

// re-send level 0's problem table after a change of its wall-value pointers
int upload_level0(eq_batch* b)
{
    Level& L = b->levels[0];
    EQ_HIP(hipStreamSynchronize(b->stream));
    const int rc = pull_well_values(b); // keep the wall values the wells wrote
    if (rc) return rc;
    EQ_HIP(hipMemcpy(L.d_probs, L.probs.data(), sizeof(LevelProb) * L.probs.size(), hipMemcpyHostToDevice));
    return EQ_OK;
}

int wall_length(const LevelProb& P, int wall) { return (wall == kTop || wall == kBottom) ? P.nx + 1 : P.ny + 1; }

} // namespace

namespace eqk {

// the channels after a solve (called by eq_batch_step)
int channels_after_step(eq_batch* b)
{
    if (b->chan_count == 0) return EQ_OK;
    double bytes = 0.0; // per channel: its values read and written, the trap's two rows read, per substep
    for (size_t p = 0; p < b->chan_desc.size(); ++p)
        for (int k = 0; k < 2; ++k)
            if (b->chan_desc[p][k].substeps > 0)
                bytes += 32.0 * (b->levels[0].probs[p].nx + 1) * b->chan_desc[p][k].substeps;
    launch(b, EQ_KERNEL_BOUNDARY, bytes, k_channels, dim3((b->chan_count + 63) / 64), dim3(64), 0, b->stream,
           (const ChannelDev*)b->d_chans, b->chan_count);
    EQ_HIP(hipGetLastError());
    return EQ_OK;
}

void free_channels(eq_batch* b)
{
    for (auto& w : b->wall_vals)
        for (double* p : w) hipFree(p);
    b->wall_vals.clear();
    hipFree(b->d_chans);
    b->d_chans = nullptr;
    for (double* p : b->chan_scratch) hipFree(p);
    b->chan_scratch.clear();
}

} // namespace eqk

namespace {

// rebuild the device channel table from the host descriptions
int rebuild_channels(eq_batch* b)
{
    EQ_HIP(hipStreamSynchronize(b->stream)); // the old table and scratch are no longer in use
    for (double* p : b->chan_scratch) hipFree(p);
    b->chan_scratch.clear();
    std::vector<ChannelDev> v;
    const Level& L = b->levels[0];
    for (int p = 0; p < b->nprob; ++p) {
        for (int side = 0; side < 2; ++side) {
            const int wall = side == 0 ? kTop : kBottom;
            if (b->chan_desc[p][side].substeps == 0) continue; // no channel
            const LevelProb& P = L.probs[p];
            ChannelDev c{};
            c.ch = b->chan_desc[p][side];
            c.c = b->wall_vals[p][wall];
            const size_t n = (size_t)P.nx + 1;
            double* scratch = nullptr;
            if (hipMalloc((void**)&scratch, sizeof(double) * 2 * n) != hipSuccess)
                return set_error(EQ_ERR_NOMEM, "channel scratch");
            b->chan_scratch.push_back(scratch);
            c.cp = scratch;
            c.dp = scratch + n;
            c.u = b->u + P.origin;
            c.nx = P.nx;
            c.ny = P.ny;
            c.pitch = P.pitch;
            c.top = side == 0 ? 1 : 0;
            c.hc = b->descs[p].width / P.nx;
            c.D = b->descs[p].D;
            c.dt = b->descs[p].dt;
            v.push_back(c);
        }
    }
    hipFree(b->d_chans);
    b->d_chans = nullptr;
    b->chan_count = (int32_t)v.size();
    if (v.empty()) return EQ_OK;
    if (hipMalloc(&b->d_chans, sizeof(ChannelDev) * v.size()) != hipSuccess)
        return set_error(EQ_ERR_NOMEM, "channel table");
    EQ_HIP(hipMemcpy(b->d_chans, v.data(), sizeof(ChannelDev) * v.size(), hipMemcpyHostToDevice));
    return EQ_OK;
}

// device array of one wall's per-node values (allocated on first use)
int ensure_wall_array(eq_batch* b, int32_t prob, int32_t wall, double** out)
{
    if ((int)b->wall_vals.size() < b->nprob) b->wall_vals.resize(b->nprob, std::array<double*, 4>{});
    double*& a = b->wall_vals[prob][wall];
    if (!a) {
        const int n = wall_length(b->levels[0].probs[prob], wall);
        if (hipMalloc((void**)&a, sizeof(double) * n) != hipSuccess) return set_error(EQ_ERR_NOMEM, "wall values");
        EQ_HIP(hipMemset(a, 0, sizeof(double) * n));
    }
    *out = a;
    return EQ_OK;
}

} // namespace

extern "C" {

int eq_batch_set_wall_values(eq_batch* b, int32_t prob, int32_t wall, const double* values, size_t n)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    if (wall < 0 || wall > 3) return set_error(EQ_ERR_INVALID, "wall index out of range");
    hipSetDevice(b->device);
    LevelProb& P = b->levels[0].probs[prob];
    if (!(P.dmask & (1 << wall))) return set_error(EQ_ERR_INVALID, "the wall is not Dirichlet");
    if ((wall == kTop || wall == kBottom) && (int)b->chan_desc.size() > prob &&
        b->chan_desc[prob][wall == kTop ? 0 : 1].substeps != 0)
        return set_error(EQ_ERR_INVALID, "the wall's values come from its channel");
    if (!values) {
        P.gw[wall] = nullptr;
        return upload_level0(b);
    }
    if (n != (size_t)wall_length(P, wall)) return set_error(EQ_ERR_INVALID, "wrong number of wall values");
    double* a = nullptr;
    int rc = ensure_wall_array(b, prob, wall, &a);
    if (rc) return rc;
    EQ_HIP(hipStreamSynchronize(b->stream));
    EQ_HIP(hipMemcpy(a, values, sizeof(double) * n, hipMemcpyHostToDevice));
    P.gw[wall] = a;
    return upload_level0(b);
}

int eq_batch_set_channels(eq_batch* b, int32_t prob, const eq_channel_desc* top, const eq_channel_desc* bottom)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    hipSetDevice(b->device);
    LevelProb& P = b->levels[0].probs[prob];
    const eq_channel_desc* d[2] = {top, bottom};
    const int walls[2] = {kTop, kBottom};
    for (int side = 0; side < 2; ++side) {
        if (!d[side]) continue;
        if (!(P.dmask & (1 << walls[side]))) return set_error(EQ_ERR_INVALID, "a channel's wall must be Dirichlet");
        if (!(d[side]->well_scaling > 0) || d[side]->substeps < 1)
            return set_error(EQ_ERR_INVALID, "channel needs well_scaling > 0 and substeps >= 1");
    }
    if ((int)b->chan_desc.size() < b->nprob) b->chan_desc.resize(b->nprob);
    for (int side = 0; side < 2; ++side) {
        eq_channel_desc& slot = b->chan_desc[prob][side];
        if (!d[side]) {
            if (slot.substeps != 0) P.gw[walls[side]] = nullptr;
            slot = eq_channel_desc{};
            continue;
        }
        double* a = nullptr;
        int rc = ensure_wall_array(b, prob, walls[side], &a);
        if (rc) return rc;
        EQ_HIP(hipMemset(a, 0, sizeof(double) * wall_length(P, walls[side]))); // channels start at zero
        slot = *d[side];
        P.gw[walls[side]] = a;
    }
    int rc = upload_level0(b);
    if (rc) return rc;
    return rebuild_channels(b);
}

int eq_batch_channel_values(eq_batch* b, int32_t prob, double* top, double* bottom)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    hipSetDevice(b->device);
    EQ_HIP(hipStreamSynchronize(b->stream));
    const LevelProb& P = b->levels[0].probs[prob];
    double* out[2] = {top, bottom};
    for (int side = 0; side < 2; ++side) {
        if (!out[side]) continue;
        if ((int)b->chan_desc.size() <= prob || b->chan_desc[prob][side].substeps == 0)
            return set_error(EQ_ERR_INVALID, "no channel on that side");
        EQ_HIP(hipMemcpy(out[side], b->wall_vals[prob][side == 0 ? kTop : kBottom], sizeof(double) * (P.nx + 1),
                         hipMemcpyDeviceToHost));
    }
    return EQ_OK;
}

} // extern "C"
