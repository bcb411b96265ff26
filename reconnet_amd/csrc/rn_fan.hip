// 2D fan-beam projector kernels for gfx950 (wave64).
//
// Forward projectors are line driven: one lane per detector bin, a loop over
// the image lines (rows for X views, columns for Y views) that integrates a
// contiguous image line over the ray's footprint by direct overlap sums.
// There are no prefix sums (no FP32 cancellation) and no texture hardware.
// Backprojectors are pixel driven gathers: one lane per pixel, a loop over
// views, a register accumulator and a single store; no atomics.
//
// Reference counterparts: DDFPFanKernelX/Y (DistanceDriven.cu:64-196),
// DDBPFanKernelX/Y (:500-614), PreweightBPKernelX/Y (:441-477),
// SiddonFanProjectionKernel / SiddonFanBackprojectionKernel (Siddon.cu:240-382),
// bpFanKernel3D (PixelDriven.cu:11-75).

#include "rn_fan_device.h"
#include "rn_tune.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace rn {

namespace {

using namespace fan;

// ---------------------------------------------------------------------------
// forward projectors (line driven)
// ---------------------------------------------------------------------------

// XCD > 0: 1-D grid whose workgroups are renumbered per XCD (as in
// k_siddon_fan_fp); S slices per lane share the ray's lines and shares.
template <int BLK, int XCD, int S>
__global__ __launch_bounds__(BLK) void k_dd_fan_fp(FanK g, const float* __restrict__ RX,
                                                   const float* __restrict__ RY, float* __restrict__ prj,
                                                   int view0, int nviews, int ns)
{
    int tile, vi, sg;
    if (XCD > 0) {
        const int per = gridDim.x / XCD;
        const int l = (blockIdx.x % XCD) * per + blockIdx.x / XCD;
        const int tiles = (g.nu + BLK - 1) / BLK;
        tile = l % tiles;
        vi = (l / tiles) % nviews;
        sg = l / (tiles * nviews);
    } else {
        tile = blockIdx.x;
        vi = blockIdx.y;
        sg = blockIdx.z;
    }
    const int iu = tile * BLK + threadIdx.x;
    const bool active = iu < g.nu;
    const long plane = (long)g.nx * g.ny;
    const int s0 = sg * S;
    const int nsl = min(S, ns - s0);
    float acc[S];
    dd_fp_ray_slices<float, S>(g, RX + s0 * plane, RY + s0 * plane, plane, nsl, view0 + vi,
                               active ? iu : g.nu - 1, active, acc);
    if (!active) return;
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (q >= nsl) break;
        prj[((long)(s0 + q) * nviews + vi) * g.nu + iu] = acc[q];
    }
}

// DD FP of G interleaved groups of 4 slices per lane (padded quad averaged
// planes), H parts of a wave per ray; writes the plain sinograms of the real
// slices [s][nviews][nu].
template <int BLK, int XCD, int G, int H>
__global__ __launch_bounds__(BLK) void k_dd_fan_fp_q(FanK g, const float4* __restrict__ QX,
                                                     const float4* __restrict__ QY, float* __restrict__ prj,
                                                     int view0, int nviews, int ns)
{
    int tile, vi, sg;
    if (XCD > 0) {
        const int per = gridDim.x / XCD;
        const int l = (blockIdx.x % XCD) * per + blockIdx.x / XCD;
        const int tiles = (g.nu + BLK / H - 1) / (BLK / H);
        tile = l % tiles;
        vi = (l / tiles) % nviews;
        sg = l / (tiles * nviews);
    } else {
        tile = blockIdx.x;
        vi = blockIdx.y;
        sg = blockIdx.z;
    }
    const int iu = tile * (BLK / H) + (int)(threadIdx.x >> 6) * (64 / H) + (int)(threadIdx.x & (64 / H - 1));
    const bool active = iu < g.nu;
    const int nquads = (ns + 3) / 4;
    const int q0 = sg * G;
    const int nq = min(G, nquads - q0);
    float4 acc[G];
    dd_fp_ray_quads<G, H>(g, QX + (long)q0 * g.ny * qpitch(g.nx), QY + (long)q0 * g.nx * qpitch(g.ny), nq,
                          view0 + vi, active ? iu : g.nu - 1, active, acc);
    if (!active || (threadIdx.x & 63) >= 64 / H) return;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const float v[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
        #pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int sl = 4 * (q0 + k) + c;
            if (sl < ns) prj[((long)sl * nviews + vi) * g.nu + iu] = v[c];
        }
    }
}

// With meas != nullptr the epilogue writes the weighted residual
// wgt * (Ax - meas) of the OS-SQS data step (ossqs.cu:101-122) instead of Ax.
// Each lane projects its ray through S consecutive slices (the ray's lines,
// cells and weights are shared; the grid's slice index counts groups of S).
template <class T, int BLK, int XCD, int S>
__global__ __launch_bounds__(BLK) void k_siddon_fan_fp(FanK g, const T* __restrict__ Pyx,
                                                       const T* __restrict__ Pxy, T* __restrict__ prj,
                                                       int view0, int nviews, int ns,
                                                       const float* __restrict__ meas, const float* __restrict__ wgt,
                                                       int meas_nviews)
{
    // XCD > 0: 1-D grid of tiles * nviews * slice groups workgroups (a multiple
    // of XCD).  The hardware deals workgroups round-robin to the XCDs, so
    // workgroup h runs on XCD h % XCD; h is renumbered so that each XCD walks a
    // contiguous, slice-major range of (slice group, view, bin tile) and its L2
    // holds the images of only the slices it is working on.
    int tile, vi, sg;
    if (XCD > 0) {
        const int per = gridDim.x / XCD;
        const int l = (blockIdx.x % XCD) * per + blockIdx.x / XCD;
        const int tiles = (g.nu + BLK - 1) / BLK;
        tile = l % tiles;
        vi = (l / tiles) % nviews;
        sg = l / (tiles * nviews);
    } else {
        tile = blockIdx.x;
        vi = blockIdx.y;
        sg = blockIdx.z;
    }
    const int iu = tile * BLK + threadIdx.x;
    const bool active = iu < g.nu;
    const int view = view0 + vi;
    const long plane = (long)g.nx * g.ny;
    const int s0 = sg * S;
    const int nsl = min(S, ns - s0);
    T acc[S];
    siddon_fp_ray_slices<T, S>(g, Pyx + s0 * plane, Pxy + s0 * plane, plane, nsl, view, active ? iu : g.nu - 1,
                               active, acc);
    if (!active) return;
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (q >= nsl) break;
        const int s = s0 + q;
        T v = acc[q];
        if (meas != nullptr) {
            const long mi = ((long)s * meas_nviews + view) * g.nu + iu;
            v = (T)wgt[mi] * (v - (T)meas[mi]);
        }
        prj[((long)s * nviews + vi) * g.nu + iu] = v;
    }
}

// ---------------------------------------------------------------------------
// backprojectors (pixel driven gathers)
// ---------------------------------------------------------------------------

// One lane per pixel of S consecutive slices (the grid's z index counts
// groups of S): the bin ranges and weights are shared by the slices.
template <class T, int BX, int BY, int NB, int S>
__global__ __launch_bounds__(BX * BY) void k_siddon_fan_bp(FanK g, const T* __restrict__ prj,
                                                       T* __restrict__ img, int view0, int nviews, int ns,
                                                       int accumulate)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int s0 = blockIdx.z * S;
    const int nsl = min(S, ns - s0);
    if (ix >= g.nx || iy >= g.ny) return;
    const long sstride = (long)nviews * g.nu;
    T acc[S];
    siddon_bp_pixel_slices<T, NB, S>(g, prj + s0 * sstride, sstride, nsl, view0, nviews, ix, iy, acc);
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (q >= nsl) break;
        T* o = img + (long)(s0 + q) * g.nx * g.ny + (long)ix * g.ny + iy;
        *o = accumulate ? *o + acc[q] : acc[q];
    }
}

// Quad-layout Siddon FP + residual: each lane carries G groups of 4 slices
// (one float4 load per cell and group); the grid's slice index counts groups
// of G quads.  res = w * (Ax - meas) per component.
template <int BLK, int XCD, int G, int H>
__global__ __launch_bounds__(BLK) void k_siddon_fan_fp_q(FanK g, const float4* __restrict__ Qyx,
                                                         const float4* __restrict__ Qxy, float4* __restrict__ res,
                                                         int view0, int nviews, int nquads,
                                                         const float4* __restrict__ meas,
                                                         const float4* __restrict__ wgt, int meas_nviews,
                                                         int res_vstride, int res_voff)
{
    int tile, vi, sg;
    if (XCD > 0) {
        const int per = gridDim.x / XCD;
        const int l = (blockIdx.x % XCD) * per + blockIdx.x / XCD;
        const int tiles = (g.nu + BLK / H - 1) / (BLK / H);
        tile = l % tiles;
        vi = (l / tiles) % nviews;
        sg = l / (tiles * nviews);
    } else {
        tile = blockIdx.x;
        vi = blockIdx.y;
        sg = blockIdx.z;
    }
    // H > 1: each wave carries 64 / H rays H times (one part of the lines each)
    const int iu = tile * (BLK / H) + (int)(threadIdx.x >> 6) * (64 / H) + (int)(threadIdx.x & (64 / H - 1));
    const bool active = iu < g.nu;
    const int view = view0 + vi;
    const int q0 = sg * G;
    const int nq = min(G, nquads - q0);
    float4 acc[G];
    siddon_fp_ray_quads<G, H>(g, Qyx + (long)q0 * g.ny * qpitch(g.nx), Qxy + (long)q0 * g.nx * qpitch(g.ny), nq, view,
                           active ? iu : g.nu - 1, active, acc);
    if (!active || (threadIdx.x & 63) >= 64 / H) return;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const long mi = ((long)(q0 + k) * meas_nviews + view) * g.nu + iu;
        const float4 m = meas[mi], w = wgt[mi];
        res[((long)(q0 + k) * res_vstride + res_voff + vi) * g.nu + iu] =
            make_float4(w.x * (acc[k].x - m.x), w.y * (acc[k].y - m.y), w.z * (acc[k].z - m.z),
                        w.w * (acc[k].w - m.w));
    }
}

// Quad-layout Siddon BP: one lane per pixel of G groups of 4 slices (one
// float4 load per bin and group); writes the plain images of the real slices.
template <int BX, int BY, int NB, int G>
__global__ __launch_bounds__(BX * BY) void k_siddon_fan_bp_q(FanK g, const float4* __restrict__ prj,
                                                         float* __restrict__ img, int view0, int nviews, int ns)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int nquads = (ns + 3) / 4;
    const int q0 = blockIdx.z * G;
    const int nq = min(G, nquads - q0);
    if (ix >= g.nx || iy >= g.ny) return;
    const long sstride = (long)nviews * g.nu;
    float4 acc[G];
    siddon_bp_pixel_slices<float4, NB, G>(g, prj + q0 * sstride, sstride, nq, view0, nviews, ix, iy, acc);
    const long plane = (long)g.nx * g.ny, pix = (long)ix * g.ny + iy;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const int s = 4 * (q0 + k);
        const float v[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
        #pragma unroll
        for (int c = 0; c < 4; ++c)
            if (s + c < ns) img[(long)(s + c) * plane + pix] = v[c];
    }
}

// k_siddon_fan_bp_q with each wave's bin window staged in LDS.  Per view the
// wave's 64 pixels need the bins [wlo, whi] (the lane minimum and maximum of
// their ranges); the G groups' values and the view's ray table over CAP bins
// at a time are loaded once per wave with coalesced dwordx4 loads and written
// to the wave's own LDS rows, from which every pixel gathers its bins: a
// fraction of the per-pixel gathers' vector-memory traffic.  The first CAP
// bins of a view are loaded one view ahead into registers; a wider window
// takes further chunks of CAP bins, loaded in turn (the next view's loads are
// then issued after them).  Every view runs the one gather code (no branch
// around the accumulators, which would cost copies of all of them).  A wave
// touches only its own rows and the LDS instructions of one wave execute in
// order, so no barrier is needed; the fences keep the compiler from moving
// LDS accesses across the handovers.  A wave clears its rows at the start, so
// that a view with an empty window (every lane then takes weight 0 from slot
// 0) reads zeros or an earlier view's bins, never unwritten memory.  The
// group rows are addressed as 32-bit byte offsets from the groups' base (the
// launcher checks that G groups fit).
//
// UPD: the epilogue applies the OS-SQS update to the backprojected gradient
// instead of storing it (sqs_update_core per slice, the arithmetic of
// k_sqs_update_q): the current image and its 4-neighbours are read as float4
// groups from the interleaved copy u.Qold, the normaliser from u.D4, and the
// new image goes to u.Xn and, as float4 per group, to the other interleaved
// copies u.Qxy, u.Qyx (this launch reads none of the buffers it writes).
template <int BX, int BY, int NB, int G, int CAP, bool UPD = false>
__global__ __launch_bounds__(BX * BY, G >= 8 ? 4 : 1) void k_siddon_fan_bp_ql(FanK g, const float4* __restrict__ prj,
                                                          float* __restrict__ img, int view0, int nviews, int ns,
                                                          SqsUpd u)
{
    static_assert((BX * BY) % 64 == 0 && CAP <= 64, "whole waves, one load for the ray table row");
    // LDS rows 0 .. G - 1: the groups; row G: the ray table
    constexpr int kRows = G + 1, kLoads = (G * CAP + 63) / 64;
    __shared__ float4 stage[(BX * BY / 64) * kRows * CAP];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int lane = tid & 63;
    float4* st = stage + (tid >> 6) * kRows * CAP;
    for (int i = lane; i < kRows * CAP; i += 64) st[i] = vzero<float4>();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const bool inside = ix < g.nx && iy < g.ny;   // lanes outside stage and gather nothing
    const int nquads = (ns + 3) / 4;
    const int q0 = blockIdx.z * G;
    const int nq = min(G, nquads - q0);
    const long sstride = (long)nviews * g.nu;
    const float4* P = prj + q0 * sstride;
    const int cx = min(ix, g.nx - 1), cy = min(iy, g.ny - 1);
    const float xc = (cx - 0.5f * (g.nx - 1)) * g.dx, yc = (cy - 0.5f * (g.ny - 1)) * g.dy;
    const float rpix = 0.5f * sqrtf(g.dx * g.dx + g.dy * g.dy);
    const float inv_da = 1.0f / g.da;
    const float yr = (float)cy - 0.5f * g.ny, xr = (float)cx - 0.5f * g.nx;
    // group staging load j of this lane: element i = 64 j + lane of the group
    // rows, that is group i / CAP, bin i % CAP of the chunk (LDS slot i), at
    // byte offset roff[j] + (view row + bin) * 16 from P; the ray table row is
    // one more load by the first CAP lanes
    int bj[kLoads];
    unsigned roff[kLoads];
    bool has[kLoads];
    #pragma unroll
    for (int j = 0; j < kLoads; ++j) {
        const int i = j * 64 + lane, r = i / CAP;
        bj[j] = i % CAP;
        has[j] = r < G;
        roff[j] = (unsigned)(min(r, nq - 1) * sstride) * 16u;
    }
    const float4* const sray_v = g.sray + (long)view0 * g.nu;
    float4 acc[G];
    #pragma unroll
    for (int q = 0; q < G; ++q) acc[q] = vzero<float4>();
    float4 rg[kLoads], rr;
    // the pixel's bins of view vi (empty outside the image or the detector)
    auto range = [&](int vi, int& k0, int& k1) {
        if (g.unit_angle) bin_range<true>(g, xc, yc, rpix, inv_da, view0 + vi, k0, k1);
        else bin_range<false>(g, xc, yc, rpix, inv_da, view0 + vi, k0, k1);
        if (!inside || k1 < k0) { k0 = g.nu; k1 = -1; }
    };
    // loads of the CAP bins from `base` of view vi, and their LDS handover
    auto issue = [&](int vi, int base) {
        #pragma unroll
        for (int j = 0; j < kLoads; ++j)
            if (has[j]) rg[j] = ld_at(P, roff[j] + (unsigned)(vi * g.nu + min(base + bj[j], g.nu - 1)) * 16u);
        if (lane < CAP) rr = ld_at(sray_v, (unsigned)(vi * g.nu + min(base + lane, g.nu - 1)) * 16u);
    };
    auto handover = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        #pragma unroll
        for (int j = 0; j < kLoads; ++j)
            if (has[j]) st[j * 64 + lane] = rg[j];
        if (lane < CAP) st[G * CAP + lane] = rr;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    int k0, k1;
    range(0, k0, k1);
    int wlo = wave_min(k0), whi = wave_max(k1);
    if (whi >= wlo) issue(0, wlo);
    for (int vi = 0; vi < nviews; ++vi) {
        const bool any = whi >= wlo;
        const int nch = any ? (whi - wlo) / CAP + 1 : 1;   // an empty window: one chunk of weight-0 reads
        if (any) handover();
        int n0 = g.nu, n1 = -1, nlo = g.nu, nhi = -1;
        if (vi + 1 < nviews) {
            range(vi + 1, n0, n1);
            nlo = wave_min(n0);
            nhi = wave_max(n1);
        }
        const bool pre = nhi >= nlo;
        if (pre && nch == 1) issue(vi + 1, nlo);
        const int view = view0 + vi;
        const bool xv = g.vcls[view] == 0;
        const float lr = xv ? yr : xr;
        const float cell = xv ? (float)cx : (float)cy;
        siddon_gather_staged<NB, G, CAP, false>(st, k0 - wlo, min(k1 - wlo, CAP - 1), lr, cell, acc);
        for (int c = 1; c < nch; ++c) {
            const int base = wlo + c * CAP;
            issue(vi, base);
            handover();
            siddon_gather_staged<NB, G, CAP, false>(st, max(k0 - base, 0), min(k1 - base, CAP - 1), lr, cell, acc);
        }
        if (pre && nch > 1) issue(vi + 1, nlo);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        k0 = n0; k1 = n1; wlo = nlo; whi = nhi;
    }
    if (!inside) return;
    const long plane = (long)g.nx * g.ny, pix = (long)ix * g.ny + iy;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const int s = 4 * (q0 + k);
        float v[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
        if (UPD) {
            const int pxy = qpitch(g.ny), pyx = qpitch(g.nx);
            const float4* qo = u.Qold + (long)(q0 + k) * g.nx * pxy;
            const long c0 = (long)ix * pxy + kQPadL + iy;
            const float4 xc = qo[c0];
            // neighbours for the TV terms (padded lines: iy - 1 and iy + 1 stay inside
            // the line; the border cases below use the pixel value itself, as the
            // plain update)
            float4 xl = xc, xr = xc, yl = xc, yr = xc;
            if (u.beta2 != 0.0f) {
                if (ix > 0) xl = qo[c0 - pxy];
                if (ix < g.nx - 1) xr = qo[c0 + pxy];
                yl = qo[c0 - 1];
                yr = qo[c0 + 1];
            }
            const float4 d4 = u.D4[(long)(q0 + k) * plane + pix];
            const float xs[4] = {xc.x, xc.y, xc.z, xc.w}, xls[4] = {xl.x, xl.y, xl.z, xl.w};
            const float xrs[4] = {xr.x, xr.y, xr.z, xr.w}, yls[4] = {yl.x, yl.y, yl.z, yl.w};
            const float yrs[4] = {yr.x, yr.y, yr.z, yr.w}, ds[4] = {d4.x, d4.y, d4.z, d4.w};
            #pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float x = xs[c];
                const float dx0 = ix == 0 ? x : x - xls[c], dx1 = ix == g.nx - 1 ? x : x - xrs[c];
                const float dy0 = iy == 0 ? x : x - yls[c], dy1 = iy == g.ny - 1 ? x : x - yrs[c];
                float xn = 0.0f;
                if (s + c < ns) {
                    xn = sqs_update_core(x, dx0, dx1, dy0, dy1, v[c], ds[c], u.step, u.beta2, u.eps, u.lb, u.ub,
                                         u.crop);
                    u.Xn[(long)(s + c) * plane + pix] = xn;
                }
                v[c] = xn;
            }
            const float4 f = make_float4(v[0], v[1], v[2], v[3]);
            u.Qxy[(long)(q0 + k) * g.nx * pxy + (long)ix * pxy + kQPadL + iy] = f;
            u.Qyx[(long)(q0 + k) * g.ny * pyx + (long)iy * pyx + kQPadL + ix] = f;
        } else {
            #pragma unroll
            for (int c = 0; c < 4; ++c)
                if (s + c < ns) img[(long)(s + c) * plane + pix] = v[c];
        }
    }
}

template <int BX, int BY, int S, bool UNIT>
__global__ __launch_bounds__(BX * BY) void k_dd_fan_adj(FanK g, const float* __restrict__ prj,
                                                    float* __restrict__ img, int view0, int nviews, int ns,
                                                    int accumulate)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int s0 = blockIdx.z * S;
    const int nsl = min(S, ns - s0);
    if (ix >= g.nx || iy >= g.ny) return;
    const long sstride = (long)nviews * g.nu;
    float acc[S];
    dd_adj_pixel_slices<UNIT, S>(g, prj + s0 * sstride, sstride, nsl, view0, nviews, ix, iy, acc);
    const long plane = (long)g.nx * g.ny;
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (q >= nsl) break;
        float* o = img + (long)(s0 + q) * plane + (long)ix * g.ny + iy;
        *o = accumulate ? *o + acc[q] : acc[q];
    }
}

// PreweightBPKernelX/Y: length factor for the iterative (non-FBP) DD BP.
__global__ void k_dd_preweight(FanK g, const float* __restrict__ prj, float* __restrict__ wprj)
{
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int view = blockIdx.y;
    const int s = blockIdx.z;
    if (iu >= g.nu) return;
    const float ct = g.vcos[view], st = g.vsin[view];
    float f;
    if (g.vcls[view] == 0) f = g.dy / fabsf(ct * g.bcos[iu] + st * g.bsin[iu]);   // cos(t - a)
    else f = g.dx / fabsf(st * g.bcos[iu] - ct * g.bsin[iu]);                      // sin(t - a)
    const long i = ((long)s * g.nview + view) * g.nu + iu;
    wprj[i] = prj[i] * f;
}

// Reference-parity DD BP (DDBPFanKernelX/Y): pixel edges projected through
// atan, average of the (pre-weighted) detector row over the footprint.  One
// lane per pixel of S consecutive slices (grid z counts groups of S): the
// footprint and the FBP weight are computed once for the S slices (their
// quotients through 1-ulp reciprocals).
// The pixel's S slices (value type T: float, or float4 for 4 interleaved
// slices) over every view: rows[q] are the wave-uniform bases of the S
// pre-weighted [view][nu] sinograms.
template <class T, int S>
__device__ __forceinline__ void dd_bp_pixel(const FanK& g, const T* const (&rows)[S], int ix, int iy, int fbp,
                                            T (&acc)[S])
{
    const float x = (ix - 0.5f * (g.nx - 1)) * g.dx, y = (iy - 0.5f * (g.ny - 1)) * g.dy;
    const float inv_da = 1.0f / g.da;
    const float boff = 0.5f * g.nu - g.off_a;
    #pragma unroll
    for (int q = 0; q < S; ++q) acc[q] = vzero<T>();
    for (int view = 0; view < g.nview; ++view) {
        const float ct = g.vcos[view], st = g.vsin[view];
        const float rx = x * ct + y * st, ry = -x * st + y * ct;
        float ex, ey;   // edge offset (rotated frame) of +-half a pixel
        if (g.vcls[view] == 0) { ex = 0.5f * g.dx * ct; ey = -0.5f * g.dx * st; }
        else                   { ex = 0.5f * g.dy * st; ey = 0.5f * g.dy * ct; }
        const float b1 = boff - atanf((rx - ex) * __builtin_amdgcn_rcpf(ry - ey + g.dso)) * inv_da;
        const float b2 = boff - atanf((rx + ex) * __builtin_amdgcn_rcpf(ry + ey + g.dso)) * inv_da;
        float fw = 1.0f;
        if (fbp) {
            const float d = g.dso + ry;
            fw = g.dso * g.dso * __builtin_amdgcn_rcpf(fmaf(rx, rx, d * d));
        }
        seg_acc_rows<T, S>(rows, (unsigned)(view * g.nu) * (unsigned)sizeof(T), g.nu, fminf(b1, b2),
                           fmaxf(b1, b2), fw, acc);
    }
}

template <int BX, int BY, int S>
__global__ __launch_bounds__(BX * BY) void k_dd_fan_bp(FanK g, const float* __restrict__ wprj,
                                                   float* __restrict__ img, int ns, int fbp)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int s0 = blockIdx.z * S;
    const int nsl = min(S, ns - s0);
    if (ix >= g.nx || iy >= g.ny) return;
    const long sstride = (long)g.nview * g.nu;
    const float* rows[S];   // wave-uniform slice bases
    #pragma unroll
    for (int q = 0; q < S; ++q) rows[q] = wprj + (long)(s0 + min(q, nsl - 1)) * sstride;
    float acc[S];
    dd_bp_pixel<float, S>(g, rows, ix, iy, fbp, acc);
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (q >= nsl) break;
        img[(long)(s0 + q) * g.nx * g.ny + (long)ix * g.ny + iy] = acc[q];
    }
}

// The same on 4-slice interleaved sinograms [q][view][nu] (float4): G groups
// per lane; writes the plain images of the real slices.
template <int BX, int BY, int G>
__global__ __launch_bounds__(BX * BY) void k_dd_fan_bp_q(FanK g, const float4* __restrict__ wprj,
                                                     float* __restrict__ img, int ns, int fbp)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int nquads = (ns + 3) / 4;
    const int q0 = blockIdx.z * G;
    const int nq = min(G, nquads - q0);
    if (ix >= g.nx || iy >= g.ny) return;
    const long sstride = (long)g.nview * g.nu;
    const float4* rows[G];
    #pragma unroll
    for (int k = 0; k < G; ++k) rows[k] = wprj + (long)(q0 + min(k, nq - 1)) * sstride;
    float4 acc[G];
    dd_bp_pixel<float4, G>(g, rows, ix, iy, fbp, acc);
    const long plane = (long)g.nx * g.ny, pix = (long)ix * g.ny + iy;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const int s = 4 * (q0 + k);
        const float v[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
        #pragma unroll
        for (int c = 0; c < 4; ++c)
            if (s + c < ns) img[(long)(s + c) * plane + pix] = v[c];
    }
}

// k_dd_fan_bp_q with each wave's bin window staged in LDS (as
// k_siddon_fan_bp_ql): per view the wave's 64 pixels read the bins [wlo, whi]
// of the G groups' pre-weighted rows; those CAP bins from wlo are loaded once
// per wave with coalesced dwordx4 loads, issued one view ahead into
// registers, written to the wave's own LDS rows and averaged from there with
// the arithmetic of dd_bp_pixel.  Views whose window is wider than CAP bins
// read the rows directly.
template <int BX, int BY, int G, int CAP, int NB>
__global__ __launch_bounds__(BX * BY, G >= 8 ? 4 : 1) void k_dd_fan_bp_ql(FanK g, const float4* __restrict__ wprj,
                                                      float* __restrict__ img, int ns, int fbp)
{
    static_assert((BX * BY) % 64 == 0, "whole waves");
    constexpr int kLoads = (G * CAP + 63) / 64;
    __shared__ float4 stage[(BX * BY / 64) * G * CAP];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int lane = tid & 63;
    float4* st = stage + (tid >> 6) * G * CAP;
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const bool inside = ix < g.nx && iy < g.ny;   // lanes outside stage but gather nothing
    const int nquads = (ns + 3) / 4;
    const int q0 = blockIdx.z * G;
    const int nq = min(G, nquads - q0);
    const long sstride = (long)g.nview * g.nu;
    const float4* P = wprj + q0 * sstride;
    const float4* rows[G];
    float4 acc[G];
    #pragma unroll
    for (int q = 0; q < G; ++q) {
        rows[q] = P + (long)min(q, nq - 1) * sstride;
        acc[q] = vzero<float4>();
    }
    // staging load j of this lane: element i = 64 j + lane, row i / CAP, bin
    // i % CAP (LDS slot i)
    int bj[kLoads];
    const float4* src[kLoads];
    bool has[kLoads];
    #pragma unroll
    for (int j = 0; j < kLoads; ++j) {
        const int i = j * 64 + lane, r = i / CAP;
        bj[j] = i % CAP;
        has[j] = r < G;
        src[j] = P + (long)min(r, nq - 1) * sstride;
    }
    const int cx = min(ix, g.nx - 1), cy = min(iy, g.ny - 1);
    const float x = (cx - 0.5f * (g.nx - 1)) * g.dx, y = (cy - 0.5f * (g.ny - 1)) * g.dy;
    const float inv_da = 1.0f / g.da;
    const float boff = 0.5f * g.nu - g.off_a;
    // the footprint [lo, hi] of the pixel in view `view` (dd_bp_pixel), its
    // bins and FBP weight
    auto geom = [&](int view, float& lo, float& hi, float& fw, int& m0, int& m1) {
        const float ct = g.vcos[view], sn = g.vsin[view];
        const float rx = x * ct + y * sn, ry = -x * sn + y * ct;
        float ex, ey;
        if (g.vcls[view] == 0) { ex = 0.5f * g.dx * ct; ey = -0.5f * g.dx * sn; }
        else                   { ex = 0.5f * g.dy * sn; ey = 0.5f * g.dy * ct; }
        const float b1 = boff - atanf((rx - ex) * __builtin_amdgcn_rcpf(ry - ey + g.dso)) * inv_da;
        const float b2 = boff - atanf((rx + ex) * __builtin_amdgcn_rcpf(ry + ey + g.dso)) * inv_da;
        lo = fminf(b1, b2);
        hi = fmaxf(b1, b2);
        const float d = g.dso + ry;
        fw = g.dso * g.dso * __builtin_amdgcn_rcpf(fmaf(rx, rx, d * d));
        seg_avg_bins(g.nu, lo, hi, m0, m1);
        if (!inside) { m0 = g.nu; m1 = -1; }
    };
    float4 rg[kLoads];
    auto issue = [&](int view, int wlo) {
        #pragma unroll
        for (int j = 0; j < kLoads; ++j)
            if (has[j]) rg[j] = ld_at(src[j], (unsigned)(view * g.nu + min(wlo + bj[j], g.nu - 1)) * 16u);
    };
    float lo, hi, fw;
    int m0, m1;
    geom(0, lo, hi, fw, m0, m1);
    int wlo = wave_min(m0), whi = wave_max(m1);
    if (whi >= wlo && whi - wlo < CAP) issue(0, wlo);
    for (int view = 0; view < g.nview; ++view) {
        const bool any = whi >= wlo;            // some lane reads a bin of this view
        const bool staged = whi - wlo < CAP;
        if (any && staged) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            #pragma unroll
            for (int j = 0; j < kLoads; ++j)
                if (has[j]) st[j * 64 + lane] = rg[j];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        float nlo_f = 0.0f, nhi_f = 0.0f, nfw = 1.0f;
        int n0 = g.nu, n1 = -1, nlo = g.nu, nhi = -1;
        if (view + 1 < g.nview) {
            geom(view + 1, nlo_f, nhi_f, nfw, n0, n1);
            nlo = wave_min(n0);
            nhi = wave_max(n1);
            if (nhi >= nlo && nhi - nlo < CAP) issue(view + 1, nlo);
        }
        if (any) {
            const float sc = fbp ? fw : 1.0f;
            if (staged) {
                // lanes outside the image read nothing: their interval is made empty
                seg_acc_staged<G, CAP, NB>(st, wlo, g.nu, inside ? lo : -2.0f, inside ? hi : -1.0f, sc, acc);
            } else if (inside) {
                seg_acc_rows<float4, G>(rows, (unsigned)(view * g.nu) * 16u, g.nu, lo, hi, sc, acc);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        lo = nlo_f; hi = nhi_f; fw = nfw; wlo = nlo; whi = nhi;
    }
    if (!inside) return;
    const long plane = (long)g.nx * g.ny, pix = (long)ix * g.ny + iy;
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= nq) break;
        const int s = 4 * (q0 + k);
        const float v[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
        #pragma unroll
        for (int c = 0; c < 4; ++c)
            if (s + c < ns) img[(long)(s + c) * plane + pix] = v[c];
    }
}

// bpFanKernel3D with isFBP = true (the only mode the reference launches).
// Pixel-driven FBP backprojection (PixelDriven.cu:11-75): one pixel of S
// slices per lane; the view's detector coordinate (an atan), the linear
// interpolation weights and the FBP weight dso^2 / (rx^2 + (dso + ry)^2) are
// computed once and applied to the S slices, each with the arithmetic of a
// one-slice call (a tail group reads its last slice again and discards it).
template <int BX, int BY, int S>
__global__ __launch_bounds__(BX * BY) void k_pixel_fan_bp(FanK g, const float* __restrict__ prj,
                                                      float* __restrict__ img, int ns)
{
    const int iy = blockIdx.x * BX + threadIdx.x;
    const int ix = blockIdx.y * BY + threadIdx.y;
    const int s0 = blockIdx.z * S;
    if (ix >= g.nx || iy >= g.ny) return;
    const int nsl = min(S, ns - s0);
    const float x = (ix - 0.5f * (g.nx - 1)) * g.dx, y = (iy - 0.5f * (g.ny - 1)) * g.dy;
    const float inv_da = 1.0f / g.da;
    const long sstride = (long)g.nview * g.nu;
    float acc[S];
    #pragma unroll
    for (int q = 0; q < S; ++q) acc[q] = 0.0f;
    for (int view = 0; view < g.nview; ++view) {
        const float ct = g.vcos[view], st = g.vsin[view];
        const float rx = x * ct + y * st, ry = -x * st + y * ct;
        const float d = ry + g.dso;
        const float pu = g.cu - g.off_a - atanf(rx * __builtin_amdgcn_rcpf(d)) * inv_da;
        const float f0 = floorf(pu);
        const int i0 = (int)f0;
        const float f = pu - f0;
        const float w = g.dso * g.dso * __builtin_amdgcn_rcpf(fmaf(rx, rx, d * d));
        const bool in0 = i0 >= 0 && i0 < g.nu, in1 = i0 + 1 >= 0 && i0 + 1 < g.nu;
        const float* p = prj + ((long)s0 * g.nview + view) * g.nu;
        #pragma unroll
        for (int q = 0; q < S; ++q) {
            const float* pq = p + (long)min(q, nsl - 1) * sstride;
            const float v0 = in0 ? pq[i0] : 0.0f;
            const float v1 = in1 ? pq[i0 + 1] : 0.0f;
            acc[q] += ((1.0f - f) * v0 + f * v1) * w;
        }
    }
    #pragma unroll
    for (int q = 0; q < S; ++q)
        if (q < nsl) img[(long)(s0 + q) * g.nx * g.ny + (long)ix * g.ny + iy] = acc[q];
}

// ---------------------------------------------------------------------------
// layout helpers
// ---------------------------------------------------------------------------

// q_same[q][i][j] = (in[4q][i][j], .., in[4q+3][i][j]) (zero past nslices) and,
// if q_tr, q_tr[q][j][i] = the same value; 32 x 32 tiles through LDS.  Output
// lines have pitch p_same / p_tr and start off cells into the line.
__global__ __launch_bounds__(256) void k_pack_quads(const float* __restrict__ in, float4* __restrict__ q_same,
                                                    float4* __restrict__ q_tr, int n1, int n2, int ns, int p_same,
                                                    int p_tr, int off)
{
    __shared__ float4 tile[32][33];
    const int j0 = blockIdx.x * 32, i0 = blockIdx.y * 32, q = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const long plane = (long)n1 * n2;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < n1 && j < n2) {
            float v[4];
            #pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int s = 4 * q + c;
                v[c] = s < ns ? in[(long)s * plane + (long)i * n2 + j] : 0.0f;
            }
            const float4 f = make_float4(v[0], v[1], v[2], v[3]);
            tile[r][tx] = f;
            if (q_same != nullptr) q_same[(long)q * n1 * p_same + (long)i * p_same + off + j] = f;
        }
    }
    if (q_tr == nullptr) return;
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + r, i = i0 + tx;
        if (i < n1 && j < n2) q_tr[(long)q * n2 * p_tr + (long)j * p_tr + off + i] = tile[tx][r];
    }
}

__global__ __launch_bounds__(256) void k_permute_rows(const float* __restrict__ in, float* __restrict__ out,
                                                      const int* __restrict__ perm, int rows, int cols)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, b = blockIdx.z;
    if (c >= cols) return;
    out[((long)b * rows + j) * cols + c] = in[((long)b * rows + perm[j]) * cols + c];
}

__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in, float* __restrict__ out,
                                                   int rows, int cols)
{
    __shared__ float tile[32][33];
    const long base = (long)blockIdx.z * rows * cols;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[k][threadIdx.x] = in[base + (long)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[base + (long)c * rows + r] = tile[threadIdx.x][k];
    }
}

__global__ void k_avg_prev(const float* __restrict__ in, float* __restrict__ out, int n0, int n1, int axis)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)n0 * n1;
    const long base = (long)blockIdx.y * per;
    if (i >= per) return;
    const int a = (int)(i / n1), b = (int)(i % n1);
    const float cur = in[base + i];
    float prev = 0.0f;
    if (axis == 0) { if (a > 0) prev = in[base + i - n1]; }
    else           { if (b > 0) prev = in[base + i - 1]; }
    out[base + i] = 0.5f * (prev + cur);
}

__global__ void k_mix_gather(const float* __restrict__ src, float* __restrict__ dst, const float* __restrict__ M,
                             int n, int np, int nc, int nout)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int sd = blockIdx.y;  // (b*nout + o)*nc + c
    if (i >= n) return;
    const int c = sd % nc, bo = sd / nc, o = bo % nout, b = bo / nout;
    float v = 0.0f;
    const float* sp = src + (((long)b * n + i) * np) * nc + c;
    for (int p = 0; p < np; ++p) {
        const float m = M[o * np + p];
        if (m != 0.0f) v = fmaf(m, sp[(long)p * nc], v);
    }
    dst[(long)sd * n + i] = v;
}

__global__ void k_mix_scatter(const float* __restrict__ src, float* __restrict__ dst, const float* __restrict__ M,
                              int n, int np, int nc, int nout, int accumulate)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int dd = blockIdx.y;  // (b*nout + o)*nc + c
    if (i >= n) return;
    const int c = dd % nc, bo = dd / nc, o = bo % nout, b = bo / nout;
    float v = 0.0f;
    for (int p = 0; p < np; ++p) {
        const float m = M[o * np + p];
        if (m != 0.0f) v = fmaf(m, src[((long)(b * np + p) * nc + c) * n + i], v);
    }
    float* d = dst + (((long)b * n + i) * nout + o) * nc + c;
    *d = accumulate ? *d + v : v;
}

__global__ void k_fill(float* __restrict__ p, float v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_axpy(float* __restrict__ y, const float* __restrict__ x, float a, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fmaf(a, x[i], y[i]);
}

} // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void FanTables::build(const FanGeo& g, const float* deg_host, hipStream_t st)
{
    RN_REQUIRE(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.nu > 0 && g.nview > 0 && g.nv > 0, "fan dimensions must be positive");
    RN_REQUIRE(g.dx > 0 && g.dy > 0 && g.dz > 0 && g.da != 0 && g.dv > 0 && g.dso > 0, "fan spacings must be positive");
    RN_REQUIRE(deg_host != nullptr, "angle array is null");
    geo = g;
    deg.assign(deg_host, deg_host + g.nview);
    cls.resize(g.nview);
    const int nu = g.nu, nview = g.nview;
    // layout: vcos[nview] vsin[nview] vcls[nview] esin[nu+1] ecos[nu+1] bsin[nu] bcos[nu]
    const size_t nf = 3 * (size_t)nview + 2 * (size_t)(nu + 1) + 2 * (size_t)nu;
    host_.assign(nf, 0.0f);
    float* vcos = host_.data();
    float* vsin = vcos + nview;
    int* vcls = reinterpret_cast<int*>(vsin + nview);
    float* esin = vsin + 2 * nview;
    float* ecos = esin + (nu + 1);
    float* bsin = ecos + (nu + 1);
    float* bcos = bsin + nu;
    for (int v = 0; v < nview; v++) {
        const float d = deg_host[v];
        vcos[v] = (float)std::cos((double)d);
        vsin[v] = (float)std::sin((double)d);
        // the float32 host decision of DistanceDriven.cu:302 (glibc cosf/sinf)
        cls[v] = std::fabs(cosf(d)) > std::fabs(sinf(d)) ? 0 : 1;
        vcls[v] = cls[v];
    }
    const double da = (double)g.da, cu = (nu - 1) / 2.0, off = (double)g.off_a;
    for (int e = 0; e <= nu; e++) {
        const double a = (-(e - cu) - off) * da + da / 2.0;
        esin[e] = (float)std::sin(a);
        ecos[e] = (float)std::cos(a);
    }
    for (int k = 0; k < nu; k++) {
        const double a = (-(k - cu) - off) * da;
        bsin[k] = (float)std::sin(a);
        bcos[k] = (float)std::cos(a);
    }
    buf_.alloc(nf * sizeof(float));
    RN_HIP_CHECK(hipMemcpyAsync(buf_.get(), host_.data(), nf * sizeof(float), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    float* d = buf_.as<float>();
    k.nx = g.nx; k.ny = g.ny; k.nu = nu; k.nview = nview;
    k.dx = g.dx; k.dy = g.dy; k.da = g.da; k.off_a = g.off_a; k.dso = g.dso; k.dsd = g.dsd;
    k.cu = 0.5f * (nu - 1);
    k.vcos = d; k.vsin = d + nview; k.vcls = reinterpret_cast<const int*>(d + 2 * nview);
    k.esin = d + 3 * nview; k.ecos = k.esin + (nu + 1);
    k.bsin = k.ecos + (nu + 1); k.bcos = k.bsin + nu;

    // Siddon ray lines per (view, bin), FP64 then rounded
    std::vector<float4> ray((size_t)nview * nu);
    for (int v = 0; v < nview; v++) {
        const double t = (double)deg_host[v];
        for (int kb = 0; kb < nu; kb++) {
            const double a = (-(kb - cu) - off) * da;
            const double c = std::cos(a - t), sn = std::sin(a - t), sa = std::sin(a);
            double A, B, L;
            if (cls[v] == 0) {
                A = 0.5 * g.nx + ((double)g.dso / g.dx) * sa / c;
                B = ((double)g.dy / g.dx) * sn / c;
                L = (double)g.dy / std::fabs(c);
            } else {
                A = 0.5 * g.ny - ((double)g.dso / g.dy) * sa / sn;
                B = ((double)g.dx / g.dy) * c / sn;
                L = (double)g.dx / std::fabs(sn);
            }
            ray[(size_t)v * nu + kb] = make_float4((float)A, (float)B, (float)L, 0.0f);
        }
    }
    // per view: 1 if every ray's line interval has a positive FP32 width on
    // every image line, i.e. |B| exceeds the rounding of the end points
    // A + B (j - n/2) (a few ulp of their largest magnitude) by a wide margin
    std::vector<int> wide(nview, 1);
    for (int v = 0; v < nview; v++) {
        const int nl = cls[v] == 0 ? g.ny : g.nx;
        for (int kb = 0; kb < nu && wide[v]; kb++) {
            const float4 r = ray[(size_t)v * nu + kb];
            const double emax = std::fabs((double)r.x) + std::fabs((double)r.y) * (0.5 * nl + 1.0) + 1.0;
            if (!(std::fabs((double)r.y) > 1e-6 * emax)) wide[v] = 0;
        }
    }
    const size_t ray_bytes = ray.size() * sizeof(float4);
    ray_.alloc(ray_bytes + wide.size() * sizeof(int));
    RN_HIP_CHECK(hipMemcpyAsync(ray_.get(), ray.data(), ray_bytes, hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(ray_.get()) + ray_bytes, wide.data(), wide.size() * sizeof(int),
                                hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    k.sray = reinterpret_cast<const float4*>(ray_.get());
    // DD bin-edge rays per (view, bin), FP64 then rounded: edge e has angle
    // (-(e - cu) - off_a) da + da / 2, so bin k lies between edges k and k + 1
    // (angles a_k + da/2 and a_k - da/2)
    std::vector<double> eA((size_t)nu + 1), eB((size_t)nu + 1);
    std::vector<float4> dray((size_t)nview * nu);
    for (int v = 0; v < nview; v++) {
        const double t = (double)deg_host[v];
        for (int e = 0; e <= nu; e++) {
            const double a = (-(e - cu) - off) * da + da / 2.0;
            const double c = std::cos(a - t), sn = std::sin(a - t), sa = std::sin(a);
            double A, B;
            if (cls[v] == 0) {
                A = 0.5 * g.nx + ((double)g.dso / g.dx) * sa / c;
                B = ((double)g.dy / g.dx) * sn / c;
            } else {
                A = 0.5 * g.ny - ((double)g.dso / g.dy) * sa / sn;
                B = ((double)g.dx / g.dy) * c / sn;
            }
            eA[e] = A;
            eB[e] = B;
        }
        for (int kb = 0; kb < nu; kb++)
            dray[(size_t)v * nu + kb] = make_float4((float)eA[kb], (float)eB[kb], (float)eA[kb + 1], (float)eB[kb + 1]);
    }
    dray_.alloc(dray.size() * sizeof(float4));
    RN_HIP_CHECK(hipMemcpyAsync(dray_.get(), dray.data(), dray.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    k.dray = reinterpret_cast<const float4*>(dray_.get());
    // pixel centres lie within R of the centre: |rx| / (dso - |ry|) <= R / (dso - R) < 1 if R < 0.45 dso
    const double R = 0.5 * std::sqrt((double)g.nx * g.dx * g.nx * g.dx + (double)g.ny * g.dy * g.ny * g.dy);
    k.unit_angle = R < 0.45 * (double)g.dso ? 1 : 0;
    k.vwide = reinterpret_cast<const int*>(static_cast<const char*>(ray_.get()) + ray_bytes);
}

bool FanTables::matches(const FanGeo& g, const float* deg_host) const
{
    if (buf_.get() == nullptr) return false;
    if (std::memcmp(&g, &geo, sizeof(FanGeo)) != 0) return false;
    return std::memcmp(deg.data(), deg_host, sizeof(float) * g.nview) == 0;
}

// ---- launch shapes: candidates for the autotuner (identical results for every shape) ----

namespace {

// MI355X: 8 XCDs, each with its own L2
constexpr int kXcds = 8;

template <int BLK, int XCD, int S>
void go_dd_fp(const FanK& g, const float* RX, const float* RY, float* prj, int ns, int view0, int nviews,
              hipStream_t st)
{
    const int ngroups = cdiv(ns, S);
    const long nblk = (long)cdiv(g.nu, BLK) * nviews * ngroups;
    if (XCD > 0 && nblk % XCD == 0 && nblk <= 0x7fffffffL) {
        hipLaunchKernelGGL((k_dd_fan_fp<BLK, XCD, S>), dim3((unsigned)nblk), dim3(BLK), 0, st, g, RX, RY, prj, view0,
                           nviews, ns);
    } else {
        hipLaunchKernelGGL((k_dd_fan_fp<BLK, 0, S>), dim3(cdiv(g.nu, BLK), nviews, ngroups), dim3(BLK), 0, st, g, RX,
                           RY, prj, view0, nviews, ns);
    }
}

template <int BLK, int XCD, int G, int H = 1>
void go_dd_fp_q(const FanK& g, const float4* QX, const float4* QY, float* prj, int ns, int view0, int nviews,
                hipStream_t st)
{
    const int ngroups = cdiv(cdiv(ns, 4), G);
    const long nblk = (long)cdiv(g.nu, BLK / H) * nviews * ngroups;
    if (XCD > 0 && nblk % XCD == 0 && nblk <= 0x7fffffffL) {
        hipLaunchKernelGGL((k_dd_fan_fp_q<BLK, XCD, G, H>), dim3((unsigned)nblk), dim3(BLK), 0, st, g, QX, QY, prj,
                           view0, nviews, ns);
    } else {
        hipLaunchKernelGGL((k_dd_fan_fp_q<BLK, 0, G, H>), dim3(cdiv(g.nu, BLK / H), nviews, ngroups), dim3(BLK), 0, st,
                           g, QX, QY, prj, view0, nviews, ns);
    }
}

template <int BLK, int XCD, int S>
void go_siddon_fp(const FanK& g, const float* Pyx, const float* Pxy, float* prj, int ns, int view0, int nviews,
                  const float* meas, const float* w, hipStream_t st)
{
    const int ngroups = cdiv(ns, S);
    const long nblk = (long)cdiv(g.nu, BLK) * nviews * ngroups;
    if (XCD > 0 && nblk % XCD == 0 && nblk <= 0x7fffffffL) {
        hipLaunchKernelGGL((k_siddon_fan_fp<float, BLK, XCD, S>), dim3((unsigned)nblk), dim3(BLK), 0, st, g, Pyx,
                           Pxy, prj, view0, nviews, ns, meas, w, g.nview);
    } else {
        hipLaunchKernelGGL((k_siddon_fan_fp<float, BLK, 0, S>), dim3(cdiv(g.nu, BLK), nviews, ngroups), dim3(BLK), 0,
                           st, g, Pyx, Pxy, prj, view0, nviews, ns, meas, w, g.nview);
    }
}

template <int BX, int BY, int NB, int S>
void go_siddon_bp(const FanK& g, const float* prj, float* img, int ns, int view0, int nviews, int acc,
                  hipStream_t st)
{
    hipLaunchKernelGGL((k_siddon_fan_bp<float, BX, BY, NB, S>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(ns, S)),
                       dim3(BX, BY), 0, st, g, prj, img, view0, nviews, ns, acc);
}

template <int BX, int BY, int S>
void go_dd_adj(const FanK& g, const float* prj, float* img, int ns, int view0, int nviews, int acc, hipStream_t st)
{
    const dim3 grd(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(ns, S));
    if (g.unit_angle)
        hipLaunchKernelGGL((k_dd_fan_adj<BX, BY, S, true>), grd, dim3(BX, BY), 0, st, g, prj, img, view0, nviews, ns,
                           acc);
    else
        hipLaunchKernelGGL((k_dd_fan_adj<BX, BY, S, false>), grd, dim3(BX, BY), 0, st, g, prj, img, view0, nviews,
                           ns, acc);
}

template <int BX, int BY, int S>
void go_dd_bp(const FanK& g, const float* wprj, float* img, int ns, int fbp, hipStream_t st)
{
    hipLaunchKernelGGL((k_dd_fan_bp<BX, BY, S>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(ns, S)), dim3(BX, BY), 0,
                       st, g, wprj, img, ns, fbp);
}

using DdFpFn = void (*)(const FanK&, const float*, const float*, float*, int, int, int, hipStream_t);
using SidFpFn = void (*)(const FanK&, const float*, const float*, float*, int, int, int, const float*, const float*,
                         hipStream_t);
using GatherFn = void (*)(const FanK&, const float*, float*, int, int, int, int, hipStream_t);
using DdBpFn = void (*)(const FanK&, const float*, float*, int, int, hipStream_t);

// 1-D blocks along the detector; 2-D pixel tiles (x lanes along y, the contiguous image axis)
const DdFpFn kDdFp[] = {go_dd_fp<128, kXcds, 1>, go_dd_fp<64, kXcds, 1>, go_dd_fp<256, kXcds, 1>,
                        go_dd_fp<128, 0, 1>,     go_dd_fp<64, 0, 1>,     go_dd_fp<128, kXcds, 2>,
                        go_dd_fp<64, kXcds, 2>,  go_dd_fp<128, 0, 2>};
constexpr int kNDdFp = sizeof(kDdFp) / sizeof(kDdFp[0]);
using DdFpQFn = void (*)(const FanK&, const float4*, const float4*, float*, int, int, int, hipStream_t);
const DdFpQFn kDdFpQ[] = {go_dd_fp_q<64, kXcds, 1, 2>,  go_dd_fp_q<128, kXcds, 1, 2>, go_dd_fp_q<256, kXcds, 1, 2>,
                          go_dd_fp_q<64, kXcds, 1, 4>,  go_dd_fp_q<128, kXcds, 1>,    go_dd_fp_q<64, kXcds, 1>,
                          go_dd_fp_q<256, kXcds, 1>,    go_dd_fp_q<128, 0, 1>,        go_dd_fp_q<64, 0, 1>,
                          go_dd_fp_q<128, kXcds, 2>,    go_dd_fp_q<64, kXcds, 2, 2>,  go_dd_fp_q<128, 0, 2>};
constexpr int kNDdFpQ = sizeof(kDdFpQ) / sizeof(kDdFpQ[0]);
// S: slices per lane (1 = one slice per lane; larger S shares the geometry work)
const SidFpFn kSidFp[] = {go_siddon_fp<128, 0, 1>,     go_siddon_fp<64, 0, 1>,     go_siddon_fp<256, 0, 1>,
                          go_siddon_fp<128, kXcds, 1>, go_siddon_fp<64, kXcds, 1>, go_siddon_fp<256, kXcds, 1>,
                          go_siddon_fp<128, 0, 2>,     go_siddon_fp<128, kXcds, 2>, go_siddon_fp<64, 0, 4>,
                          go_siddon_fp<128, 0, 4>,     go_siddon_fp<64, kXcds, 4>, go_siddon_fp<128, kXcds, 4>,
                          go_siddon_fp<256, kXcds, 4>, go_siddon_fp<128, kXcds, 8>};
const GatherFn kSidBp[] = {go_siddon_bp<16, 16, 3, 1>, go_siddon_bp<8, 8, 3, 1>,  go_siddon_bp<16, 16, 2, 1>,
                           go_siddon_bp<8, 8, 2, 1>,   go_siddon_bp<16, 16, 4, 1>, go_siddon_bp<8, 8, 4, 1>,
                           go_siddon_bp<8, 8, 2, 2>,   go_siddon_bp<8, 8, 2, 4>,   go_siddon_bp<16, 16, 2, 4>,
                           go_siddon_bp<8, 8, 3, 4>,   go_siddon_bp<8, 8, 2, 8>,   go_siddon_bp<16, 16, 2, 8>,
                           go_siddon_bp<8, 8, 3, 8>,   go_siddon_bp<8, 8, 2, 16>,  go_siddon_bp<16, 16, 2, 16>};
constexpr int kNSidFp = sizeof(kSidFp) / sizeof(kSidFp[0]);
constexpr int kNSidBp = sizeof(kSidBp) / sizeof(kSidBp[0]);
const GatherFn kDdAdj[] = {go_dd_adj<16, 16, 1>, go_dd_adj<8, 32, 1>, go_dd_adj<32, 8, 1>, go_dd_adj<8, 8, 1>,
                           go_dd_adj<16, 16, 2>, go_dd_adj<8, 8, 2>,  go_dd_adj<16, 16, 4>, go_dd_adj<8, 8, 4>};
constexpr int kNDdAdj = sizeof(kDdAdj) / sizeof(kDdAdj[0]);
const DdBpFn kDdBp[] = {go_dd_bp<16, 16, 1>, go_dd_bp<8, 32, 1>, go_dd_bp<32, 8, 1>, go_dd_bp<8, 8, 1>,
                       go_dd_bp<8, 8, 4>,   go_dd_bp<16, 16, 4>, go_dd_bp<8, 8, 8>, go_dd_bp<16, 16, 8>,
                       go_dd_bp<8, 8, 16>};
constexpr int kNDdBp = sizeof(kDdBp) / sizeof(kDdBp[0]);

template <int BX, int BY, int G>
void go_dd_bp_q(const FanK& g, const float4* wprj, float* img, int ns, int fbp, hipStream_t st)
{
    hipLaunchKernelGGL((k_dd_fan_bp_q<BX, BY, G>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(cdiv(ns, 4), G)),
                       dim3(BX, BY), 0, st, g, wprj, img, ns, fbp);
}
template <int BX, int BY, int G, int CAP, int NB = 2>
void go_dd_bp_ql(const FanK& g, const float4* wprj, float* img, int ns, int fbp, hipStream_t st)
{
    hipLaunchKernelGGL((k_dd_fan_bp_ql<BX, BY, G, CAP, NB>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(cdiv(ns, 4), G)),
                       dim3(BX, BY), 0, st, g, wprj, img, ns, fbp);
}
using DdBpQFn = void (*)(const FanK&, const float4*, float*, int, int, hipStream_t);
// LDS-staged shapes only with 8 slice groups per lane: staged 4- and 2-group
// kernels timed within a few per cent of the direct ones back to back, so the
// tuner took them at times, but after the preweighting and packing launches
// of a backward call they ran 0.496 and 0.639 ms against 0.410 ms (32 slices
// of 256^2) and 3.4 and 5.3 ms against 2.4 ms (32 slices of 512^2)
const DdBpQFn kDdBpQ[] = {go_dd_bp_ql<8, 8, 8, 24>, go_dd_bp_ql<8, 8, 8, 32>, go_dd_bp_ql<8, 32, 8, 32>,
                          go_dd_bp_ql<8, 32, 8, 24>, go_dd_bp_q<8, 8, 4>,       go_dd_bp_q<16, 16, 2>,
                          go_dd_bp_q<8, 8, 2>,       go_dd_bp_q<16, 16, 1>,     go_dd_bp_q<8, 8, 1>};
constexpr int kNDdBpQ = sizeof(kDdBpQ) / sizeof(kDdBpQ[0]);

// The Siddon and DD gather kernels address one slice's image plane, ray table
// and sinogram rows with 32-bit byte offsets.
void require_offsets32(const FanK& g, int nviews, size_t elem)
{
    const long lim = 0x7fffffffL;
    RN_REQUIRE((long)g.nx * g.ny * (long)elem < lim && (long)g.nview * g.nu * 16L < lim &&
                   (long)nviews * g.nu * (long)elem < lim,
               "image plane, ray table or sinogram of one slice exceeds 2 GiB");
}

template <int BLK, int XCD, int G, int H = 1>
void go_siddon_fp_q(const FanK& g, const float4* Qyx, const float4* Qxy, float4* res, int nquads, int view0,
                    int nviews, const float4* meas, const float4* w, int res_vstride, int res_voff, hipStream_t st)
{
    const int ngroups = cdiv(nquads, G);
    const long nblk = (long)cdiv(g.nu, BLK / H) * nviews * ngroups;
    if (XCD > 0 && nblk % XCD == 0 && nblk <= 0x7fffffffL) {
        hipLaunchKernelGGL((k_siddon_fan_fp_q<BLK, XCD, G, H>), dim3((unsigned)nblk), dim3(BLK), 0, st, g, Qyx, Qxy,
                           res, view0, nviews, nquads, meas, w, g.nview, res_vstride, res_voff);
    } else {
        hipLaunchKernelGGL((k_siddon_fan_fp_q<BLK, 0, G, H>), dim3(cdiv(g.nu, BLK / H), nviews, ngroups), dim3(BLK), 0, st,
                           g, Qyx, Qxy, res, view0, nviews, nquads, meas, w, g.nview, res_vstride, res_voff);
    }
}

template <int BX, int BY, int NB, int G>
void go_siddon_bp_q(const FanK& g, const float4* prj, float* img, int ns, int view0, int nviews, hipStream_t st)
{
    hipLaunchKernelGGL((k_siddon_fan_bp_q<BX, BY, NB, G>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(cdiv(ns, 4), G)),
                       dim3(BX, BY), 0, st, g, prj, img, view0, nviews, ns);
}

template <int BX, int BY, int NB, int G, int CAP>
void go_siddon_bp_ql(const FanK& g, const float4* prj, float* img, int ns, int view0, int nviews, hipStream_t st)
{
    hipLaunchKernelGGL((k_siddon_fan_bp_ql<BX, BY, NB, G, CAP>),
                       dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(cdiv(ns, 4), G)), dim3(BX, BY), 0, st, g, prj, img,
                       view0, nviews, ns, SqsUpd{});
}

template <int BX, int BY, int NB, int G, int CAP>
void go_siddon_bp_qlu(const FanK& g, const float4* prj, int ns, int view0, int nviews, const SqsUpd& u,
                      hipStream_t st)
{
    hipLaunchKernelGGL((k_siddon_fan_bp_ql<BX, BY, NB, G, CAP, true>),
                       dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(cdiv(ns, 4), G)), dim3(BX, BY), 0, st, g, prj,
                       nullptr, view0, nviews, ns, u);
}

using SidFpQFn = void (*)(const FanK&, const float4*, const float4*, float4*, int, int, int, const float4*,
                          const float4*, int, int, hipStream_t);
using SidBpQFn = void (*)(const FanK&, const float4*, float*, int, int, int, hipStream_t);
// G: groups of 4 slices per lane; H: parts of the line range per ray (the
// split shortens the waves of a launch that would leave a partial second
// round on the chip); candidates in order of preference (the tuner keeps the
// first within 3 % of the fastest)
const SidFpQFn kSidFpQ[] = {go_siddon_fp_q<64, kXcds, 1, 2>,  go_siddon_fp_q<128, kXcds, 1, 2>,
                            go_siddon_fp_q<256, kXcds, 1, 2>, go_siddon_fp_q<64, kXcds, 1, 4>,
                            go_siddon_fp_q<128, kXcds, 1>,    go_siddon_fp_q<64, kXcds, 1>,
                            go_siddon_fp_q<256, kXcds, 1>,    go_siddon_fp_q<128, 0, 1>,
                            go_siddon_fp_q<64, 0, 1>,         go_siddon_fp_q<128, kXcds, 2>,
                            go_siddon_fp_q<128, 0, 2>};
const SidBpQFn kSidBpQ[] = {go_siddon_bp_ql<8, 8, 2, 8, 24>,  go_siddon_bp_ql<8, 32, 2, 8, 24>,
                            go_siddon_bp_ql<8, 8, 2, 8, 32>,  go_siddon_bp_ql<8, 8, 2, 8, 21>,
                            go_siddon_bp_ql<8, 8, 2, 8, 16>,  go_siddon_bp_ql<8, 8, 2, 4, 24>,
                            go_siddon_bp_ql<8, 8, 2, 2, 32>,
                            go_siddon_bp_q<16, 16, 2, 8>,     go_siddon_bp_q<8, 8, 2, 8>,
                            go_siddon_bp_q<16, 16, 2, 4>,     go_siddon_bp_q<8, 8, 2, 4>,
                            go_siddon_bp_q<8, 8, 2, 2>,       go_siddon_bp_q<16, 16, 2, 2>,
                            go_siddon_bp_q<8, 8, 2, 1>,       go_siddon_bp_q<16, 16, 2, 1>,
                            go_siddon_bp_q<8, 8, 3, 4>,       go_siddon_bp_q<8, 8, 3, 2>};
constexpr int kNSidFpQ = sizeof(kSidFpQ) / sizeof(kSidFpQ[0]);
constexpr int kNSidBpQ = sizeof(kSidBpQ) / sizeof(kSidBpQ[0]);
constexpr int kNSidBpStaged = 7;   // the first seven are k_siddon_fan_bp_ql
using SidBpQUFn = void (*)(const FanK&, const float4*, int, int, int, const SqsUpd&, hipStream_t);
// the staged shapes of kSidBpQ, in the same order, with the OS-SQS update in the epilogue
const SidBpQUFn kSidBpQU[kNSidBpStaged] = {
    go_siddon_bp_qlu<8, 8, 2, 8, 24>, go_siddon_bp_qlu<8, 32, 2, 8, 24>, go_siddon_bp_qlu<8, 8, 2, 8, 32>,
    go_siddon_bp_qlu<8, 8, 2, 8, 21>, go_siddon_bp_qlu<8, 8, 2, 8, 16>, go_siddon_bp_qlu<8, 8, 2, 4, 24>,
    go_siddon_bp_qlu<8, 8, 2, 2, 32>};

} // namespace

void launch_dd_fan_fp(const FanK& g, const float* RX, const float* RY, float* prj_vu,
                      int nslices, int view0, int nviews, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    const int c = tune_choose({TK_DD_FP, g.nx, g.ny, g.nu, nviews, nslices}, kNDdFp,
                              [&](int i) { kDdFp[i](g, RX, RY, prj_vu, nslices, view0, nviews, st); }, st, true);
    kDdFp[c](g, RX, RY, prj_vu, nslices, view0, nviews, st);
    check_launch("dd_fan_fp");
}

void launch_dd_fan_fp_q(const FanK& g, const float4* QX, const float4* QY, float* prj_vu, int nslices, int view0,
                        int nviews, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    RN_REQUIRE((long)std::max(g.nx * qpitch(g.ny), g.ny * qpitch(g.nx)) * 16L < 0x7fffffffL,
               "padded image plane of one slice group exceeds 2 GiB");
    const int c = tune_choose({TK_DD_FP_Q, g.nx, g.ny, g.nu, nviews, nslices}, kNDdFpQ,
                              [&](int i) { kDdFpQ[i](g, QX, QY, prj_vu, nslices, view0, nviews, st); }, st, true);
    kDdFpQ[c](g, QX, QY, prj_vu, nslices, view0, nviews, st);
    check_launch("dd_fan_fp_q");
}

void launch_siddon_fan_fp(const FanK& g, const float* Pyx, const float* Pxy, float* prj_vu,
                          int nslices, int view0, int nviews, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    const int c = tune_choose({TK_SIDDON_FP, g.nx, g.ny, g.nu, nviews, nslices}, kNSidFp, [&](int i) {
        kSidFp[i](g, Pyx, Pxy, prj_vu, nslices, view0, nviews, nullptr, nullptr, st);
    }, st, true);
    kSidFp[c](g, Pyx, Pxy, prj_vu, nslices, view0, nviews, nullptr, nullptr, st);
    check_launch("siddon_fan_fp");
}

void launch_siddon_fan_fp_residual(const FanK& g, const float* Pyx, const float* Pxy, float* res_vu,
                                   int nslices, int view0, int nviews, const float* meas_vu, const float* w_vu,
                                   hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    const int c = tune_choose({TK_SIDDON_FP, g.nx, g.ny, g.nu, nviews, nslices}, kNSidFp, [&](int i) {
        kSidFp[i](g, Pyx, Pxy, res_vu, nslices, view0, nviews, meas_vu, w_vu, st);
    }, st, true);
    kSidFp[c](g, Pyx, Pxy, res_vu, nslices, view0, nviews, meas_vu, w_vu, st);
    check_launch("siddon_fan_fp_residual");
}

void launch_siddon_fan_fp_residual_q(const FanK& g, const float4* Qyx, const float4* Qxy, float4* res_q,
                                     int nslices, int view0, int nviews, const float4* meas_q, const float4* w_q,
                                     int res_vstride, int res_voff, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    RN_REQUIRE(res_voff >= 0 && res_voff + nviews <= res_vstride, "residual view range out of its rows");
    require_offsets32(g, res_vstride, sizeof(float4));
    RN_REQUIRE((long)std::max(g.nx * qpitch(g.ny), g.ny * qpitch(g.nx)) * 16L < 0x7fffffffL,
               "padded image plane of one slice group exceeds 2 GiB");
    const int nq = cdiv(nslices, 4);
    const int c = tune_choose({TK_SIDDON_FP_Q, g.nx, g.ny, g.nu, nviews, nslices}, kNSidFpQ, [&](int i) {
        kSidFpQ[i](g, Qyx, Qxy, res_q, nq, view0, nviews, meas_q, w_q, res_vstride, res_voff, st);
    }, st, true);
    kSidFpQ[c](g, Qyx, Qxy, res_q, nq, view0, nviews, meas_q, w_q, res_vstride, res_voff, st);
    check_launch("siddon_fan_fp_residual_q");
}

void launch_siddon_fan_bp_q(const FanK& g, const float4* prj_q, float* img_xy, int nslices, int view0,
                            int nviews, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(float4));
    // the LDS-staged candidates address 8 groups' rows with 32-bit offsets
    const int first = 9L * nviews * g.nu * 16L < 0x7fffffffL ? 0 : kNSidBpStaged;
    const int c = first + tune_choose({TK_SIDDON_BP_Q, g.nx, g.ny, g.nu, nviews, nslices}, kNSidBpQ - first,
                                      [&](int i) { kSidBpQ[first + i](g, prj_q, img_xy, nslices, view0, nviews, st); },
                                      st, true);
    kSidBpQ[c](g, prj_q, img_xy, nslices, view0, nviews, st);
    check_launch("siddon_fan_bp_q");
}

bool launch_siddon_fan_bp_q_update(const FanK& g, const float4* prj_q, float* img_xy, int nslices, int view0,
                                   int nviews, const SqsUpd& u, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return false;
    require_offsets32(g, nviews, sizeof(float4));
    const int first = 9L * nviews * g.nu * 16L < 0x7fffffffL ? 0 : kNSidBpStaged;
    // the launch shape is tuned on the plain backprojection (img_xy as its output)
    const int c = first + tune_choose({TK_SIDDON_BP_Q, g.nx, g.ny, g.nu, nviews, nslices}, kNSidBpQ - first,
                                      [&](int i) { kSidBpQ[first + i](g, prj_q, img_xy, nslices, view0, nviews, st); },
                                      st, true);
    if (c < kNSidBpStaged) {
        kSidBpQU[c](g, prj_q, nslices, view0, nviews, u, st);
        check_launch("siddon_fan_bp_q_update");
        return true;
    }
    kSidBpQ[c](g, prj_q, img_xy, nslices, view0, nviews, st);
    check_launch("siddon_fan_bp_q");
    return false;
}

void launch_siddon_fan_fp_f64(const FanK& g, const double* Pyx, const double* Pxy, double* prj_vu,
                              int nslices, int view0, int nviews, hipStream_t st)
{
    if (nslices <= 0 || nviews <= 0) return;
    require_offsets32(g, nviews, sizeof(double));
    dim3 blk(128), grd(cdiv(g.nu, 128), nviews, nslices);
    hipLaunchKernelGGL((k_siddon_fan_fp<double, 128, 0, 1>), grd, blk, 0, st, g, Pyx, Pxy, prj_vu, view0, nviews,
                       nslices, (const float*)nullptr, (const float*)nullptr, 0);
    check_launch("siddon_fan_fp_f64");
}

void launch_siddon_fan_bp_f64(const FanK& g, const double* prj_vu, double* img_xy, int nslices,
                              int view0, int nviews, hipStream_t st)
{
    if (nslices <= 0) return;
    require_offsets32(g, nviews, sizeof(double));
    dim3 blk(16, 16), grd(cdiv(g.ny, 16), cdiv(g.nx, 16), nslices);
    hipLaunchKernelGGL((k_siddon_fan_bp<double, 16, 16, 4, 1>), grd, blk, 0, st, g, prj_vu, img_xy, view0, nviews,
                       nslices, 0);
    check_launch("siddon_fan_bp_f64");
}

void launch_siddon_fan_bp(const FanK& g, const float* prj_vu, float* img_xy, int nslices,
                          int view0, int nviews, bool accumulate, hipStream_t st)
{
    if (nslices <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    const int c = tune_choose({TK_SIDDON_BP, g.nx, g.ny, g.nu, nviews, nslices}, kNSidBp, [&](int i) {
        kSidBp[i](g, prj_vu, img_xy, nslices, view0, nviews, 0, st);
    }, st, !accumulate);
    kSidBp[c](g, prj_vu, img_xy, nslices, view0, nviews, accumulate ? 1 : 0, st);
    check_launch("siddon_fan_bp");
}

void launch_dd_fan_adj(const FanK& g, const float* prj_vu, float* img_xy, int nslices,
                       int view0, int nviews, bool accumulate, hipStream_t st)
{
    if (nslices <= 0) return;
    require_offsets32(g, nviews, sizeof(float));
    const int c = tune_choose({TK_DD_ADJ, g.nx, g.ny, g.nu, nviews, nslices}, kNDdAdj, [&](int i) {
        kDdAdj[i](g, prj_vu, img_xy, nslices, view0, nviews, 0, st);
    }, st, !accumulate);
    kDdAdj[c](g, prj_vu, img_xy, nslices, view0, nviews, accumulate ? 1 : 0, st);
    check_launch("dd_fan_adj");
}

void launch_dd_preweight(const FanK& g, const float* prj_vu, float* wprj_vu, int nslices, hipStream_t st)
{
    if (nslices <= 0) return;
    dim3 blk(256), grd(cdiv(g.nu, 256), g.nview, nslices);
    hipLaunchKernelGGL(k_dd_preweight, grd, blk, 0, st, g, prj_vu, wprj_vu);
    check_launch("dd_preweight");
}

void launch_dd_fan_bp(const FanK& g, const float* wprj_vu, float* img_xy, int nslices, bool fbp, hipStream_t st)
{
    if (nslices <= 0) return;
    require_offsets32(g, g.nview, sizeof(float));
    const int c = tune_choose({fbp ? TK_DD_FBP : TK_DD_BP, g.nx, g.ny, g.nu, g.nview, nslices}, kNDdBp,
                              [&](int i) { kDdBp[i](g, wprj_vu, img_xy, nslices, fbp ? 1 : 0, st); }, st, true);
    kDdBp[c](g, wprj_vu, img_xy, nslices, fbp ? 1 : 0, st);
    check_launch("dd_fan_bp");
}

void launch_dd_fan_bp_q(const FanK& g, const float4* wprj_q, float* img_xy, int nslices, bool fbp, hipStream_t st)
{
    if (nslices <= 0) return;
    require_offsets32(g, g.nview, sizeof(float4));
    const int c = tune_choose({fbp ? TK_DD_FBP_Q : TK_DD_BP_Q, g.nx, g.ny, g.nu, g.nview, nslices}, kNDdBpQ,
                              [&](int i) { kDdBpQ[i](g, wprj_q, img_xy, nslices, fbp ? 1 : 0, st); }, st, true);
    kDdBpQ[c](g, wprj_q, img_xy, nslices, fbp ? 1 : 0, st);
    check_launch("dd_fan_bp_q");
}

template <int BX, int BY, int S>
void go_pixel_bp(const FanK& g, const float* prj_vu, float* img_xy, int nslices, hipStream_t st)
{
    hipLaunchKernelGGL((k_pixel_fan_bp<BX, BY, S>), dim3(cdiv(g.ny, BX), cdiv(g.nx, BY), cdiv(nslices, S)),
                       dim3(BX, BY), 0, st, g, prj_vu, img_xy, nslices);
}

using PixBpFn = void (*)(const FanK&, const float*, float*, int, hipStream_t);
const PixBpFn kPixBp[] = {go_pixel_bp<16, 16, 1>, go_pixel_bp<8, 8, 1>,   go_pixel_bp<16, 16, 2>,
                          go_pixel_bp<16, 16, 4>, go_pixel_bp<8, 8, 4>,   go_pixel_bp<16, 16, 8>,
                          go_pixel_bp<8, 32, 8>,  go_pixel_bp<16, 16, 16>};
constexpr int kNPixBp = sizeof(kPixBp) / sizeof(kPixBp[0]);

void launch_pixel_fan_bp(const FanK& g, const float* prj_vu, float* img_xy, int nslices, hipStream_t st)
{
    if (nslices <= 0) return;
    RN_REQUIRE(cdiv(g.nx, 8) <= 65535 && cdiv(nslices, 1) <= 65535, "pixel_fan_bp: grid too large");
    const int c = tune_choose({TK_PIXEL_BP, g.nx, g.ny, g.nu, g.nview, nslices}, kNPixBp,
                              [&](int i) { kPixBp[i](g, prj_vu, img_xy, nslices, st); }, st, true);
    kPixBp[c](g, prj_vu, img_xy, nslices, st);
    check_launch("pixel_fan_bp");
}

void launch_pack_quads(const float* in, float4* q_same, float4* q_tr, int n1, int n2, int nslices, bool padded,
                       hipStream_t st)
{
    if (nslices <= 0 || n1 <= 0 || n2 <= 0) return;
    dim3 grd(cdiv(n2, 32), cdiv(n1, 32), cdiv(nslices, 4));
    hipLaunchKernelGGL(k_pack_quads, grd, dim3(256), 0, st, in, q_same, q_tr, n1, n2, nslices,
                       padded ? qpitch(n2) : n2, padded ? qpitch(n1) : n1, padded ? kQPadL : 0);
    check_launch("pack_quads");
}

void launch_permute_rows(const float* in, float* out, const int* perm, int rows, int cols, int batch,
                         hipStream_t st)
{
    if (rows <= 0 || cols <= 0 || batch <= 0) return;
    hipLaunchKernelGGL(k_permute_rows, dim3(cdiv(cols, 256), rows, batch), dim3(256), 0, st, in, out, perm, rows,
                       cols);
    check_launch("permute_rows");
}

void launch_transpose(const float* in, float* out, int rows, int cols, int batch, hipStream_t st)
{
    if (batch <= 0 || rows <= 0 || cols <= 0) return;
    dim3 blk(32, 8), grd(cdiv(cols, 32), cdiv(rows, 32), batch);
    hipLaunchKernelGGL(k_transpose, grd, blk, 0, st, in, out, rows, cols);
    check_launch("transpose");
}

void launch_avg_prev(const float* in, float* out, int n0, int n1, int batch, int axis, hipStream_t st)
{
    if (batch <= 0) return;
    dim3 blk(256), grd(cdiv((long)n0 * n1, 256), batch);
    hipLaunchKernelGGL(k_avg_prev, grd, blk, 0, st, in, out, n0, n1, axis);
    check_launch("avg_prev");
}

void launch_mix_gather(const float* src, float* dst, const float* M, int nB, int n, int np, int nc, int nout,
                       hipStream_t st)
{
    dim3 blk(256), grd(cdiv(n, 256), nB * nout * nc);
    hipLaunchKernelGGL(k_mix_gather, grd, blk, 0, st, src, dst, M, n, np, nc, nout);
    check_launch("mix_gather");
}

void launch_mix_scatter(const float* src, float* dst, const float* M, int nB, int n, int np, int nc, int nout,
                        bool accumulate, hipStream_t st)
{
    dim3 blk(256), grd(cdiv(n, 256), nB * nout * nc);
    hipLaunchKernelGGL(k_mix_scatter, grd, blk, 0, st, src, dst, M, n, np, nc, nout, accumulate ? 1 : 0);
    check_launch("mix_scatter");
}

void launch_fill(float* p, float v, size_t n, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3(cdiv((long)n, 256)), dim3(256), 0, st, p, v, n);
    check_launch("fill");
}

void launch_axpy(float* y, const float* x, float a, size_t n, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_axpy, dim3(cdiv((long)n, 256)), dim3(256), 0, st, y, x, a, n);
    check_launch("axpy");
}

} // namespace rn
