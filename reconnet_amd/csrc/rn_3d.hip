// 3D projectors for gfx950: arbitrary-geometry Siddon (parallel / cone),
// DD tomosynthesis, the cone OS-SQS data step and the helical BP.
//
// Reference counterparts: Siddon.cu:11-351 (tracing), SiddonParallel.cu:11-105,
// SiddonCone.cu:11-68, DistanceDriven3D.cu:139-227 / 409-559, ossqs.cu:163-333,
// Recon3D/HelicalBP.cu:10-83.
//
// Siddon tracing keeps the reference's voxel walk, but each crossing
// parameter is computed directly from the boundary index, (b - s) / d, instead
// of by repeated FP32 addition, so rounding does not accumulate along a ray.
// The arbitrary-geometry BP is the exact transpose of that walk; with detector
// axes and sources free per view there is no cheap footprint inversion, so it
// scatters with hardware float atomics (global_atomic_add_f32).  The DD
// tomosynthesis pair integrates footprints by direct 2D overlap sums in FP32
// instead of the reference's FP64 integral images.

#include "rn_3d.h"
#include "rn_fan.h"
#include "rn_tune.h"

#include <cmath>
#include <vector>

namespace rn {

namespace {

struct Grid3K {
    int nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
};

struct F3 { float x, y, z; };

__device__ __forceinline__ F3 ld3(const float* p, int i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// MoveSourceDstNearGrid (SiddonParallel.cu:35-55, cudaMath.h:25-60)
__device__ void move_near_grid(F3& s, F3& t, const Grid3K& g)
{
    const float ex = g.nx * g.dx, ey = g.ny * g.dy, ez = g.nz * g.dz;
    const float r = sqrtf(ex * ex + ey * ey + ez * ez);
    float dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    if (len >= 1e-6f) { dx /= len; dy /= len; dz /= len; }
    const float t0 = (g.cx - s.x) * dx + (g.cy - s.y) * dy + (g.cz - s.z) * dz;
    const float t1 = (t.x - g.cx) * dx + (t.y - g.cy) * dy + (t.z - g.cz) * dz;
    if (t0 > r) { s.x += (t0 - r) * dx; s.y += (t0 - r) * dy; s.z += (t0 - r) * dz; }
    if (t1 > r) { t.x -= (t1 - r) * dx; t.y -= (t1 - r) * dy; t.z -= (t1 - r) * dz; }
}

__device__ __forceinline__ void arb_ray(const ArbGeo& a, const Grid3K& g, bool parallel, const float* C,
                                        const float* U, const float* V, const float* S, int view, int iu, int iv,
                                        F3& src, F3& dst)
{
    const float u = (iu - a.off_u - (a.nu - 1) / 2.0f) * a.du;
    const float v = (iv - a.off_v - (a.nv - 1) / 2.0f) * a.dv;
    const F3 c = ld3(C, view), du = ld3(U, view), dv = ld3(V, view), sd = ld3(S, view);
    dst = F3{c.x + du.x * u + dv.x * v, c.y + du.y * u + dv.y * v, c.z + du.z * u + dv.z * v};
    if (parallel) {
        // GetSrcDstForParallel (SiddonParallel.cu:11-32)
        const float m1 = sqrtf((g.cx - c.x) * (g.cx - c.x) + (g.cy - c.y) * (g.cy - c.y) + (g.cz - c.z) * (g.cz - c.z));
        const float ex = g.nx * g.dx, ey = g.ny * g.dy, ez = g.nz * g.dz;
        const float m2 = sqrtf(ex * ex + ey * ey + ez * ez) / 2;
        const float m = fmaxf(m1, m2) * 1.1f;
        src = F3{c.x + sd.x * m + du.x * u + dv.x * v, c.y + sd.y * m + du.y * u + dv.y * v,
                 c.z + sd.z * m + du.z * u + dv.z * v};
    } else {
        src = sd;
    }
    move_near_grid(src, dst, g);
}

// Siddon walk (Siddon.cu:90-238).  f(voxel, length), voxel the linear index
// x*ny*nz + y*nz + z, or with OFF32 that index times 4 nC as a 32-bit byte
// offset, kept by adding the axis stride at each step (a slice's volume below
// 4 GiB: checked on the host).
template <bool OFF32, class F>
__device__ void siddon_walk(F3 src, F3 dst, const Grid3K& g, int nC, F f)
{
    const float eps = 1e-6f;
    const float dist = sqrtf((src.x - dst.x) * (src.x - dst.x) + (src.y - dst.y) * (src.y - dst.y) +
                             (src.z - dst.z) * (src.z - dst.z));
    const float sx = (src.x - g.cx) / g.dx + g.nx / 2.0f, sy = (src.y - g.cy) / g.dy + g.ny / 2.0f,
                sz = (src.z - g.cz) / g.dz + g.nz / 2.0f;
    const float tx = (dst.x - g.cx) / g.dx + g.nx / 2.0f, ty = (dst.y - g.cy) / g.dy + g.ny / 2.0f,
                tz = (dst.z - g.cz) / g.dz + g.nz / 2.0f;
    const float ddx = tx - sx, ddy = ty - sy, ddz = tz - sz;
    float a0x = 1, a1x = 0, a0y = 1, a1y = 0, a0z = 1, a1z = 0;
    if (fabsf(ddx) > eps) { a0x = -sx / ddx; a1x = (g.nx - sx) / ddx; } else if (sx < 0 || sx > g.nx) return;
    if (fabsf(ddy) > eps) { a0y = -sy / ddy; a1y = (g.ny - sy) / ddy; } else if (sy < 0 || sy > g.ny) return;
    if (fabsf(ddz) > eps) { a0z = -sz / ddz; a1z = (g.nz - sz) / ddz; } else if (sz < 0 || sz > g.nz) return;
    const float amin = fmaxf(0.f, fmaxf(fminf(a0x, a1x), fmaxf(fminf(a0y, a1y), fminf(a0z, a1z))));
    const float amax = fminf(1.f, fminf(fmaxf(a0x, a1x), fminf(fmaxf(a0y, a1y), fmaxf(a0z, a1z))));
    if (amax <= amin) return;
    int gx = (int)(sx + amin * ddx), gy = (int)(sy + amin * ddy), gz = (int)(sz + amin * ddz);
    if (fabsf(ddx) > eps) gx = min(max(gx, 0), g.nx - 1);
    if (fabsf(ddy) > eps) gy = min(max(gy, 0), g.ny - 1);
    if (fabsf(ddz) > eps) gz = min(max(gz, 0), g.nz - 1);
    if (gx < 0 || gx >= g.nx || gy < 0 || gy >= g.ny || gz < 0 || gz >= g.nz) return;
    const int px = ddx > 0, py = ddy > 0, pz = ddz > 0;
    // next boundary index per axis (OutboundFirstVoxel) and its crossing parameter
    float bx = (ddx > eps) ? gx + 1 : gx, by = (ddy > eps) ? gy + 1 : gy, bz = (ddz > eps) ? gz + 1 : gz;
    const bool mx = fabsf(ddx) > eps, my = fabsf(ddy) > eps, mz = fabsf(ddz) > eps;
    const float ix = mx ? dist / ddx : 0.f, iy = my ? dist / ddy : 0.f, iz = mz ? dist / ddz : 0.f;
    const float stx = px ? 1.f : -1.f, sty = py ? 1.f : -1.f, stz = pz ? 1.f : -1.f;
    float ax = mx ? (bx - sx) * ix : 2.f * dist;
    float ay = my ? (by - sy) * iy : 2.f * dist;
    float az = mz ? (bz - sz) * iz : 2.f * dist;
    // with no motion along an axis the reference keeps alpha = 2*dist + k*0
    float prev = amin * dist;
    const int maxit = g.nx + g.ny + g.nz + 4;
    // byte offset of the current voxel and the signed steps along each axis
    // (unsigned arithmetic: a step out of the grid ends the walk unused)
    const unsigned oz = 4u * (unsigned)nC, oy = oz * (unsigned)g.nz, ox = oy * (unsigned)g.ny;
    unsigned off = OFF32 ? ((unsigned)gx * (unsigned)g.ny + (unsigned)gy) * oy + (unsigned)gz * oz : 0u;
    const unsigned dox = px ? ox : 0u - ox, doy = py ? oy : 0u - oy, doz = pz ? oz : 0u - oz;
    for (int it = 0; it < maxit; ++it) {
        const int cx = gx, cy = gy, cz = gz;
        const unsigned coff = off;
        float now;
        bool done = false;
        if (ax < ay && ax < az) {
            now = ax;
            bx += stx; ax = mx ? (bx - sx) * ix : ax;
            gx += px ? 1 : -1; done = (gx >= g.nx || gx < 0);
            off += dox;
        } else if (ay < az) {
            now = ay;
            by += sty; ay = my ? (by - sy) * iy : ay;
            gy += py ? 1 : -1; done = (gy >= g.ny || gy < 0);
            off += doy;
        } else {
            now = az;
            bz += stz; az = mz ? (bz - sz) * iz : az;
            gz += pz ? 1 : -1; done = (gz >= g.nz || gz < 0);
            off += doz;
        }
        if (OFF32) f((long)coff, now - prev);
        else f(((long)cx * g.ny + cy) * g.nz + cz, now - prev);
        prev = now;
        if (done) break;
    }
}

// One ray per lane; a wave holds 64 adjacent detector rows of one column, so
// its rays share their in-plane path when the rows run along z and read
// adjacent z voxels (z is the innermost image index), and its outputs are
// adjacent; 4 columns per workgroup.
template <bool OFF32>
__global__ __launch_bounds__(256) void k_siddon_arb_fp(ArbGeo a, Grid3K g, int parallel, const float* __restrict__ img,
                                                       float* __restrict__ prj, const float* C, const float* U,
                                                       const float* V, const float* S, int nC)
{
    const int iv = blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int iu = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int view = blockIdx.z % a.nview, bc = blockIdx.z / a.nview;
    const int c = bc % nC, b = bc / nC;
    if (iu >= a.nu || iv >= a.nv) return;
    F3 s, t;
    arb_ray(a, g, parallel != 0, C, U, V, S, view, iu, iv, s, t);
    const float* base = img + (long)b * g.nx * g.ny * g.nz * nC + c;
    float val = 0.f;
    if (OFF32) {
        const char* bb = reinterpret_cast<const char*>(base);
        siddon_walk<true>(s, t, g, nC, [&](long off, float len) {
            val += *reinterpret_cast<const float*>(bb + (unsigned)off) * len;
        });
    } else {
        siddon_walk<false>(s, t, g, nC, [&](long vox, float len) { val += base[vox * nC] * len; });
    }
    prj[((((long)b * a.nu + iu) * a.nview + view) * a.nv + iv) * nC + c] = val;
}

// Siddon backprojection as a gather, the transpose of k_siddon_arb_fp: a
// voxel collects, per view, the rays of the detector cells that lie in the
// projection of its 8 corners, each weighted by the ray's chord through the
// voxel.  The ray is the walk's: from its start point (the source; for
// parallel views the detector point moved back by m along invRayDir) through
// the grid and on past the detector point, and it is used only if it enters
// the grid before the detector point (the walk's amin < amax with amax =
// min(1, exit)); moving the start point to the grid (MoveSourceDstNearGrid)
// keeps it outside the grid, so it changes no chord.  One wave holds a
// column of 64 voxels along z with (x, y) wave-uniform; no atomics, every
// voxel written once.

// [enter, exit] of the ray o + t d in the box [lo, hi] (empty: enter > exit);
// an axis with d = 0 keeps the interval if lo <= o < hi (the walk's voxel
// index of a stationary coordinate)
__device__ __forceinline__ void slab(const F3& o, const F3& d, const F3& lo, const F3& hi, float& t0, float& t1)
{
    t0 = -INFINITY;
    t1 = INFINITY;
    const float oc[3] = {o.x, o.y, o.z}, dc[3] = {d.x, d.y, d.z};
    const float lc[3] = {lo.x, lo.y, lo.z}, hc[3] = {hi.x, hi.y, hi.z};
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (dc[k] != 0.0f) {
            const float inv = __builtin_amdgcn_rcpf(dc[k]);
            const float a0 = (lc[k] - oc[k]) * inv, a1 = (hc[k] - oc[k]) * inv;
            t0 = fmaxf(t0, fminf(a0, a1));
            t1 = fminf(t1, fmaxf(a0, a1));
        } else if (!(oc[k] >= lc[k] && oc[k] < hc[k])) {
            t1 = -INFINITY;
        }
    }
}

// [s0, s1] on the ray p(s) = s d - r (a voxel-centred frame) inside the box
// |p| <= h, as slab() with the half-open faces: an axis with d = 0 keeps the
// interval if -h < r <= h.  r is small (the ray passes within a few voxels of
// the centre), so the parameters carry float precision relative to the chord
// instead of relative to the source distance.
__device__ __forceinline__ void slab_local(const F3& r, const F3& d, const F3& h, float& s0, float& s1)
{
    s0 = -INFINITY;
    s1 = INFINITY;
    const float rc[3] = {r.x, r.y, r.z}, dc[3] = {d.x, d.y, d.z}, hc[3] = {h.x, h.y, h.z};
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (dc[k] != 0.0f) {
            const float inv = __builtin_amdgcn_rcpf(dc[k]);
            const float a0 = (rc[k] - hc[k]) * inv, a1 = (rc[k] + hc[k]) * inv;
            s0 = fmaxf(s0, fminf(a0, a1));
            s1 = fminf(s1, fmaxf(a0, a1));
        } else if (!(rc[k] > -hc[k] && rc[k] <= hc[k])) {
            s1 = -INFINITY;
        }
    }
}

// the wave's copy of a value every active lane holds alike (an SGPR operand)
__device__ __forceinline__ float wave_uniform(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double wave_uniform(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Per view, in FP64: the detector normal N = U x V, the dual basis (Ud, Vd) of
// (U, V) and the projection constants: cone A = (C - S).N, su, sv = (S - C).Ud,
// (S - C).Vd; parallel sdn, sdu, sdv = invRayDir.(N, Ud, Vd) and the start
// distance m of GetSrcDstForParallel (SiddonParallel.cu:11-32).
__global__ void k_siddon_arb_views(ArbGeo a, Grid3K g, int parallel, const float* C, const float* U, const float* V,
                                   const float* S, float4* __restrict__ vt)
{
    const int view = blockIdx.x * blockDim.x + threadIdx.x;
    if (view >= a.nview) return;
    double c[3], u[3], v[3], sd[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = C[3 * view + k]; u[k] = U[3 * view + k]; v[k] = V[3 * view + k]; sd[k] = S[3 * view + k];
    }
    const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const double idet = 1.0 / (uu * vv - uv * uv);
    double ud[3], vd[3];
    for (int k = 0; k < 3; ++k) { ud[k] = (vv * u[k] - uv * v[k]) * idet; vd[k] = (uu * v[k] - uv * u[k]) * idet; }
    auto dot = [](const double* p, const double* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
    double w, wu, wv, m = 0.0;
    if (!parallel) {
        const double cs[3] = {c[0] - sd[0], c[1] - sd[1], c[2] - sd[2]};
        w = dot(cs, n); wu = -dot(cs, ud); wv = -dot(cs, vd);
    } else {
        w = dot(sd, n); wu = dot(sd, ud); wv = dot(sd, vd);
        // the walk's start distance, in its float arithmetic
        const float fx = g.cx - C[3 * view], fy = g.cy - C[3 * view + 1], fz = g.cz - C[3 * view + 2];
        const float ex = g.nx * g.dx, ey = g.ny * g.dy, ez = g.nz * g.dz;
        m = fmaxf(sqrtf(fx * fx + fy * fy + fz * fz), sqrtf(ex * ex + ey * ey + ez * ez) / 2) * 1.1f;
    }
    vt[4 * view + 0] = make_float4((float)n[0], (float)n[1], (float)n[2], (float)w);
    vt[4 * view + 1] = make_float4((float)ud[0], (float)ud[1], (float)ud[2], (float)wu);
    vt[4 * view + 2] = make_float4((float)vd[0], (float)vd[1], (float)vd[2], (float)wv);
    vt[4 * view + 3] = make_float4((float)m, 0.0f, 0.0f, 0.0f);
}

// What one lane's voxels take from one cone view whose detector rows run
// along z and whose columns lie in a horizontal plane (V = (0, 0, v),
// U.z = 0).  N and Ud are then horizontal and Vd = (0, 0, 1/v): a voxel's
// detector columns depend on (x, y) only, so the column range and each
// column's in-plane ray direction and x/y slabs are the wave's (a wave holds
// one (x, y)); the rows scale with z and each adds its z slab.  A lane holds
// ZPL voxels stacked in z from vlz (bottom face) and vcz (centre of the
// lowest, FP64); every row is loaded once for all of them.
struct SidBpCol {
    F3 vlo, vhi;          // the wave's x/y faces (z unused)
    F3 glo, ghi, hw;      // grid box, voxel half widths
    double vcx, vcy;      // the wave's voxel centre (x, y), FP64
    float cu, cv, gdz;
};

template <int ZPL>
__device__ __forceinline__ void siddon_rows_z_view(const ArbGeo& a, const SidBpCol& w, int view, const F3& Cv,
                                                   const F3& Uv, const F3& Vv, const F3& Sv, const float4& t0v,
                                                   const float4& t1v, const float4& t2v, float vlz, double vcz,
                                                   const float* __restrict__ q, int nC, float (&acc)[ZPL])
{
    const F3 N{t0v.x, t0v.y, t0v.z}, Ud{t1v.x, t1v.y, t1v.z};
    const float A = t0v.w, su = t1v.w, sv = t2v.w, vdz = t2v.z;
    float umin = INFINITY, umax = -INFINITY, tmin = INFINITY, tmax = -INFINITY;
    bool ok = true;
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float wxk = ((k & 1) ? w.vhi.x : w.vlo.x) - Sv.x, wyk = ((k & 2) ? w.vhi.y : w.vlo.y) - Sv.y;
        const float den = fmaf(wxk, N.x, wyk * N.y);
        ok = ok && (den * A > 0.f);
        const float t = A * __builtin_amdgcn_rcpf(den);
        const float u = fmaf(t, fmaf(wxk, Ud.x, wyk * Ud.y), su);
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        tmin = fminf(tmin, t); tmax = fmaxf(tmax, t);
    }
    if (!ok) return;
    // rows: the corners' v over the lane's stack of voxels
    const float z0 = (vlz - Sv.z) * vdz, z1 = (vlz + ZPL * w.gdz - Sv.z) * vdz;
    const float vmin = fminf(fminf(tmin * z0, tmax * z0), fminf(tmin * z1, tmax * z1)) + sv;
    const float vmax = fmaxf(fmaxf(tmin * z0, tmax * z0), fmaxf(tmin * z1, tmax * z1)) + sv;
    const float fu0 = umin / a.du + w.cu, fu1 = umax / a.du + w.cu;
    const float fv0 = vmin / a.dv + w.cv, fv1 = vmax / a.dv + w.cv;
    const int iu0 = max((int)ceilf(fminf(fu0, fu1) - 1e-3f), 0), iu1 = min((int)floorf(fmaxf(fu0, fu1) + 1e-3f), a.nu - 1);
    const int iv0 = max((int)ceilf(fminf(fv0, fv1) - 1e-3f), 0), iv1 = min((int)floorf(fmaxf(fv0, fv1) + 1e-3f), a.nv - 1);
    const double wx = wave_uniform(w.vcx - Sv.x), wy = wave_uniform(w.vcy - Sv.y), wz = vcz - Sv.z;
    // rows in float: the lowest voxel's z offset from the source split in two
    // floats, the detector centre's height above the source rounded once
    const float wz_hi = (float)wz, wz_lo = (float)(wz - wz_hi), czs = (float)((double)Cv.z - Sv.z);
    for (int iu = iu0; iu <= iu1; ++iu) {
        const double uc = (double)(iu - w.cu) * a.du;
        const double dxd = wave_uniform(fma((double)Uv.x, uc, (double)Cv.x) - Sv.x);
        const double dyd = wave_uniform(fma((double)Uv.y, uc, (double)Cv.y) - Sv.y);
        const float dx = wave_uniform((float)dxd), dy = wave_uniform((float)dyd);
        const float dxy2 = wave_uniform(fmaf(dx, dx, dy * dy));
        // tp: the column's in-plane closest approach to the voxel (any nearby
        // value serves: r is exact for the tp used); the ray's points are
        // S + (tp + s) d, at s d - r from the voxel centre
        const float tp = wave_uniform((float)(wx * dxd + wy * dyd) * __builtin_amdgcn_rcpf(dxy2));
        const float rx = wave_uniform((float)fma(-(double)tp, dxd, wx));
        const float ry = wave_uniform((float)fma(-(double)tp, dyd, wy));
        float xy0 = -INFINITY, xy1 = INFINITY;
        if (dx != 0.f) {
            const float inv = __builtin_amdgcn_rcpf(dx);
            const float a0 = (rx - w.hw.x) * inv, a1 = (rx + w.hw.x) * inv;
            xy0 = fmaxf(xy0, fminf(a0, a1)); xy1 = fminf(xy1, fmaxf(a0, a1));
        } else if (!(rx > -w.hw.x && rx <= w.hw.x)) {
            xy1 = -INFINITY;
        }
        if (dy != 0.f) {
            const float inv = __builtin_amdgcn_rcpf(dy);
            const float a0 = (ry - w.hw.y) * inv, a1 = (ry + w.hw.y) * inv;
            xy0 = fmaxf(xy0, fminf(a0, a1)); xy1 = fminf(xy1, fmaxf(a0, a1));
        } else if (!(ry > -w.hw.y && ry <= w.hw.y)) {
            xy1 = -INFINITY;
        }
        xy0 = wave_uniform(fmaxf(xy0, -tp));          // t >= 0: from the source on
        xy1 = wave_uniform(xy1);
        if (!(xy1 > xy0)) continue;
        // a voxel can lie beyond the detector point only if the column's rays
        // are still inside the voxel column there
        const bool far = xy1 > 1.f - tp;
        const float* qc = q + ((long)iu * a.nview + view) * a.nv * nC;
        // the row's ray height is affine in the row index
        const float dzs = Vv.z * a.dv, dz_0 = fmaf(-w.cv, dzs, czs);
        for (int iv = iv0; iv <= iv1; ++iv) {
            const float dz = fmaf(dzs, (float)iv, dz_0), rz = fmaf(-tp, dz, wz_hi) + wz_lo;
            const float len = __builtin_amdgcn_sqrtf(fmaf(dz, dz, dxy2));      // hardware sqrt, 1 ulp
            const float val = qc[(long)iv * nC];
            // the lowest voxel's z slab; voxel k's is shifted by k dz_vox / dz
            float zl, zh, st;
            if (dz != 0.f) {
                const float inv = __builtin_amdgcn_rcpf(dz);
                const float a0 = (rz - w.hw.z) * inv, a1 = (rz + w.hw.z) * inv;
                zl = fminf(a0, a1); zh = fmaxf(a0, a1); st = w.gdz * inv;
            } else {
                // a ray at a voxel's height for all t: voxel k holds it if
                // -h < rz + k dz_vox <= h
                zl = -INFINITY; zh = INFINITY; st = 0.f;
            }
            #pragma unroll
            for (int k = 0; k < ZPL; ++k) {
                if (dz == 0.f) {
                    const float rzk = fmaf((float)k, w.gdz, rz);
                    if (!(rzk > -w.hw.z && rzk <= w.hw.z)) continue;
                }
                const float s0 = fmaxf(xy0, fmaf((float)k, st, zl)), s1 = fminf(xy1, fmaf((float)k, st, zh));
                if (!(s1 > s0)) continue;
                if (far && s0 >= 1.f - tp) {
                    // a voxel beyond the detector point: the ray counts only if
                    // it entered the grid before that point
                    float g0, g1;
                    slab(Sv, F3{dx, dy, dz}, w.glo, w.ghi, g0, g1);
                    if (!(fminf(g1, 1.f) > fmaxf(g0, 0.f))) continue;
                }
                acc[k] = fmaf((s1 - s0) * len, val, acc[k]);
            }
        }
    }
}

// ROWS_Z: the caller has found every view to be an upright cone view (the
// kernel then has no general-view path: its FP64 ray set-up would double
// the register count); ZPL voxels per lane stacked in z (ROWS_Z only).
template <bool ROWS_Z, int ZPL>
__global__ __launch_bounds__(256) void k_siddon_arb_bp(ArbGeo a, Grid3K g, int parallel, const float* __restrict__ prj,
                                                       float* __restrict__ img, const float* C, const float* U,
                                                       const float* V, const float* S, const float4* __restrict__ vt,
                                                       int nC)
{
    static_assert(ROWS_Z || ZPL == 1, "stacked voxels need upright cone views");
    const int iz = (blockIdx.x * 64 + (threadIdx.x & 63)) * ZPL;
    const int iy = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (int)(threadIdx.x >> 6));
    const int ix = blockIdx.z % g.nx, bc = blockIdx.z / g.nx;
    if (iz >= g.nz || iy >= g.ny) return;
    const int cc = bc % nC, b = bc / nC;
    const float* q = prj + (long)b * a.nu * a.nview * a.nv * nC + cc;
    // wave-uniform values are kept in scalar registers (wave_uniform)
    SidBpCol w;
    const float vlz = g.cz + (iz - g.nz / 2.0f) * g.dz;
    w.vlo = F3{wave_uniform(g.cx + (ix - g.nx / 2.0f) * g.dx), wave_uniform(g.cy + (iy - g.ny / 2.0f) * g.dy), vlz};
    w.vhi = F3{wave_uniform(w.vlo.x + g.dx), wave_uniform(w.vlo.y + g.dy), vlz + g.dz};
    w.glo = F3{wave_uniform(g.cx - g.nx / 2.0f * g.dx), wave_uniform(g.cy - g.ny / 2.0f * g.dy),
               wave_uniform(g.cz - g.nz / 2.0f * g.dz)};
    w.ghi = F3{wave_uniform(g.cx + g.nx / 2.0f * g.dx), wave_uniform(g.cy + g.ny / 2.0f * g.dy),
               wave_uniform(g.cz + g.nz / 2.0f * g.dz)};
    // the voxel centre in FP64 and its half widths: chords are measured from a
    // point of the ray next to the voxel, because parameters counted from
    // the source (~600 mm away) lose ~5e-5 mm to float rounding
    w.vcx = wave_uniform(g.cx + (ix + 0.5 - g.nx / 2.0) * g.dx);
    w.vcy = wave_uniform(g.cy + (iy + 0.5 - g.ny / 2.0) * g.dy);
    const double vcz = g.cz + (iz + 0.5 - g.nz / 2.0) * g.dz;
    w.hw = F3{wave_uniform(0.5f * g.dx), wave_uniform(0.5f * g.dy), wave_uniform(0.5f * g.dz)};
    w.cu = wave_uniform((a.nu - 1) / 2.0f + a.off_u);
    w.cv = wave_uniform((a.nv - 1) / 2.0f + a.off_v);
    w.gdz = wave_uniform(g.dz);
    auto dot = [](const F3& p, const F3& r) { return p.x * r.x + p.y * r.y + p.z * r.z; };
    float acc[ZPL];
    #pragma unroll
    for (int k = 0; k < ZPL; ++k) acc[k] = 0.f;
    for (int view = 0; view < a.nview; ++view) {
        const F3 Cv = ld3(C, view), Uv = ld3(U, view), Vv = ld3(V, view), Sv = ld3(S, view);
        const float4 t0v = vt[4 * view], t1v = vt[4 * view + 1], t2v = vt[4 * view + 2];
        if (ROWS_Z || (!parallel && Vv.x == 0.f && Vv.y == 0.f && Uv.z == 0.f)) {
            siddon_rows_z_view<ZPL>(a, w, view, Cv, Uv, Vv, Sv, t0v, t1v, t2v, vlz, vcz, q, nC, acc);
            continue;
        }
        if constexpr (!ROWS_Z) {
            // any other view: the voxel's eight corners on the detector (u, v in mm)
            const F3 N{t0v.x, t0v.y, t0v.z}, Ud{t1v.x, t1v.y, t1v.z}, Vd{t2v.x, t2v.y, t2v.z};
            const F3 vlo = w.vlo, vhi = w.vhi;
            float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
            bool ok = true;
            float m = 0.f;
            if (!parallel) {
                const float A = t0v.w, su = t1v.w, sv = t2v.w;
                #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const F3 wk{((k & 1) ? vhi.x : vlo.x) - Sv.x, ((k & 2) ? vhi.y : vlo.y) - Sv.y,
                                ((k & 4) ? vhi.z : vlo.z) - Sv.z};
                    const float den = dot(wk, N);
                    ok = ok && (den * A > 0.f);
                    const float t = A * __builtin_amdgcn_rcpf(den);
                    const float u = fmaf(t, dot(wk, Ud), su), v = fmaf(t, dot(wk, Vd), sv);
                    umin = fminf(umin, u); umax = fmaxf(umax, u);
                    vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
                }
            } else {
                const float sdn = t0v.w, sdu = t1v.w, sdv = t2v.w;
                ok = sdn != 0.f;
                const float isdn = -__builtin_amdgcn_rcpf(sdn);
                #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const F3 wk{((k & 1) ? vhi.x : vlo.x) - Cv.x, ((k & 2) ? vhi.y : vlo.y) - Cv.y,
                                ((k & 4) ? vhi.z : vlo.z) - Cv.z};
                    const float tau = dot(wk, N) * isdn;
                    const float u = fmaf(tau, sdu, dot(wk, Ud)), v = fmaf(tau, sdv, dot(wk, Vd));
                    umin = fminf(umin, u); umax = fmaxf(umax, u);
                    vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
                }
                m = vt[4 * view + 3].x;
            }
            if (!ok) continue;
            // cells whose centre ray meets the corners' bounding box
            const float fu0 = umin / a.du + w.cu, fu1 = umax / a.du + w.cu;
            const float fv0 = vmin / a.dv + w.cv, fv1 = vmax / a.dv + w.cv;
            const int iu0 = max((int)ceilf(fminf(fu0, fu1) - 1e-3f), 0);
            const int iu1 = min((int)floorf(fmaxf(fu0, fu1) + 1e-3f), a.nu - 1);
            const int iv0 = max((int)ceilf(fminf(fv0, fv1) - 1e-3f), 0);
            const int iv1 = min((int)floorf(fmaxf(fv0, fv1) + 1e-3f), a.nv - 1);
            for (int iu = iu0; iu <= iu1; ++iu) {
                const double uc = (double)(iu - w.cu) * a.du;
                const float* qc = q + ((long)iu * a.nview + view) * a.nv * nC;
                for (int iv = iv0; iv <= iv1; ++iv) {
                    const double vc = (double)(iv - w.cv) * a.dv;
                    const double dst[3] = {Cv.x + Uv.x * uc + Vv.x * vc, Cv.y + Uv.y * uc + Vv.y * vc,
                                           Cv.z + Uv.z * uc + Vv.z * vc};
                    const double o[3] = {parallel ? dst[0] + (double)Sv.x * m : (double)Sv.x,
                                         parallel ? dst[1] + (double)Sv.y * m : (double)Sv.y,
                                         parallel ? dst[2] + (double)Sv.z * m : (double)Sv.z};
                    const double dd[3] = {dst[0] - o[0], dst[1] - o[1], dst[2] - o[2]};
                    const double wv[3] = {w.vcx - o[0], w.vcy - o[1], vcz - o[2]};
                    const float tp = (float)((wv[0] * dd[0] + wv[1] * dd[1] + wv[2] * dd[2]) /
                                             (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]));
                    const F3 r{(float)fma(-(double)tp, dd[0], wv[0]), (float)fma(-(double)tp, dd[1], wv[1]),
                               (float)fma(-(double)tp, dd[2], wv[2])};
                    const F3 d{(float)dd[0], (float)dd[1], (float)dd[2]};
                    float s0, s1;
                    slab_local(r, d, w.hw, s0, s1);
                    s0 = fmaxf(s0, -tp);
                    if (!(s1 > s0)) continue;
                    if (s0 >= 1.f - tp) {
                        // a voxel beyond the detector point: the ray counts only
                        // if it entered the grid before that point
                        const F3 of{(float)o[0], (float)o[1], (float)o[2]};
                        float g0, g1;
                        slab(of, d, w.glo, w.ghi, g0, g1);
                        if (!(fminf(g1, 1.f) > fmaxf(g0, 0.f))) continue;
                    }
                    acc[0] = fmaf((s1 - s0) * __builtin_amdgcn_sqrtf(dot(d, d)), qc[(long)iv * nC], acc[0]);
                }
            }
        }
    }
    float* out = img + ((((long)b * g.nx + ix) * g.ny + iy) * g.nz + iz) * nC + cc;
    #pragma unroll
    for (int k = 0; k < ZPL; ++k)
        if (iz + k < g.nz) out[(long)k * nC] = acc[k];
}

// ---- DD tomosynthesis ----

// average over [x1,x2] x [y1,y2] of a [ny][nx] plane (zero outside)
__device__ __forceinline__ float box_avg(const float* __restrict__ P, int nx, int ny, float x1, float x2, float y1,
                                         float y2)
{
    const float area = (x2 - x1) * (y2 - y1);
    const float lx = fmaxf(x1, 0.f), hx = fminf(x2, (float)nx);
    const float ly = fmaxf(y1, 0.f), hy = fminf(y2, (float)ny);
    if (!(hx > lx) || !(hy > ly)) return 0.f;
    const int m0 = (int)lx, m1 = min((int)hx, nx - 1), n0 = (int)ly, n1 = min((int)hy, ny - 1);
    float s = 0.f;
    for (int n = n0; n <= n1; ++n) {
        const float oy = fminf(hy, (float)(n + 1)) - fmaxf(ly, (float)n);
        if (oy <= 0.f) continue;
        float r = 0.f;
        for (int m = m0; m <= m1; ++m) {
            const float ox = fminf(hx, (float)(m + 1)) - fmaxf(lx, (float)m);
            r = fmaf(P[(long)n * nx + m], fmaxf(ox, 0.f), r);
        }
        s = fmaf(r, oy, s);
    }
    return s / area;
}

// vol_zyx: [slice][nz][ny][nx]
__global__ __launch_bounds__(128) void k_dd_tomo_fp(ArbGeo a, Grid3K g, const float* __restrict__ vol_zyx,
                                                    float* __restrict__ prj, const float* C, const float* S, int nC)
{
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int view = blockIdx.y;
    const int bc = blockIdx.z;
    const int c = bc % nC, biv = bc / nC, iv = biv % a.nv, b = biv / a.nv;
    if (iu >= a.nu) return;
    const F3 cc = ld3(C, view), s0 = ld3(S, view);
    const float u = (iu - a.off_u - (a.nu - 1) / 2.0f) * a.du;
    const float v = (iv - a.off_v - (a.nv - 1) / 2.0f) * a.dv;
    auto tix = [&](float px) { return (px - g.cx) / g.dx + g.nx / 2.0f; };
    auto tiy = [&](float py) { return (py - g.cy) / g.dy + g.ny / 2.0f; };
    const float zc = (cc.z - g.cz) / g.dz + g.nz / 2.0f;
    const float sx = tix(s0.x), sy = tiy(s0.y), sz = (s0.z - g.cz) / g.dz + g.nz / 2.0f;
    float x1 = tix(cc.x + u - a.du / 2), x2 = tix(cc.x + u + a.du / 2);
    float y1 = tiy(cc.y + v - a.dv / 2), y2 = tiy(cc.y + v + a.dv / 2);
    if (x1 > x2) { const float t = x1; x1 = x2; x2 = t; }
    if (y1 > y2) { const float t = y1; y1 = y2; y2 = t; }
    const float dzs = zc - sz;
    const float rx1 = (x1 - sx) / dzs, rx2 = (x2 - sx) / dzs, ry1 = (y1 - sy) / dzs, ry2 = (y2 - sy) / dzs;
    const long plane = (long)g.nx * g.ny;
    const float* vol = vol_zyx + (long)(b * nC + c) * plane * g.nz;
    float val = 0.f;
    for (int iz = 0; iz < g.nz; ++iz) {
        const float h = (float)iz - sz;
        val += box_avg(vol + iz * plane, g.nx, g.ny, fmaf(rx1, h, sx), fmaf(rx2, h, sx), fmaf(ry1, h, sy),
                       fmaf(ry2, h, sy));
    }
    const float ddx = s0.x - (cc.x + u), ddy = s0.y - (cc.y + v), ddz = s0.z - cc.z;
    val *= g.dz / fabsf(ddz) * sqrtf(ddz * ddz + ddy * ddy + ddx * ddx);
    prj[((((long)b * a.nu + iu) * a.nview + view) * a.nv + iv) * nC + c] = val;
}

// wprj: [slice][nview][nv][nu] pre-weighted; out img reference layout
__global__ void k_dd_tomo_preweight(ArbGeo a, const float* __restrict__ prj_ref, float* __restrict__ w,
                                    const float* C, const float* S, int nC)
{
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int view = blockIdx.y;
    const int bc = blockIdx.z;   // (b*nv + iv)*nC + c
    const int c = bc % nC, biv = bc / nC, iv = biv % a.nv, b = biv / a.nv;
    if (iu >= a.nu) return;
    const F3 cc = ld3(C, view), s0 = ld3(S, view);
    const float u = (iu - a.off_u - (a.nu - 1) / 2.0f) * a.du;
    const float v = (iv - a.off_v - (a.nv - 1) / 2.0f) * a.dv;
    const float ddx = s0.x - (cc.x + u), ddy = s0.y - (cc.y + v), ddz = s0.z - cc.z;
    // PreweightBPCartKernelXY (DistanceDriven3D.cu:429-430)
    const float wgt = a.dz / fabsf(ddz) * sqrtf(ddz * ddz + ddy * ddy + ddx * ddx);
    const float pv = prj_ref[((((long)b * a.nu + iu) * a.nview + view) * a.nv + iv) * nC + c];
    w[((((long)b * nC + c) * a.nview + view) * a.nv + iv) * a.nu + iu] = pv * wgt;
}

__global__ __launch_bounds__(256) void k_dd_tomo_bp(ArbGeo a, Grid3K g, const float* __restrict__ w,
                                                    float* __restrict__ img, const float* C, const float* S, int nC)
{
    const int iy = blockIdx.x * 16 + threadIdx.x;
    const int ix = blockIdx.y * 16 + threadIdx.y;
    const int bz = blockIdx.z;   // (b*nC + c)*nz + iz
    const int iz = bz % g.nz, bc = bz / g.nz, c = bc % nC, b = bc / nC;
    if (ix >= g.nx || iy >= g.ny) return;
    const float x = (ix - (g.nx - 1) / 2.0f) * g.dx + g.cx;
    const float y = (iy - (g.ny - 1) / 2.0f) * g.dy + g.cy;
    const float z = (iz - (g.nz - 1) / 2.0f) * g.dz + g.cz;
    const float cu = (a.nu - 1) / 2.0f + a.off_u, cv = (a.nv - 1) / 2.0f + a.off_v;
    const float* wb = w + (long)(b * nC + c) * a.nview * a.nv * a.nu;
    float val = 0.f;
    for (int view = 0; view < a.nview; ++view) {
        const F3 cc = ld3(C, view), s0 = ld3(S, view);
        // ProjectConeToDetCart (DistanceDriven3D.cu:478-489); its output is a bin-centre
        // index used directly as an integral-image coordinate (DistanceDriven3D.cu:544-553)
        const float r = (cc.z - s0.z) / (z - s0.z);
        float u1 = (s0.x - cc.x + r * (x - g.dx / 2 - s0.x)) / a.du + cu;
        float u2 = (s0.x - cc.x + r * (x + g.dx / 2 - s0.x)) / a.du + cu;
        float v1 = (s0.y - cc.y + r * (y - g.dy / 2 - s0.y)) / a.dv + cv;
        float v2 = (s0.y - cc.y + r * (y + g.dy / 2 - s0.y)) / a.dv + cv;
        if (u1 > u2) { const float t = u1; u1 = u2; u2 = t; }
        if (v1 > v2) { const float t = v1; v1 = v2; v2 = t; }
        val += box_avg(wb + (long)view * a.nv * a.nu, a.nu, a.nv, u1, u2, v1, v2);
    }
    img[((((long)b * g.nx + ix) * g.ny + iy) * g.nz + iz) * nC + c] = val;
}

// ---- circular cone-beam DD with a per-view main axis ----
//
// X views (|C.x - S.x| >= |C.y - S.y|, decided on the host) integrate over the
// planes x = voxel-centre x; Y views over y planes.  On plane p the footprint
// of detector cell (iu, iv) is [e1(p), e2(p)] x [z1(p), z2(p)] in index
// coordinates of the in-plane axis and of z: the plane crossings of the rays
// through the cell's u edges (at the cell's v centre) and through its v edges
// (at its u centre).  Each crossing is affine in the plane index, so a ray is
// described by four (alpha, beta) pairs and its per-plane length L.  The
// forward value is sum_p (rectangle average of plane p) * L; the adjoint
// gathers, per voxel, exactly those weights (overlap area / rectangle area * L).

struct ConeRay {
    float a[4], b[4];   // e1, e2, z1, z2 = a + b * plane
    float L;
    bool ok;
};

__device__ __forceinline__ ConeRay cone_ray(const ArbGeo& a, const Grid3K& g, const F3& c, const F3& U, const F3& V,
                                            const F3& s, bool xv, int iu, int iv)
{
    ConeRay r;
    const float uc = (iu - a.off_u - (a.nu - 1) / 2.0f) * a.du, vc = (iv - a.off_v - (a.nv - 1) / 2.0f) * a.dv;
    auto P = [&](float u, float v) {
        return F3{fmaf(v, V.x, fmaf(u, U.x, c.x)), fmaf(v, V.y, fmaf(u, U.y, c.y)), fmaf(v, V.z, fmaf(u, U.z, c.z))};
    };
    const F3 pc = P(uc, vc);
    const F3 pts[4] = {P(uc - 0.5f * a.du, vc), P(uc + 0.5f * a.du, vc), P(uc, vc - 0.5f * a.dv), P(uc, vc + 0.5f * a.dv)};
    // main axis m, in-plane axis n (plane coordinate of plane 0 and its step)
    const float sm = xv ? s.x : s.y, sn = xv ? s.y : s.x;
    const float dm = xv ? g.dx : g.dy, dn = xv ? g.dy : g.dx;
    const float cn = xv ? g.cy : g.cx;
    const int nn = xv ? g.ny : g.nx, nm = xv ? g.nx : g.ny;
    const float m0 = (0.5f - 0.5f * nm) * dm + (xv ? g.cx : g.cy);
    const float dpc = (xv ? pc.x : pc.y) - sm;
    r.ok = fabsf(dpc) > 1e-6f;
    const float lx = pc.x - s.x, ly = pc.y - s.y, lz = pc.z - s.z;
    r.L = dm * sqrtf(fmaf(lx, lx, fmaf(ly, ly, lz * lz))) / fabsf(dpc);
    #pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float em = (xv ? pts[e].x : pts[e].y) - sm;
        const float eo = e < 2 ? (xv ? pts[e].y : pts[e].x) - sn : pts[e].z - s.z;
        const float k = eo / em;                                  // slope along the main axis
        const float so = e < 2 ? sn : s.z, co = e < 2 ? cn : g.cz, d = e < 2 ? dn : g.dz;
        const float no = e < 2 ? (float)nn : (float)g.nz;
        r.b[e] = k * dm / d;
        r.a[e] = (fmaf(m0 - sm, k, so) - co) / d + 0.5f * no;
        r.ok = r.ok && fabsf(em) > 1e-6f;
    }
    return r;
}

// vol, prj: reference layouts [b][nx][ny][nz][c], [b][nu][nview][nv][c]
// XCD > 0: 1-D grid (padded to a multiple of XCD) renumbered so that each XCD
// walks a contiguous range of (row block, column, view): the adjacent detector
// columns of a view, whose footprints share plane cells, meet in one L2.
// OFF32: the 2 x 2 path addresses a plane's cells as a wave-uniform plane
// base plus 32-bit byte offsets from 24-bit multiplies (byte strides below
// 2^24, one slice's volume below 4 GiB: checked on the host).
template <int XCD, bool OFF32>
__global__ __launch_bounds__(64) void k_dd_cone_fp(ArbGeo a, Grid3K g, const float* __restrict__ vol,
                                                   float* __restrict__ prj, const float* Cp, const float* Up,
                                                   const float* Vp, const float* Sp, const int* __restrict__ cls,
                                                   int nC, int nblocks)
{
    int bx, iu, bz;
    if (XCD > 0) {
        const int per = gridDim.x / XCD;
        const int l = (blockIdx.x % XCD) * per + blockIdx.x / XCD;
        if (l >= nblocks) return;
        const int nvb = (a.nv + 63) / 64;
        bx = l % nvb;
        iu = (l / nvb) % a.nu;
        bz = l / (nvb * a.nu);
    } else {
        bx = blockIdx.x;
        iu = blockIdx.y;
        bz = blockIdx.z;
    }
    const int iv = bx * 64 + threadIdx.x;
    const int view = bz % a.nview, bc = bz / a.nview;
    if (iv >= a.nv) return;
    const bool xv = cls[view] == 0;
    ConeRay r = cone_ray(a, g, ld3(Cp, view), ld3(Up, view), ld3(Vp, view), ld3(Sp, view), xv, iu, iv);
    const int nm = xv ? g.nx : g.ny, nn = xv ? g.ny : g.nx;
    const long sP = xv ? (long)g.ny * g.nz : g.nz, sJ = xv ? g.nz : (long)g.ny * g.nz;
    const int c = bc % nC, b = bc / nC;
    const float* base = vol + (long)b * g.nx * g.ny * g.nz * nC + c;   // reference layout, channel innermost
    float acc = 0.f;
    // the in-plane edges do not depend on the row (detV = (0, 0, v)): the
    // wave's, in scalar registers
    #pragma unroll
    for (int e = 0; e < 2; ++e) {
        r.a[e] = wave_uniform(r.a[e]);
        r.b[e] = wave_uniform(r.b[e]);
    }
    // footprints at most one cell wide and high on every plane (the widths
    // are affine in the plane index: the end planes bound them): at most 2 x
    // 2 cells, taken without loops
    const float wl = nm - 1.0f;
    const bool small = fmaxf(fabsf(r.a[1] - r.a[0]), fabsf(fmaf(r.b[1] - r.b[0], wl, r.a[1] - r.a[0]))) <= 1.0f &&
                       fmaxf(fabsf(r.a[3] - r.a[2]), fabsf(fmaf(r.b[3] - r.b[2], wl, r.a[3] - r.a[2]))) <= 1.0f;
    if (r.ok && small) {
        // the two edges of a pair never cross between planes (the footprint
        // would vanish): order each pair once
        const bool se = r.a[0] > r.a[1], sz = r.a[2] > r.a[3];
        const float ea1 = se ? r.a[1] : r.a[0], eb1 = se ? r.b[1] : r.b[0];
        const float ea2 = se ? r.a[0] : r.a[1], eb2 = se ? r.b[0] : r.b[1];
        const float za1 = sz ? r.a[3] : r.a[2], zb1 = sz ? r.b[3] : r.b[2];
        const float za2 = sz ? r.a[2] : r.a[3], zb2 = sz ? r.b[2] : r.b[3];
        for (int p = 0; p < nm; ++p) {
            const float fp = (float)p;
            const float e1 = fmaf(eb1, fp, ea1), e2 = fmaf(eb2, fp, ea2);
            const float le = fmaxf(e1, 0.f), he = fminf(e2, (float)nn);
            if (!(he > le)) continue;
            const float z1 = fmaf(zb1, fp, za1), z2 = fmaf(zb2, fp, za2);
            const float lz = fmaxf(z1, 0.f), hz = fminf(z2, (float)g.nz);
            const float area = (e2 - e1) * (z2 - z1);
            if (!(hz > lz) || !(area > 0.f)) continue;
            // cells j0, j0 + 1 and k0, k0 + 1; a second cell's share is 0 when
            // the footprint ends in the first (its index is then clamped)
            const int j0 = (int)le, k0 = (int)lz;
            const float oe0 = fminf(he, (float)(j0 + 1)) - le, oe1 = fmaxf(he - (float)(j0 + 1), 0.f);
            const float oz0 = fminf(hz, (float)(k0 + 1)) - lz, oz1 = fmaxf(hz - (float)(k0 + 1), 0.f);
            const int j1 = min(j0 + 1, nn - 1), k1 = min(k0 + 1, g.nz - 1);
            const float* P = base + p * sP * nC;
            float v00, v01, v10, v11;
            if (OFF32) {
                const unsigned sj4 = (unsigned)(sJ * nC * 4), sk4 = (unsigned)(nC * 4);
                const unsigned oj0 = __umul24((unsigned)j0, sj4), oj1 = __umul24((unsigned)j1, sj4);
                const unsigned ok0 = __umul24((unsigned)k0, sk4), ok1 = __umul24((unsigned)k1, sk4);
                const char* Pc = reinterpret_cast<const char*>(P);
                v00 = *reinterpret_cast<const float*>(Pc + (oj0 + ok0));
                v01 = *reinterpret_cast<const float*>(Pc + (oj0 + ok1));
                v10 = *reinterpret_cast<const float*>(Pc + (oj1 + ok0));
                v11 = *reinterpret_cast<const float*>(Pc + (oj1 + ok1));
            } else {
                v00 = P[(j0 * sJ + k0) * nC];
                v01 = P[(j0 * sJ + k1) * nC];
                v10 = P[(j1 * sJ + k0) * nC];
                v11 = P[(j1 * sJ + k1) * nC];
            }
            const float sum = fmaf(fmaf(v01, oz1, v00 * oz0), oe0, fmaf(v11, oz1, v10 * oz0) * oe1);
            acc = fmaf(sum, __builtin_amdgcn_rcpf(area), acc);
        }
        acc *= r.L;
    } else if (r.ok) {
        for (int p = 0; p < nm; ++p) {
            const float fp = (float)p;
            float e1 = fmaf(r.b[0], fp, r.a[0]), e2 = fmaf(r.b[1], fp, r.a[1]);
            float z1 = fmaf(r.b[2], fp, r.a[2]), z2 = fmaf(r.b[3], fp, r.a[3]);
            if (e1 > e2) { const float t = e1; e1 = e2; e2 = t; }
            if (z1 > z2) { const float t = z1; z1 = z2; z2 = t; }
            const float area = (e2 - e1) * (z2 - z1);
            const float le = fmaxf(e1, 0.f), he = fminf(e2, (float)nn);
            const float lz = fmaxf(z1, 0.f), hz = fminf(z2, (float)g.nz);
            if (!(area > 0.f) || !(he > le) || !(hz > lz)) continue;
            const float* P = base + p * sP * nC;
            const int j0 = (int)le, j1 = min((int)he, nn - 1), k0 = (int)lz, k1 = min((int)hz, g.nz - 1);
            float sum = 0.f;
            for (int j = j0; j <= j1; ++j) {
                const float oe = fminf(he, (float)(j + 1)) - fmaxf(le, (float)j);
                if (oe <= 0.f) continue;
                float row = 0.f;
                for (int k = k0; k <= k1; ++k) {
                    const float oz = fminf(hz, (float)(k + 1)) - fmaxf(lz, (float)k);
                    row = fmaf(P[(j * sJ + k) * nC], fmaxf(oz, 0.f), row);
                }
                sum = fmaf(row, oe, sum);
            }
            acc += sum / area;
        }
        acc *= r.L;
    }
    prj[((((long)b * a.nu + iu) * a.nview + view) * a.nv + iv) * nC + c] = acc;
}

// Tables of the table-driven cone DD adjoint (k_dd_cone_adj_tab).
// Row sums: per (slice, view, detector column) the nv + 1 FP64 sums
// G[r] = sum over rows below r of len(row) q(row), len the length of the ray
// through the cell centre (the adjoint's row weight).
// ConeView: per view the constants of the plane-crossing geometry, resolved
// for the view's main axis m (x or y; n the other in-plane axis, p the plane
// index along m and j the voxel index along n).
struct alignas(16) ConeView {
    int xv;             // 1: main axis x
    float mp0, dm;      // plane p's main-axis offset from the source: mp0 + p dm
    float qn0, dn;      // n edge j relative to the source: qn0 + j dn
    float A, Cc, Um, Un;// detector centre minus source (n, m) and detU (m, n)
    float o;            // n index of the point on the plane at the source's n
    float a0;           // z index of the source
    float pad;
};
// ConeColumn: per (view, detector column) the ray constants that do not
// depend on the plane: the n index of the column's two u-edge rays on plane
// mp is mp k_e + o; the z index of row 0's lower edge on plane mp is
// mp A1 + a0; a row spans mp / T3 in z index; W the column weight dm /
// |main-axis run to the detector|, signed with the rows' z direction.
// Columns that cannot be hit (an edge ray parallel to the planes) hold
// k0 = k1 = 0 (an empty footprint) and W = 0.
struct alignas(32) ConeColumn { float k0, k1, A1, T3, W, pad[3]; };

// Row sums: one wave per (slice, view, column), in table order; lanes take
// 64 rows at a time, an FP64 inclusive scan across the wave, then a carry.
__global__ __launch_bounds__(256) void k_dd_cone_rowsum(ArbGeo a, const float* __restrict__ prj, const float* Cp,
                                                        const float* Up, const float* Vp, const float* Sp, int nC,
                                                        long ncols, double* __restrict__ T)
{
    const long col = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= ncols) return;
    const int lane = (int)(threadIdx.x & 63);
    const int iu = (int)(col % a.nu);
    const long rest = col / a.nu;
    const int view = (int)(rest % a.nview);
    const long bc = rest / a.nview;
    const int cc = (int)(bc % nC), b = (int)(bc / nC);
    const F3 c = ld3(Cp, view), U = ld3(Up, view), V = ld3(Vp, view), s = ld3(Sp, view);
    const float cu = (a.nu - 1) / 2.0f + a.off_u, cv = (a.nv - 1) / 2.0f + a.off_v;
    const float uc = ((float)iu - cu) * a.du;
    const float lx = fmaf(uc, U.x, c.x) - s.x, ly = fmaf(uc, U.y, c.y) - s.y;
    const float lxy2 = fmaf(lx, lx, ly * ly);
    const float zc0 = fmaf(uc, U.z, c.z) - s.z;
    const float lzs = V.z * a.dv, lz_0 = fmaf(-cv, lzs, zc0);
    const float* qcol = prj + (((long)b * a.nu + iu) * a.nview + view) * a.nv * nC + cc;
    double* G = T + col * (a.nv + 1);
    if (lane == 0) G[0] = 0.0;
    double carry = 0.0;
    for (int r0 = 0; r0 < a.nv; r0 += 64) {
        const int iv = r0 + lane;
        double f = 0.0;
        if (iv < a.nv) {
            const float lz = fmaf(lzs, (float)iv, lz_0);
            f = (double)(sqrtf(fmaf(lz, lz, lxy2)) * qcol[(long)iv * nC]);
        }
        #pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(f, d);
            if (lane >= d) f += y;
        }
        if (iv < a.nv) G[iv + 1] = carry + f;
        carry += __shfl(f, 63);
    }
}

// Exact transpose of k_dd_cone_fp: voxel (ix, iy, iz) gathers, per view, the
// cells whose footprint on its plane overlaps it, with the forward pass's
// weight overlap / area * L.  A wave holds one (ix, iy) column of 64 voxels
// along z, so the view's plane, the candidate detector columns and each
// column's in-plane footprint [e1, e2] (the rays through its u edges, which
// do not depend on v because the detector rows run along z) are
// wave-uniform.  Candidate columns come from inverting the plane crossing of
// the u-edge rays, u = (A - q Cc) / (q D - B) with q the in-plane slope of the
// voxel edge.  Along a column the rows' z edges on plane p are affine in v,
// z(v) = alpha + beta v (the forward pass's z-edge crossings), which gives
// each lane its candidate rows and every row its footprint with a few FMAs.
// Reference layouts.
template <int ZPL>
__global__ __launch_bounds__(256) void k_dd_cone_adj(ArbGeo a, Grid3K g, const float* __restrict__ prj,
                                                     float* __restrict__ img, const float* Cp, const float* Up,
                                                     const float* Vp, const float* Sp, const int* __restrict__ cls,
                                                     int nC)
{
    const int iz = (blockIdx.x * 64 + (threadIdx.x & 63)) * ZPL;
    const int iy = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (int)(threadIdx.x >> 6));
    const int ix = blockIdx.z % g.nx, bc = blockIdx.z / g.nx;
    if (iz >= g.nz || iy >= g.ny) return;
    const int cc = bc % nC, b = bc / nC;
    const float* q = prj + (long)b * a.nu * a.nview * a.nv * nC + cc;
    const float cu = (a.nu - 1) / 2.0f + a.off_u, cv = (a.nv - 1) / 2.0f + a.off_v;
    const float margin = 0.05f;
    const float fz = (float)iz;
    const float idu = 1.0f / a.du, idv = 1.0f / a.dv, idz = 1.0f / g.dz;
    float acc[ZPL];
    #pragma unroll
    for (int k = 0; k < ZPL; ++k) acc[k] = 0.f;
    for (int view = 0; view < a.nview; ++view) {
        const bool xv = cls[view] == 0;
        const F3 c = ld3(Cp, view), U = ld3(Up, view), V = ld3(Vp, view), s = ld3(Sp, view);
        const int p = xv ? ix : iy, j = xv ? iy : ix;
        const int nm = xv ? g.nx : g.ny, nn = xv ? g.ny : g.nx;
        const float dm = xv ? g.dx : g.dy, dn = xv ? g.dy : g.dx;
        const float idn = 1.0f / dn;
        const float sm = xv ? s.x : s.y, sn = xv ? s.y : s.x;
        const float cn = xv ? g.cy : g.cx;
        const float cm = xv ? c.x : c.y, cnd = xv ? c.y : c.x, Um = xv ? U.x : U.y, Un = xv ? U.y : U.x;
        // plane p, relative to the source along the main axis
        const float mp = ((float)p + 0.5f - 0.5f * nm) * dm + (xv ? g.cx : g.cy) - sm;
        const float imp = __builtin_amdgcn_rcpf(mp);
        // u of the rays through the voxel's two in-plane edges
        const float cn0 = cn - 0.5f * nn * dn;
        const float A = cnd - sn, Cc = cm - sm;
        float ulo = 3.0e38f, uhi = -3.0e38f;
        #pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float qe = (cn0 + (float)(j + e) * dn - sn) * imp;
            const float u = (A - qe * Cc) * __builtin_amdgcn_rcpf(qe * Um - Un);
            ulo = fminf(ulo, u);
            uhi = fmaxf(uhi, u);
        }
        const float flo = fmaxf(ulo * idu + cu - 0.5f - margin, -1.0f);
        const float fhi = fminf(uhi * idu + cu + 0.5f + margin, (float)a.nu);
        if (!(fhi >= flo)) continue;                // also rejects NaN
        const int u0 = max((int)ceilf(flo), 0), u1 = min((int)floorf(fhi), a.nu - 1);
        for (int iu = u0; iu <= u1; ++iu) {
            const float uc = ((float)iu - cu) * a.du;
            // in-plane footprint of the column on plane p (index coordinates)
            bool ok = true;
            float ee[2];
            #pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float ue = fmaf(e ? 0.5f : -0.5f, a.du, uc);
                const float em = fmaf(ue, Um, cm) - sm, eo = fmaf(ue, Un, cnd) - sn;
                ok = ok && fabsf(em) > 1e-6f;
                ee[e] = (fmaf(mp, eo * __builtin_amdgcn_rcpf(em), sn) - cn) * idn + 0.5f * nn;
            }
            const float e1 = fminf(ee[0], ee[1]), e2 = fmaxf(ee[0], ee[1]);
            const float oe = fminf(e2, (float)(j + 1)) - fmaxf(e1, (float)j);
            const float dpc = fmaf(uc, Um, cm) - sm;     // centre ray's main-axis run to the detector
            if (!ok || !(fabsf(dpc) > 1e-6f) || !(oe > 0.f) || !(e2 > e1)) continue;
            const float idpc = __builtin_amdgcn_rcpf(dpc);
            const float lx = fmaf(uc, U.x, c.x) - s.x, ly = fmaf(uc, U.y, c.y) - s.y;
            const float lxy2 = fmaf(lx, lx, ly * ly);
            const float zc0 = fmaf(uc, U.z, c.z) - s.z;
            // z edges of a row at detector height v: alpha + beta v (index
            // coordinates); every row's footprint is |beta| dv high
            const float alpha = (fmaf(mp, zc0 * idpc, s.z) - g.cz) * idz + 0.5f * g.nz;
            const float beta = mp * V.z * idpc * idz;
            if (!(fabsf(beta) > 0.f)) continue;
            const float ib = __builtin_amdgcn_rcpf(beta);
            const float va = (fz - alpha) * ib, vb = (fz + (float)ZPL - alpha) * ib;
            const float glo = fmaxf(fminf(va, vb) * idv + cv - 0.5f - margin, -1.0f);
            const float ghi = fminf(fmaxf(va, vb) * idv + cv + 0.5f + margin, (float)a.nv);
            if (!(ghi >= glo)) continue;
            const int v0 = max((int)ceilf(glo), 0), v1 = min((int)floorf(ghi), a.nv - 1);
            // oe / (e2 - e1) / (|beta| dv) * dm / |dpc|: every factor of the weight but oz and the length
            const float wcol = oe * __builtin_amdgcn_rcpf((e2 - e1) * fabsf(beta) * a.dv) * dm * fabsf(idpc);
            // the row's z edges and the ray's z run are affine in the row
            // index: z1 = z1_0 + beta dv iv, z2 = z1 + |beta| dv, lz = lz_0 + V.z dv iv
            const float bdv = beta * a.dv, hb = 0.5f * a.dv;
            const float z1_0 = fmaf(beta, -cv * a.dv - hb, alpha) + fminf(bdv, 0.f), zh = fabsf(bdv);
            const float lzs = V.z * a.dv, lz_0 = fmaf(-cv, lzs, zc0);
            const float* qcol = q + ((long)iu * a.nview + view) * a.nv * nC;
            for (int iv = v0; iv <= v1; ++iv) {
                const float fiv = (float)iv;
                const float z1 = fmaf(bdv, fiv, z1_0), z2 = z1 + zh;
                const float lz = fmaf(lzs, fiv, lz_0);
                const float wv = wcol * __builtin_amdgcn_sqrtf(fmaf(lz, lz, lxy2)) * qcol[(long)iv * nC];
                #pragma unroll
                for (int k = 0; k < ZPL; ++k) {
                    const float oz = fminf(z2, fz + (float)(k + 1)) - fmaxf(z1, fz + (float)k);
                    if (oz > 0.f) acc[k] = fmaf(oz, wv, acc[k]);
                }
            }
        }
    }
    float* out = img + ((((long)b * g.nx + ix) * g.ny + iy) * g.nz + iz) * nC + cc;
    #pragma unroll
    for (int k = 0; k < ZPL; ++k)
        if (iz + k < g.nz) out[(long)k * nC] = acc[k];
}

// k_dd_cone_adj driven by tables: the rows' z footprints on a column's
// plane tile z without gaps (row r covers [zA + r bdv, zA + (r + 1) bdv],
// bdv = beta dv), so the sum over rows of the z overlap times len q is bdv
// times the difference of the row sums, interpolated linearly within a row,
// at the voxel's two z edges.  ZPL + 1 lookups of one 16-byte pair
// (G[r], G[r + 1]) give ZPL voxels; the per-view and per-column geometry
// comes from the ConeView / ConeColumn tables (wave-uniform loads).  Same
// weights as the forward pass (the transpose holds to rounding).  A NaN or
// infinity in a row reaches every voxel above that row on the column's
// plane, not only the voxels the row overlaps.
template <int ZPL>
__global__ __launch_bounds__(256) void k_dd_cone_adj_tab(ArbGeo a, Grid3K g, const double* __restrict__ T,
                                                         float* __restrict__ img, const ConeView* __restrict__ views,
                                                         const ConeColumn* __restrict__ cols, int nC)
{
    const int iz = (blockIdx.x * 64 + (threadIdx.x & 63)) * ZPL;
    const int iy = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (int)(threadIdx.x >> 6));
    const int ix = blockIdx.z % g.nx, bc = blockIdx.z / g.nx;
    if (iz >= g.nz || iy >= g.ny) return;
    const int cc = bc % nC, b = bc / nC;
    const float cu = (a.nu - 1) / 2.0f + a.off_u;
    const float margin = 0.05f;
    const float fz = (float)iz, fnv = (float)a.nv;
    const float idu = 1.0f / a.du;
    const double* Tb = T + (long)bc * a.nview * a.nu * (a.nv + 1);
    float acc[ZPL];
    #pragma unroll
    for (int k = 0; k < ZPL; ++k) acc[k] = 0.f;
    for (int view = 0; view < a.nview; ++view) {
        const ConeView vw = views[view];
        const double* Tv = Tb + (long)view * a.nu * (a.nv + 1);
        const int p = vw.xv ? ix : iy, j = vw.xv ? iy : ix;
        const float mp = fmaf((float)p, vw.dm, vw.mp0);
        const float imp = __builtin_amdgcn_rcpf(mp);
        // u of the rays through the voxel's two in-plane edges
        float ulo = 3.0e38f, uhi = -3.0e38f;
        #pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float qe = fmaf((float)(j + e), vw.dn, vw.qn0) * imp;
            const float u = fmaf(-qe, vw.Cc, vw.A) * __builtin_amdgcn_rcpf(fmaf(qe, vw.Um, -vw.Un));
            ulo = fminf(ulo, u);
            uhi = fmaxf(uhi, u);
        }
        const float flo = fmaxf(ulo * idu + cu - 0.5f - margin, -1.0f);
        const float fhi = fminf(uhi * idu + cu + 0.5f + margin, (float)a.nu);
        if (!(fhi >= flo)) continue;                // also rejects NaN
        const int u0 = max((int)ceilf(flo), 0), u1 = min((int)floorf(fhi), a.nu - 1);
        const float smp = mp > 0.f ? 1.f : -1.f;
        for (int iu = u0; iu <= u1; ++iu) {
            const ConeColumn col = cols[(long)view * a.nu + iu];
            // in-plane footprint of the column on plane p (index coordinates)
            const float ea = fmaf(mp, col.k0, vw.o), eb = fmaf(mp, col.k1, vw.o);
            const float e1 = fminf(ea, eb), e2 = fmaxf(ea, eb);
            const float oe = fminf(e2, (float)(j + 1)) - fmaxf(e1, (float)j);
            if (!(oe > 0.f) || !(e2 > e1)) continue;
            // row coordinate of z index zi: s = (zi - zA) / bdv = zi ibdv + s0
            const float ibdv = imp * col.T3;
            if (!(fabsf(ibdv) < 1e20f)) continue;
            const float s0 = -fmaf(mp, col.A1, vw.a0) * ibdv;
            const float sb = fmaf(fz, ibdv, s0), st = fmaf(fz + (float)ZPL, ibdv, s0);
            if (!(fmaxf(sb, st) > 0.f && fminf(sb, st) < fnv)) continue;
            // oe / (e2 - e1) / |bdv| * dm / |dpc| times bdv
            const float wcol = oe * __builtin_amdgcn_rcpf(e2 - e1) * col.W * smp;
            // one 16-byte load per edge: G[r] and G[r + 1] (the address
            // processing of these loads bounds the kernel, so lanes whose
            // voxels miss the rows load nothing)
            const double* Gc = Tv + (long)iu * (a.nv + 1);
            double hp = 0.0;
            #pragma unroll
            for (int k = 0; k <= ZPL; ++k) {
                const float sk = __builtin_amdgcn_fmed3f(fmaf(fz + (float)k, ibdv, s0), 0.f, fnv);
                const unsigned r = min((unsigned)sk, (unsigned)(a.nv - 1));
                const double* gp = Gc + r;
                const double g0 = gp[0], g1 = gp[1];
                const double h = fma((double)(sk - (float)r), g1 - g0, g0);
                if (k > 0) acc[k - 1] = fmaf(wcol, (float)(h - hp), acc[k - 1]);
                hp = h;
            }
        }
    }
    float* out = img + ((((long)b * g.nx + ix) * g.ny + iy) * g.nz + iz) * nC + cc;
    #pragma unroll
    for (int k = 0; k < ZPL; ++k)
        if (iz + k < g.nz) out[(long)k * nC] = acc[k];
}

// ---- cone OS-SQS helpers ----

__global__ void k_cone_residual(const float* __restrict__ fp, const float* __restrict__ prj, float* __restrict__ r,
                                int nB, int nu, int nview, int cur, int start, int nvc)
{
    // fp, r: [b][nu][cur][nv*nC]; prj: [b][nu][nview][nv*nC]
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)nu * cur * nvc;
    if (i >= per * nB) return;
    const long b = i / per, rem = i % per;
    const long iu = rem / ((long)cur * nvc), rem2 = rem % ((long)cur * nvc);
    const long vi = rem2 / nvc, k = rem2 % nvc;
    r[i] = fp[i] - prj[((b * nu + iu) * nview + start + vi) * nvc + k];
}

__global__ void k_cone_update(float* __restrict__ x, const float* __restrict__ bp, const float* __restrict__ D,
                              const float* __restrict__ mask, float step, float lb, float ub, int crop, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = (x[i] - step * (bp[i] / (D[i] + 1e-6f))) * mask[i];
    if (crop) v = v < lb ? lb : (v > ub ? ub : v);
    x[i] = v;
}

// ---- helical ----
//
// Weighted helical backprojection of parallel-cone data
// (bpHelicalParallelConebeam, Recon3D/HelicalBP.cu:10-83).  For a voxel at
// height z the views of one pi-window, i = iStart(z) .. iStart(z) + nviewPerPI - 1,
// each combine their conjugates k = i + m nviewPerPI (|m| <= mPI) as
// sum_m w_k P_k / (sum_m w_k + 1e-6), with P_k the bilinear sample at
//   u = x sin t_k - y cos t_k,
//   v = (z - zs_k(u)) dsd / (sqrt(dso^2 - u^2) - (x cos t_k + y sin t_k)),
//   zs_k(u) = zrot (k / (2 nviewPerPI) - asin(u / dso) / (2 pi)),
// and w_k the row taper of q = |2 v / (dv nv)|.
//
// Everything about a view except v is independent of z: a lane owns a
// column of ZC voxels (x, y fixed) and walks the union of their pi-windows;
// for each view it evaluates t_k, u, the magnification dsd / (...), zs_k and
// the column interpolation once, and then each of its voxels whose window
// contains the view costs one FMA for v, the taper and a bilinear fetch.  A
// wave takes an 8 x 8 block of columns (its fetches land in one detector
// neighbourhood).  The sums run in the order above (i ascending, m
// ascending), the window start in the reference's float arithmetic.
// ---------------------------------------------------------------------------
// prj_k: [slice][nview][nv][nu]
// per view k: (sin t_k, cos t_k, zk = zrot k / (2 nviewPerPI) as a float pair
// hi, lo), t_k = theta0 + k PI / nviewPerPI, in FP64
__global__ void k_helical_views(HelicalGeo h, float4* __restrict__ vt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= h.nview) return;
    const double PI = (double)3.141592657f;
    const double t = (double)h.theta0 + k * PI / h.nviewPerPI;
    const double zk = (double)h.zrot * k / (2.0 * h.nviewPerPI);
    const float zhi = (float)zk;
    vt[k] = make_float4((float)sin(t), (float)cos(t), zhi, (float)(zk - zhi));
}

// asin by its Maclaurin series, sum_k C(2k, k) / (4^k (2k + 1)) x^(2k + 1),
// through x^15 for |x| <= 0.35 (truncation below 3e-10), asinf beyond
__device__ __forceinline__ float asin_series(float x)
{
    if (!(fabsf(x) <= 0.35f)) return asinf(x);
    const float x2 = x * x;
    float p = 1.396484375e-02f;
    p = fmaf(p, x2, 1.735276442e-02f);
    p = fmaf(p, x2, 2.237215909e-02f);
    p = fmaf(p, x2, 3.038194444e-02f);
    p = fmaf(p, x2, 4.464285714e-02f);
    p = fmaf(p, x2, 7.500000000e-02f);
    p = fmaf(p, x2, 1.666666667e-01f);
    p = fmaf(p, x2, 1.000000000e+00f);
    return x * p;
}

// two adjacent floats at a 4-byte-aligned address (one 8-byte load)
typedef float PairF __attribute__((ext_vector_type(2), aligned(4)));

template <int ZC>
__global__ __launch_bounds__(256) void k_helical_bp(HelicalGeo h, const float* __restrict__ prj,
                                                    const float4* __restrict__ vt, float* __restrict__ img, int nC,
                                                    int nzc)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int iy = blockIdx.x * 16 + (wv & 1) * 8 + (lane & 7);
    const int ix = blockIdx.y * 16 + (wv >> 1) * 8 + (lane >> 3);
    const int zb = blockIdx.z % nzc, bc = blockIdx.z / nzc;
    const int c = bc % nC, b = bc / nC;
    if (ix >= h.nx || iy >= h.ny) return;
    const float PI = 3.141592657f;                      // the reference's PI (Projector.h)
    const float x = (ix - (h.nx - 1) / 2.0f) * h.dx, y = (iy - (h.ny - 1) / 2.0f) * h.dy;
    const int z0 = zb * ZC;
    float zc[ZC];
    int is[ZC];
    int iLo = 0x7fffffff, iHi = -0x7fffffff;
    #pragma unroll
    for (int j = 0; j < ZC; ++j) {
        // the reference's float arithmetic, multiply-adds as FMAs (nvcc's
        // default contraction): the start is often an integer in exact
        // arithmetic, so its truncation follows these roundings
        zc[j] = fmaf((float)(z0 + j), h.dz, h.volZ0);
        is[j] = (int)(fmaf(zc[j] / h.zrot * 2.0f, PI, -PI / 4.0f) * (float)h.nviewPerPI / PI);
        if (z0 + j < h.nz) {
            iLo = min(iLo, is[j]);
            iHi = max(iHi, is[j] + h.nviewPerPI);
        }
    }
    // only views with a conjugate inside the data can contribute
    iLo = max(iLo, -h.mPI * h.nviewPerPI);
    iHi = min(iHi, h.nview + h.mPI * h.nviewPerPI);
    const float* pb = prj + (long)(b * nC + c) * h.nview * h.nv * h.nu;
    const float zr2pi = h.zrot / (2 * PI);
    const float qs = 2.0f / (h.dv * h.nv);
    const float cu = (h.nu - 1.0f) / 2.0f + h.off_u, cv = (h.nv - 1.0f) / 2.0f + h.off_v;
    const float idso = 1.0f / h.dso, idu = 1.0f / h.du, idv = 1.0f / h.dv;
    // the taper's phase in revolutions: cos((q - Q) / (1 - Q) pi / 2) = cos(2 pi r),
    // r = (q - Q) / (4 (1 - Q)), by the hardware cosine (taper weights enter as
    // sum w P / sum w; its error is far below the parity bound)
    const float tq = 0.25f / (1.0f - h.Q);
    float acc[ZC];
    #pragma unroll
    for (int j = 0; j < ZC; ++j) acc[j] = 0.0f;
    for (int i = iLo; i < iHi; ++i) {
        float val[ZC], tw[ZC];
        #pragma unroll
        for (int j = 0; j < ZC; ++j) val[j] = tw[j] = 0.0f;
        for (int m = -h.mPI; m <= h.mPI; ++m) {
            const int k = i + m * h.nviewPerPI;
            if (k < 0 || k >= h.nview) continue;
            const float4 tv = vt[k];                          // k is wave-uniform: scalar loads
            const float st = tv.x, ct = tv.y;
            const float u = x * st - y * ct;
            const float mag = h.dsd * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(h.dso * h.dso - u * u) -
                                                            (x * ct + y * st));
            // source height zk - zr2pi asin(u / dso) as tv.z + zsl (tv.z exact in
            // float, so that z - tv.z loses nothing to the source's turn height)
            const float zsl = fmaf(-zr2pi, asin_series(u * idso), tv.w);
            const float pu = fmaf(u, idu, cu);
            const float fu = floorf(pu);
            const int u0 = (int)fu;
            const float au = pu - fu;
            const float wu0 = (u0 >= 0 && u0 < h.nu) ? 1.0f - au : 0.0f;
            const float wu1 = (u0 + 1 >= 0 && u0 + 1 < h.nu) ? au : 0.0f;
            // the two columns as one 8-byte load of the pair (ua, ua + 1),
            // ua = u0 clamped to [0, nu - 2]; the weights follow the columns
            // (u0 and u0 + 1 inside the pair where they fall in it)
            const int ua = min(max(u0, 0), h.nu - 2);
            const float wlo = (u0 == ua ? wu0 : 0.0f) + (u0 + 1 == ua ? wu1 : 0.0f);
            const float whi = (u0 == ua + 1 ? wu0 : 0.0f) + (u0 + 1 == ua + 1 ? wu1 : 0.0f);
            // the view's rows from a wave-uniform base and 32-bit lane
            // offsets; every voxel's two rows are loaded (rows clamped into
            // the detector) before any is used, and rows outside it are
            // dropped by selects
            const float* pk = pb + (long)k * h.nv * h.nu;
            float v0w[ZC], v1w[ZC], w[ZC];
            bool in0[ZC], in1[ZC];
            PairF r0[ZC], r1[ZC];
            #pragma unroll
            for (int j = 0; j < ZC; ++j) {
                const float v = ((zc[j] - tv.z) - zsl) * mag;
                const float q = fabsf(v * qs);
                float wj;
                if (q <= h.Q) wj = 1.0f;
                else if (q < 1.0f) { wj = __builtin_amdgcn_cosf((q - h.Q) * tq); wj *= wj; }
                else wj = 0.0f;
                w[j] = wj;
                const float pv = fmaf(v, idv, cv);
                const float fv = floorf(pv);
                const int v0 = (int)fv;
                const float av = pv - fv;
                in0[j] = v0 >= 0 && v0 < h.nv;
                in1[j] = v0 + 1 >= 0 && v0 + 1 < h.nv;
                v0w[j] = 1.0f - av;
                v1w[j] = av;
                const int ra = min(max(v0, 0), h.nv - 1), rb = min(max(v0 + 1, 0), h.nv - 1);
                r0[j] = *reinterpret_cast<const PairF*>(pk + (__umul24(ra, h.nu) + (unsigned)ua));
                r1[j] = *reinterpret_cast<const PairF*>(pk + (__umul24(rb, h.nu) + (unsigned)ua));
            }
            #pragma unroll
            for (int j = 0; j < ZC; ++j) {
                if (i < is[j] || i >= is[j] + h.nviewPerPI) continue;     // wave-uniform
                float smp = 0.0f;
                if (in0[j]) smp += v0w[j] * (wlo * r0[j].x + whi * r0[j].y);
                if (in1[j]) smp += v1w[j] * (wlo * r1[j].x + whi * r1[j].y);
                val[j] += smp * w[j];
                tw[j] += w[j];
            }
        }
        #pragma unroll
        for (int j = 0; j < ZC; ++j)
            if (i >= is[j] && i < is[j] + h.nviewPerPI)
                acc[j] = fmaf(val[j], __builtin_amdgcn_rcpf(tw[j] + 1e-6f), acc[j]);   // 1-ulp reciprocal
    }
    #pragma unroll
    for (int j = 0; j < ZC; ++j)
        if (z0 + j < h.nz) img[((((long)b * h.nx + ix) * h.ny + iy) * h.nz + z0 + j) * nC + c] = acc[j];
}

Grid3K grid_of(const ArbGeo& g) { return Grid3K{g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.cx, g.cy, g.cz}; }

} // namespace

bool arb_rows_along_z(const float* detU, const float* detV, int nview)
{
    if (!detU || !detV) return false;
    for (int v = 0; v < nview; v++)
        if (detV[3 * v] != 0.f || detV[3 * v + 1] != 0.f || detU[3 * v + 2] != 0.f) return false;
    return true;
}

void siddon_arb_fp(const ArbGeo& g, bool parallel, int nB, int nC, const float* img, float* prj,
                   const float* detCenter, const float* detU, const float* detV, const float* srcOrDir,
                   DevBuf& work, hipStream_t st)
{
    (void)work;
    RN_REQUIRE(detCenter && detU && detV && srcOrDir, "null geometry array");
    RN_REQUIRE((long)g.nview * nB * nC <= 65535 && cdiv(g.nu, 4) <= 65535, "siddon_arb_fp: grid too large");
    dim3 blk(256), grd(cdiv(g.nv, 64), cdiv(g.nu, 4), g.nview * nB * nC);
    // 32-bit voxel offsets when one slice's volume stays below 4 GiB
    auto k = (long)g.nx * g.ny * g.nz * nC * 4 < (1L << 32) ? k_siddon_arb_fp<true> : k_siddon_arb_fp<false>;
    hipLaunchKernelGGL(k, grd, blk, 0, st, g, grid_of(g), parallel ? 1 : 0, img, prj, detCenter, detU, detV, srcOrDir,
                       nC);
    check_launch("siddon_arb_fp");
}

void siddon_arb_bp(const ArbGeo& g, bool parallel, int nB, int nC, const float* prj, float* img,
                   const float* detCenter, const float* detU, const float* detV, const float* srcOrDir,
                   DevBuf& work, hipStream_t st)
{
    RN_REQUIRE(detCenter && detU && detV && srcOrDir, "null geometry array");
    RN_REQUIRE((long)g.nx * nB * nC <= 65535 && cdiv(g.ny, 4) <= 65535, "siddon_arb_bp: grid too large");
    work.ensure((size_t)g.nview * 4 * sizeof(float4));
    float4* vt = work.as<float4>();
    hipLaunchKernelGGL(k_siddon_arb_views, dim3(cdiv(g.nview, 256)), dim3(256), 0, st, g, grid_of(g),
                       parallel ? 1 : 0, detCenter, detU, detV, srcOrDir, vt);
    check_launch("siddon_arb_views");
    if (parallel || !g.rows_z) {
        dim3 grd(cdiv(g.nz, 64), cdiv(g.ny, 4), g.nx * nB * nC);
        hipLaunchKernelGGL((k_siddon_arb_bp<false, 1>), grd, dim3(256), 0, st, g, grid_of(g), parallel ? 1 : 0, prj, img,
                           detCenter, detU, detV, srcOrDir, vt, nC);
        check_launch("siddon_arb_bp");
        return;
    }
    // upright cone views: 1, 2 or 4 voxels per lane (the launch-shape tuner)
    auto launch = [&](int c) {
        const int zpl = 1 << c;
        dim3 grd(cdiv(g.nz, 64 * zpl), cdiv(g.ny, 4), g.nx * nB * nC);
        auto k = c == 0 ? k_siddon_arb_bp<true, 1> : c == 1 ? k_siddon_arb_bp<true, 2> : k_siddon_arb_bp<true, 4>;
        hipLaunchKernelGGL(k, grd, dim3(256), 0, st, g, grid_of(g), 0, prj, img, detCenter, detU, detV, srcOrDir, vt, nC);
        check_launch("siddon_arb_bp");
    };
    const int c = tune_choose({TK_SIDDON_CONE_BP, g.nx * g.ny, g.nz, g.nu * g.nv, g.nview, nB * nC}, 3, launch, st, true);
    launch(c);
}

void dd_tomo_fp(const ArbGeo& g, int nB, int nC, const float* img, float* prj,
                const float* detCenter, const float* src, DevBuf& work, hipStream_t st)
{
    RN_REQUIRE(detCenter && src, "null geometry array");
    const size_t nvol = (size_t)g.nx * g.ny * g.nz;
    const int ns = nB * nC;
    work.ensure(2 * ns * nvol * sizeof(float));
    float* t1 = work.as<float>();
    float* t2 = t1 + ns * nvol;
    // [b][x][y][z][c] -> [(b,c)][x][y][z] -> [(b,c)][z][x][y] -> [(b,c)][z][y][x]
    const float* base = img;
    if (nC > 1) {
        launch_transpose(img, t2, (int)nvol, nC, nB, st);
        base = t2;
    }
    launch_transpose(base, t1, g.nx * g.ny, g.nz, ns, st);
    launch_transpose(t1, t2, g.nx, g.ny, ns * g.nz, st);
    dim3 blk(128), grd(cdiv(g.nu, 128), g.nview, nB * g.nv * nC);
    hipLaunchKernelGGL(k_dd_tomo_fp, grd, blk, 0, st, g, grid_of(g), t2, prj, detCenter, src, nC);
    check_launch("dd_tomo_fp");
}

void dd_tomo_bp(const ArbGeo& g, int nB, int nC, const float* prj, float* img,
                const float* detCenter, const float* src, DevBuf& work, hipStream_t st)
{
    RN_REQUIRE(detCenter && src, "null geometry array");
    const size_t nprj = (size_t)nB * nC * g.nu * g.nview * g.nv;
    work.ensure(nprj * sizeof(float));
    dim3 b1(256), g1(cdiv(g.nu, 256), g.nview, nB * g.nv * nC);
    hipLaunchKernelGGL(k_dd_tomo_preweight, g1, b1, 0, st, g, prj, work.as<float>(), detCenter, src, nC);
    check_launch("dd_tomo_preweight");
    dim3 blk(16, 16), grd(cdiv(g.ny, 16), cdiv(g.nx, 16), nB * nC * g.nz);
    hipLaunchKernelGGL(k_dd_tomo_bp, grd, blk, 0, st, g, grid_of(g), work.as<float>(), img, detCenter, src, nC);
    check_launch("dd_tomo_bp");
}

void ConePlan::setup(const ArbGeo& g, const float* C, const float* U, const float* V, const float* S, hipStream_t st)
{
    RN_REQUIRE(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.nu > 0 && g.nview > 0 && g.nv > 0, "cone dimensions must be positive");
    RN_REQUIRE(g.dx > 0 && g.dy > 0 && g.dz > 0 && g.du > 0 && g.dv > 0, "cone spacings must be positive");
    RN_REQUIRE(C && U && V && S, "null geometry array");
    const int n = g.nview;
    std::vector<float> h((size_t)12 * n);
    std::vector<int> cls(n);
    for (int v = 0; v < n; v++) {
        const float* u = U + 3 * v;
        const float* w = V + 3 * v;
        const float un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const float wn = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        RN_REQUIRE(un > 0 && wn > 0, "detector axes must be non-zero");
        RN_REQUIRE(std::fabs(w[0]) <= 1e-6f * wn && std::fabs(w[1]) <= 1e-6f * wn,
                   "the cone DD projector needs detector columns parallel to z (detV = (0, 0, v))");
        RN_REQUIRE(std::fabs(u[2]) <= 1e-6f * un, "the cone DD projector needs detector rows normal to z (detU.z = 0)");
        cls[v] = std::fabs(C[3 * v] - S[3 * v]) >= std::fabs(C[3 * v + 1] - S[3 * v + 1]) ? 0 : 1;
    }
    std::copy(C, C + 3 * n, h.begin());
    std::copy(U, U + 3 * n, h.begin() + 3 * n);
    std::copy(V, V + 3 * n, h.begin() + 6 * n);
    std::copy(S, S + 3 * n, h.begin() + 9 * n);
    vec_.alloc(h.size() * sizeof(float));
    cls_.alloc(cls.size() * sizeof(int));
    RN_HIP_CHECK(hipMemcpyAsync(vec_.get(), h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipMemcpyAsync(cls_.get(), cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice, st));
    // view and column tables of the table-driven DD adjoint, in double
    std::vector<ConeView> vt(n);
    std::vector<ConeColumn> ct((size_t)n * g.nu);
    const double cu = (g.nu - 1) / 2.0 + g.off_u, cv = (g.nv - 1) / 2.0 + g.off_v;
    for (int v = 0; v < n; v++) {
        const int xv = cls[v] == 0;
        const double Cm = C[3 * v + (xv ? 0 : 1)], Cn = C[3 * v + (xv ? 1 : 0)], Cz = C[3 * v + 2];
        const double Sm = S[3 * v + (xv ? 0 : 1)], Sn = S[3 * v + (xv ? 1 : 0)], Sz = S[3 * v + 2];
        const double Um = U[3 * v + (xv ? 0 : 1)], Un = U[3 * v + (xv ? 1 : 0)], Uz = U[3 * v + 2];
        const double Vz = V[3 * v + 2];
        const int nm = xv ? g.nx : g.ny, nn = xv ? g.ny : g.nx;
        const double dm = xv ? g.dx : g.dy, dn = xv ? g.dy : g.dx;
        const double gm = xv ? g.cx : g.cy, gn = xv ? g.cy : g.cx;
        ConeView& w = vt[v];
        w.xv = xv;
        w.mp0 = (float)((0.5 - 0.5 * nm) * dm + gm - Sm);
        w.dm = (float)dm;
        w.qn0 = (float)(gn - 0.5 * nn * dn - Sn);
        w.dn = (float)dn;
        w.A = (float)(Cn - Sn);
        w.Cc = (float)(Cm - Sm);
        w.Um = (float)Um;
        w.Un = (float)Un;
        w.o = (float)((Sn - gn) / dn + 0.5 * nn);
        w.a0 = (float)((Sz - g.cz) / g.dz + 0.5 * g.nz);
        w.pad = 0.f;
        for (int iu = 0; iu < g.nu; iu++) {
            ConeColumn& k = ct[(size_t)v * g.nu + iu];
            k = ConeColumn{0.f, 0.f, 0.f, 0.f, 0.f, {0.f, 0.f, 0.f}};
            const double uc = (iu - cu) * g.du;
            const double e0m = (uc - 0.5 * g.du) * Um + Cm - Sm, e0n = (uc - 0.5 * g.du) * Un + Cn - Sn;
            const double e1m = (uc + 0.5 * g.du) * Um + Cm - Sm, e1n = (uc + 0.5 * g.du) * Un + Cn - Sn;
            const double dpc = uc * Um + Cm - Sm;
            if (!(std::fabs(e0m) > 1e-6) || !(std::fabs(e1m) > 1e-6) || !(std::fabs(dpc) > 1e-6)) continue;
            const double zr0 = uc * Uz + Cz - Sz - (cv + 0.5) * g.dv * Vz;   // row 0's lower edge, from the source
            k.k0 = (float)(e0n / e0m / dn);
            k.k1 = (float)(e1n / e1m / dn);
            k.A1 = (float)(zr0 / dpc / g.dz);
            k.T3 = (float)(dpc * g.dz / (Vz * g.dv));
            k.W = (float)(dm / std::fabs(dpc) * (Vz / dpc > 0 ? 1.0 : -1.0));
        }
    }
    views_.alloc(vt.size() * sizeof(ConeView));
    cols_.alloc(ct.size() * sizeof(ConeColumn));
    RN_HIP_CHECK(hipMemcpyAsync(views_.get(), vt.data(), vt.size() * sizeof(ConeView), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipMemcpyAsync(cols_.get(), ct.data(), ct.size() * sizeof(ConeColumn), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    geo = g;
    geo.rows_z = arb_rows_along_z(U, V, n) ? 1 : 0;
}

void ConePlan::dd_fp(int nB, int nC, const float* img, float* prj, hipStream_t st) const
{
    const ArbGeo& g = geo;
    RN_REQUIRE(nB > 0 && nC > 0 && (long)g.nview * nB * nC <= 65535, "cone batch too large for one launch");
    const float* v = vec_.as<float>();
    const long nblocks = (long)cdiv(g.nv, 64) * g.nu * g.nview * nB * nC;
    RN_REQUIRE(nblocks + 8 < 0x7fffffffL, "dd_cone_fp: grid too large");
    // 32-bit cell offsets when every byte stride fits 24 bits and a slice's
    // volume 4 GiB
    const bool off32 = (long)std::max(g.ny, g.nx) * g.nz * nC * 4 < (1L << 24) &&
                       (long)g.nx * g.ny * g.nz * nC * 4 < (1L << 32);
    auto launch = [&](int c) {
        if (c == 0) {
            const long n8 = (nblocks + 7) / 8 * 8;
            auto k = off32 ? k_dd_cone_fp<8, true> : k_dd_cone_fp<8, false>;
            hipLaunchKernelGGL(k, dim3((unsigned)n8), dim3(64), 0, st, g, grid_of(g), img, prj, v, v + 3 * g.nview,
                               v + 6 * g.nview, v + 9 * g.nview, cls_.as<int>(), nC, (int)nblocks);
        } else {
            dim3 grd(cdiv(g.nv, 64), g.nu, g.nview * nB * nC);
            auto k = off32 ? k_dd_cone_fp<0, true> : k_dd_cone_fp<0, false>;
            hipLaunchKernelGGL(k, grd, dim3(64), 0, st, g, grid_of(g), img, prj, v, v + 3 * g.nview, v + 6 * g.nview,
                               v + 9 * g.nview, cls_.as<int>(), nC, (int)nblocks);
        }
        check_launch("dd_cone_fp");
    };
    const int c = tune_choose({TK_DD_CONE_FP, g.nx * g.ny, g.nz, g.nu * g.nv, g.nview, nB * nC}, 2, launch, st, true);
    launch(c);
}

void ConePlan::dd_adj(int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st) const
{
    const ArbGeo& g = geo;
    RN_REQUIRE(nB > 0 && nC > 0 && (long)g.nx * nB * nC <= 65535, "cone batch too large for one launch");
    const float* v = vec_.as<float>();
    // candidates 0-2: 1, 2 or 4 voxels per lane summing the rows one by one;
    // 3-5: 1, 2 or 4 voxels per lane from the row-sum table
    const long ncols = (long)nB * nC * g.nu * g.nview;
    RN_REQUIRE((ncols + 3) / 4 < 0x7fffffffL, "dd_cone_adj: too many detector columns");
    auto launch = [&](int c) {
        const int zpl = 1 << (c % 3);
        dim3 grd(cdiv(g.nz, 64 * zpl), cdiv(g.ny, 4), g.nx * nB * nC);
        if (c < 3) {
            auto k = c == 0 ? k_dd_cone_adj<1> : c == 1 ? k_dd_cone_adj<2> : k_dd_cone_adj<4>;
            hipLaunchKernelGGL(k, grd, dim3(256), 0, st, g, grid_of(g), prj, img, v, v + 3 * g.nview,
                               v + 6 * g.nview, v + 9 * g.nview, cls_.as<int>(), nC);
        } else {
            work.ensure((size_t)ncols * (g.nv + 1) * sizeof(double));
            double* T = work.as<double>();
            hipLaunchKernelGGL(k_dd_cone_rowsum, dim3((unsigned)cdiv(ncols, 4)), dim3(256), 0, st, g, prj, v,
                               v + 3 * g.nview, v + 6 * g.nview, v + 9 * g.nview, nC, ncols, T);
            check_launch("dd_cone_rowsum");
            auto k = c == 3 ? k_dd_cone_adj_tab<1> : c == 4 ? k_dd_cone_adj_tab<2> : k_dd_cone_adj_tab<4>;
            hipLaunchKernelGGL(k, grd, dim3(256), 0, st, g, grid_of(g), T, img, views_.as<ConeView>(),
                               cols_.as<ConeColumn>(), nC);
        }
        check_launch("dd_cone_adj");
    };
    const int c = tune_choose({TK_DD_CONE_ADJ, g.nx * g.ny, g.nz, g.nu * g.nv, g.nview, nB * nC}, 6, launch, st, true);
    launch(c);
}

void ConePlan::siddon_fp(int nB, int nC, const float* img, float* prj, DevBuf& work, hipStream_t st) const
{
    const float* v = vec_.as<float>();
    siddon_arb_fp(geo, false, nB, nC, img, prj, v, v + 3 * geo.nview, v + 6 * geo.nview, v + 9 * geo.nview, work, st);
}

void ConePlan::siddon_bp(int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st) const
{
    const float* v = vec_.as<float>();
    siddon_arb_bp(geo, false, nB, nC, prj, img, v, v + 3 * geo.nview, v + 6 * geo.nview, v + 9 * geo.nview, work, st);
}

void os_sqs_cone(const ArbGeo& g, int nB, int nC, float* x, const float* prj, const float* norm, const float* mask,
                 const float* detCenter, const float* detU, const float* detV, const float* src,
                 float lb, float ub, int nviewPerSubset, int nSubsets, float lambda, DevBuf& work, hipStream_t st)
{
    const size_t nimg = (size_t)nB * g.nx * g.ny * g.nz * nC;
    const int nvc = g.nv * nC;
    const size_t nsub = (size_t)nB * g.nu * nviewPerSubset * nvc;
    DevBuf fp(nsub * sizeof(float)), r(nsub * sizeof(float)), bp(nimg * sizeof(float));
    int start = 0;
    for (int k = 0; k < nSubsets && start < g.nview; k++) {
        const int cur = std::min(nviewPerSubset, g.nview - start);
        ArbGeo gs = g;
        gs.nview = cur;
        siddon_arb_fp(gs, false, nB, nC, x, fp.as<float>(), detCenter + 3 * start, detU + 3 * start,
                      detV + 3 * start, src + 3 * start, work, st);
        const long n = (long)nB * g.nu * cur * nvc;
        hipLaunchKernelGGL(k_cone_residual, dim3(cdiv(n, 256)), dim3(256), 0, st, fp.as<float>(), prj, r.as<float>(),
                           nB, g.nu, g.nview, cur, start, nvc);
        check_launch("cone_residual");
        siddon_arb_bp(gs, false, nB, nC, r.as<float>(), bp.as<float>(), detCenter + 3 * start, detU + 3 * start,
                      detV + 3 * start, src + 3 * start, work, st);
        hipLaunchKernelGGL(k_cone_update, dim3(cdiv((long)nimg, 256)), dim3(256), 0, st, x, bp.as<float>(), norm, mask,
                           lambda * nSubsets, lb, ub, lb < ub ? 1 : 0, nimg);
        check_launch("cone_update");
        start += cur;
    }
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

// one lane per (row, channel) of a parallel ray: the rows of one (u, view)
// are contiguous in both layouts
// views per thread: the column's geometry (FP64, so that the interpolation
// weights carry no index rounding of the 2^12-view magnitudes) is set up once
constexpr int kRebinViews = 8;

__global__ __launch_bounds__(256) void k_helical_rebin(RebinGeo r, const float* __restrict__ fan,
                                                       float* __restrict__ par, int nC)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;      // iv * nC + c
    const int k0 = blockIdx.y * kRebinViews;
    const int lb = blockIdx.z;                                  // b * nu_p + l
    const int nq = r.nv_p * nC;
    if (q >= nq) return;
    const int l = lb % r.nu_p, b = lb / r.nu_p;
    const int iv = q / nC, c = q - iv * nC;
    const int kend = min(k0 + kRebinViews, r.nview_p);
    float* out = par + ((long)lb * r.nview_p) * nq + q;
    const double u = ((double)l - 0.5 * (r.nu_p - 1) - (double)r.off_u_p) * (double)r.du_p;
    const double sg = u / (double)r.dso;
    if (!(fabs(sg) < 1.0)) {
        for (int k = k0; k < kend; ++k) out[(long)k * nq] = 0.0f;
        return;
    }
    const double cg = sqrt(fma(-sg, sg, 1.0));
    // column s = dsd tan(g) and row height w = v / cos(g) on the flat detector
    const double fc = (double)r.dsd * sg / (cg * (double)r.du_f) + (double)r.off_u_f + 0.5 * (r.nu_f - 1);
    const double v = ((double)iv - 0.5 * (r.nv_p - 1) - (double)r.off_v_p) * (double)r.dv_p;
    const double fr = v / (cg * (double)r.dv_f) + (double)r.off_v_f + 0.5 * (r.nv_f - 1);
    const double gm = asin(sg) * r.idbeta;
    const double c0 = floor(fc), r0 = floor(fr);
    const float wc = (float)(fc - c0), wr = (float)(fr - r0);
    const bool in_cr = c0 > -2.0 && c0 < (double)r.nu_f && r0 > -2.0 && r0 < (double)r.nv_f;
    const int ic = in_cr ? (int)c0 : 0, ir = in_cr ? (int)r0 : 0;
    const float* base = fan + (long)b * r.nu_f * r.nview_f * r.nv_f * nC + c;
    for (int k = k0; k < kend; ++k) {
        const double fm = fma((double)k, r.kscale, r.moff) - gm;
        const double m0 = floor(fm);
        const float wm = (float)(fm - m0);
        float val = 0.0f;
        // the 8 corner samples; beyond the raw data's index range a sample is 0
        if (in_cr && m0 > -2.0 && m0 < (double)r.nview_f) {
            const int im = (int)m0;
            #pragma unroll
            for (int dc = 0; dc < 2; ++dc) {
                const int cc = ic + dc;
                if (cc < 0 || cc >= r.nu_f) continue;
                #pragma unroll
                for (int dm = 0; dm < 2; ++dm) {
                    const int mm = im + dm;
                    if (mm < 0 || mm >= r.nview_f) continue;
                    const float* row = base + ((long)cc * r.nview_f + mm) * r.nv_f * nC;
                    const float w2 = (dc ? wc : 1.0f - wc) * (dm ? wm : 1.0f - wm);
                    #pragma unroll
                    for (int dr = 0; dr < 2; ++dr) {
                        const int rr = ir + dr;
                        if (rr < 0 || rr >= r.nv_f) continue;
                        val = fmaf(row[(long)rr * nC], w2 * (dr ? wr : 1.0f - wr), val);
                    }
                }
            }
        }
        out[(long)k * nq] = val;
    }
}

void helical_rebin(const RebinGeo& r, int nB, int nC, const float* fan, float* par, hipStream_t st)
{
    RN_REQUIRE(nB > 0 && nC > 0, "nBatches and nChannels must be positive");
    RN_REQUIRE(r.nu_f > 0 && r.nview_f > 0 && r.nv_f > 0 && r.nu_p > 0 && r.nview_p > 0 && r.nv_p > 0,
               "rebinning dimensions must be positive");
    RN_REQUIRE(r.dso > 0 && r.dsd > 0 && r.du_f != 0 && r.dv_f != 0 && r.idbeta != 0 && std::isfinite(r.idbeta),
               "rebinning distances and spacings must be non-zero");
    RN_REQUIRE((long)nB * r.nu_p < 0x7fffffffL && r.nview_p < 65536, "rebinning grid too large");
    const dim3 grd(cdiv((long)r.nv_p * nC, 256), cdiv(r.nview_p, kRebinViews), nB * r.nu_p);
    hipLaunchKernelGGL(k_helical_rebin, grd, dim3(256), 0, st, r, fan, par, nC);
    check_launch("helical_rebin");
}

void helical_bp(const HelicalGeo& h, int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st)
{
    RN_REQUIRE(h.nu >= 2, "the helical backprojection needs at least 2 detector columns");
    RN_REQUIRE(h.nu < (1 << 24) && h.nv < (1 << 24) && (long)h.nu * h.nv < (1L << 31),
               "helical backprojection: detector too large");
    const size_t per = (size_t)h.nu * h.nview * h.nv;
    const size_t vt_off = (2 * (size_t)nB * nC * per + 3) / 4 * 4;          // 16-byte aligned view table
    work.ensure((vt_off + 4 * (size_t)h.nview) * sizeof(float));
    float* t1 = work.as<float>();
    float* t2 = t1 + (size_t)nB * nC * per;
    float4* vt = reinterpret_cast<float4*>(t1 + vt_off);
    hipLaunchKernelGGL(k_helical_views, dim3(cdiv(h.nview, 256)), dim3(256), 0, st, h, vt);
    check_launch("helical_views");
    // [b][u][view][v][c] -> [(b, c)][u][view*v] -> [(b, c)][view][v][u]
    const float* base = prj;
    if (nC > 1) {
        // [b][u*view*v][c] -> [b][c][u*view*v]
        launch_transpose(prj, t2, (int)per, nC, nB, st);
        base = t2;
    }
    launch_transpose(base, t1, h.nu, h.nview * h.nv, nB * nC, st);
    // candidates: columns of 4, 8 or 2 voxels per lane
    auto launch = [&](int cand) {
        auto go = [&](auto kern, int zc) {
            const int nzc = cdiv(h.nz, zc);
            RN_REQUIRE((long)nB * nC * nzc <= 65535, "helical_bp: batch too large for one launch");
            const dim3 grd(cdiv(h.ny, 16), cdiv(h.nx, 16), nB * nC * nzc);
            hipLaunchKernelGGL(kern, grd, dim3(256), 0, st, h, t1, vt, img, nC, nzc);
        };
        if (cand == 0) go(k_helical_bp<4>, 4);
        else if (cand == 1) go(k_helical_bp<8>, 8);
        else go(k_helical_bp<2>, 2);
        check_launch("helical_bp");
    };
    const int c = tune_choose({TK_HELICAL_BP, h.nx * h.ny, h.nz, h.nu * h.nv, h.nview, nB * nC}, 3, launch, st, true);
    launch(c);
}

} // namespace rn
