// CPU oracle for the ReconNet projector / reconstruction path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain host C++, the
// algorithms of the reference's CUDA kernel bodies (wudufan/ReconNet,
// directory ReconNet4D_no_tf/ plus Recon3D/HelicalBP.cu).  Every routine is a
// template on the arithmetic type R:
//   R = double : the parity oracle (every intermediate in FP64),
//   R = float  : the CPU timing baseline (same loops, FP32 arithmetic).
// Inputs are always float32 (the buffers the reference's C-ABI receives);
// outputs are R.  Nothing in the product library links this file: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it.
//
// Deliberate, documented differences from the reference binary:
//  * hardware texture fetches (9-bit fixed-point weights) are emulated with
//    exact weights; __sinf/__cosf/atanf/sqrtf map to the exact functions of R;
//  * cuFFT is replaced by exact DFTs / direct linear convolution;
//  * cOSSQS uses nviewCurrent = nview - startView for a short last subset
//    (ossqs.cu:95-98 uses nview - nviewPerSubset, which overruns its buffer);
//  * cSiddonParallel* zero their outputs and BP runs over (iu, view, iv)
//    (SiddonParallel.cu:133-134,208-218 launch over the image grid instead).
// The host decisions that the reference takes in float32 (the X/Y view split,
// DistanceDriven.cu:302,705) are taken in float32 here too, with glibc
// cosf/sinf, so oracle and product partition views identically.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::vector;

// Projector.h:5 defines PI as 3.141592657f; as a float it equals float(pi).
const double kPI = (double)3.141592657f;

template <class R> inline R rsin(R x) { return std::sin(x); }
template <class R> inline R rcos(R x) { return std::cos(x); }

struct V2 { double x, y; };

// ---------------------------------------------------------------------------
// View split: DistanceDriven.cu:299-312 (and :702-715).  float32 decision.
// ---------------------------------------------------------------------------
void split_views(const float* deg, int nview, vector<int>& vx, vector<int>& vy)
{
    vx.clear(); vy.clear();
    for (int i = 0; i < nview; i++) {
        float d = deg[i];
        if (std::fabs(cosf(d)) > std::fabs(sinf(d))) vx.push_back(i); else vy.push_back(i);
    }
}

// ---------------------------------------------------------------------------
// Emulation of tex3D<float> with cudaFilterModeLinear and unnormalised
// coordinates, exact weights.  Array element (w,h,d) lives at
// arr[(d*H + h)*W + w]; mode 0 = border (0), 1 = clamp.
// (Texture setups: DistanceDriven.cu:244-251,263-265,658-665; Projector.cu:195-200.)
// ---------------------------------------------------------------------------
template <class R>
inline R tex_fetch(const R* arr, int W, int H, int D, long w, long h, long d, int mw, int mh, int md)
{
    if (mw) w = std::min<long>(std::max<long>(w, 0), W - 1); else if (w < 0 || w >= W) return R(0);
    if (mh) h = std::min<long>(std::max<long>(h, 0), H - 1); else if (h < 0 || h >= H) return R(0);
    if (md) d = std::min<long>(std::max<long>(d, 0), D - 1); else if (d < 0 || d >= D) return R(0);
    return arr[((size_t)d * H + h) * W + w];
}

template <class R>
R tex3(const R* arr, int W, int H, int D, R cw, R ch, R cd, int mw, int mh, int md)
{
    R pw = cw - R(0.5), ph = ch - R(0.5), pd = cd - R(0.5);
    R fw0 = std::floor(pw), fh0 = std::floor(ph), fd0 = std::floor(pd);
    long iw = (long)fw0, ih = (long)fh0, id = (long)fd0;
    R a = pw - fw0, b = ph - fh0, c = pd - fd0;
    R v = 0;
    for (int kd = 0; kd < 2; kd++) {
        R wd = kd ? c : R(1) - c;
        if (wd == R(0)) continue;
        for (int kh = 0; kh < 2; kh++) {
            R wh = kh ? b : R(1) - b;
            if (wh == R(0)) continue;
            for (int kw = 0; kw < 2; kw++) {
                R ww = kw ? a : R(1) - a;
                if (ww == R(0)) continue;
                v += wd * wh * ww * tex_fetch(arr, W, H, D, iw + kw, ih + kh, id + kd, mw, mh, md);
            }
        }
    }
    return v;
}

// index helpers: img[B][nx][ny][nz][C], prj[B][nu][nview][nv][C]
inline size_t img_idx(int b, int ix, int iy, int iz, int c, int nx, int ny, int nz, int nc)
{ return ((((size_t)b * nx + ix) * ny + iy) * nz + iz) * nc + c; }
inline size_t prj_idx(int b, int iu, int iview, int iv, int c, int nu, int nview, int nv, int nc)
{ return ((((size_t)b * nu + iu) * nview + iview) * nv + iv) * nc + c; }

struct Fan {
    int nB, nC, nx, ny, nz; double dx, dy, dz;
    int nu, nview, nv; double da, dv, off_a, off_v, dsd, dso; int type;
};

// ---------------------------------------------------------------------------
// DD fan forward projection: DistanceDriven.cu:23-60 (integral images),
// :64-196 (DDFPFanKernelX/Y), :198-384 (orchestration).
// ---------------------------------------------------------------------------
template <class R>
void dd_fan_fp(R* prj, const float* img, const float* deg, const Fan& g)
{
    vector<int> vx, vy; split_views(deg, g.nview, vx, vy);
    const int nx = g.nx, ny = g.ny, nz = g.nz, nc = g.nC, nu = g.nu, nv = g.nv, nview = g.nview;
    vector<R> accX((size_t)(nx + 1) * ny * nz), accY((size_t)nx * (ny + 1) * nz);
    vector<int> cls(nview, 0);
    for (int i : vy) cls[i] = 1;
    for (int b = 0; b < g.nB; b++)
    for (int c = 0; c < nc; c++) {
        // AccumulateKernelX: accX[ix+1][iy][iz] = accX[ix][iy][iz] + img[ix][iy][iz]
        for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++) {
            accX[(size_t)iy * nz + iz] = 0;
            for (int ix = 0; ix < nx; ix++)
                accX[((size_t)(ix + 1) * ny + iy) * nz + iz] = accX[((size_t)ix * ny + iy) * nz + iz]
                    + (R)img[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)];
        }
        // AccumulateKernelY: accY[ix][iy+1][iz]
        for (int ix = 0; ix < nx; ix++) for (int iz = 0; iz < nz; iz++) {
            accY[((size_t)ix * (ny + 1)) * nz + iz] = 0;
            for (int iy = 0; iy < ny; iy++)
                accY[((size_t)ix * (ny + 1) + iy + 1) * nz + iz] = accY[((size_t)ix * (ny + 1) + iy) * nz + iz]
                    + (R)img[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)];
        }
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int iview = 0; iview < nview; iview++)
        for (int iu = 0; iu < nu; iu++) {
            const R d = (R)deg[iview];
            const R cosD = rcos(d), sinD = rsin(d);
            const R da = (R)g.da, dso = (R)g.dso, dsd = (R)g.dsd;
            const R a = (-(iu - (nu - 1) / R(2)) - (R)g.off_a) * da;
            R sx = dso * sinD, sy = -dso * cosD;
            R e1x = dsd * rsin(a - da / 2), e1y = -dso + dsd * rcos(a - da / 2);
            R e2x = dsd * rsin(a + da / 2), e2y = -dso + dsd * rcos(a + da / 2);
            R d1x = e1x * cosD - e1y * sinD, d1y = e1x * sinD + e1y * cosD;
            R d2x = e2x * cosD - e2y * sinD, d2y = e2x * sinD + e2y * cosD;
            // PhysicsToImg (cx = cy = 0 for the fan projector)
            sx = sx / (R)g.dx + nx / R(2); sy = sy / (R)g.dy + ny / R(2);
            d1x = d1x / (R)g.dx + nx / R(2); d1y = d1y / (R)g.dy + ny / R(2);
            d2x = d2x / (R)g.dx + nx / R(2); d2y = d2y / (R)g.dy + ny / R(2);
            for (int iv = 0; iv < nv; iv++) {
                R z = (iv - (nv - 1) / R(2) + (R)g.off_v) * (R)g.dv;
                z = z / (R)g.dz + nz / R(2);
                R val = 0;
                if (cls[iview] == 0) {
                    R p1x = d1x, p1y = d1y, p2x = d2x, p2y = d2y;
                    if (p1x > p2x) { std::swap(p1x, p2x); std::swap(p1y, p2y); }
                    R r1 = (p1x - sx) / (p1y - sy), r2 = (p2x - sx) / (p2y - sy);
                    for (int iy = 0; iy < ny; iy++) {
                        R x1 = sx + r1 * (iy - sy), x2 = sx + r2 * (iy - sy);
                        // texAccX: W = nz (border), H = ny (border), D = nx+1 (clamp)
                        val += (tex3<R>(accX.data(), nz, ny, nx + 1, z, (R)iy, x2 + R(0.5), 0, 0, 1)
                              - tex3<R>(accX.data(), nz, ny, nx + 1, z, (R)iy, x1 + R(0.5), 0, 0, 1)) / (x2 - x1);
                    }
                    val *= (R)g.dy / std::fabs(rcos(d - a));
                } else {
                    R p1x = d1x, p1y = d1y, p2x = d2x, p2y = d2y;
                    if (p1y > p2y) { std::swap(p1x, p2x); std::swap(p1y, p2y); }
                    R r1 = (p1y - sy) / (p1x - sx), r2 = (p2y - sy) / (p2x - sx);
                    for (int ix = 0; ix < nx; ix++) {
                        R y1 = sy + r1 * (ix - sx), y2 = sy + r2 * (ix - sx);
                        // texAccY: W = nz (border), H = ny+1 (clamp), D = nx (border)
                        val += (tex3<R>(accY.data(), nz, ny + 1, nx, z, y2 + R(0.5), (R)ix, 0, 1, 0)
                              - tex3<R>(accY.data(), nz, ny + 1, nx, z, y1 + R(0.5), (R)ix, 0, 1, 0)) / (y2 - y1);
                    }
                    val *= (R)g.dx / std::fabs(rsin(d - a));
                }
                prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nc)] = val;
            }
        }
    }
}

// DistanceDriven.cu:479-494
template <class R> inline R proj_on_det(R x, R y, R dso, R cosD, R sinD)
{
    R rx = x * cosD + y * sinD, ry = -x * sinD + y * cosD;
    return std::atan(rx / (ry + dso));
}
template <class R> inline R fbp_weight(R x, R y, R dso, R cosD, R sinD)
{
    R rx = x * cosD + y * sinD, ry = -x * sinD + y * cosD;
    return dso * dso / (rx * rx + (dso + ry) * (dso + ry));
}

// ---------------------------------------------------------------------------
// DD fan backprojection: DistanceDriven.cu:441-477 (pre-weights),
// :500-614 (DDBPFanKernelX/Y), :619-791 (orchestration).
// typeProjector == 1 -> FBP weighting, no length pre-weight.
// ---------------------------------------------------------------------------
template <class R>
void dd_fan_bp(R* imgOut, const float* prj, const float* deg, const Fan& g)
{
    vector<int> vx, vy; split_views(deg, g.nview, vx, vy);
    const bool isFBP = (g.type == 1);
    const int nx = g.nx, ny = g.ny, nz = g.nz, nc = g.nC, nu = g.nu, nv = g.nv, nview = g.nview;
    vector<int> cls(nview, 0);
    for (int i : vy) cls[i] = 1;
    vector<R> w((size_t)nu * nview * nv), accU((size_t)(nu + 1) * nview * nv);
    for (int b = 0; b < g.nB; b++)
    for (int c = 0; c < nc; c++) {
        for (int iu = 0; iu < nu; iu++) for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++) {
            R v = (R)prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nc)];
            if (!isFBP) {
                R a = (-(iu - (nu - 1) / R(2)) - (R)g.off_a) * (R)g.da;
                R d = (R)deg[iview];
                if (cls[iview] == 0) v *= (R)g.dy / std::fabs(rcos(d - a));
                else v *= (R)g.dx / std::fabs(rsin(d - a));
            }
            w[((size_t)iu * nview + iview) * nv + iv] = v;
        }
        // AccumulateKernelX reused along u: accU[iu+1][view][iv]
        for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++) {
            accU[(size_t)iview * nv + iv] = 0;
            for (int iu = 0; iu < nu; iu++)
                accU[((size_t)(iu + 1) * nview + iview) * nv + iv] = accU[((size_t)iu * nview + iview) * nv + iv]
                    + w[((size_t)iu * nview + iview) * nv + iv];
        }
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const R dso = (R)g.dso;
            R x = (ix - (nx - 1) / R(2)) * (R)g.dx;
            R y = (iy - (ny - 1) / R(2)) * (R)g.dy;
            for (int iz = 0; iz < nz; iz++) {
                R ivc = (iz - (nz - 1) / R(2)) * (R)g.dz / (R)g.dv + (R)g.off_v + (nv - 1) / R(2);
                R acc = 0;
                for (int pass = 0; pass < 2; pass++) {
                    const vector<int>& lst = pass == 0 ? vx : vy;
                    for (int iview : lst) {
                        R d = (R)deg[iview];
                        R cosD = rcos(d), sinD = rsin(d);
                        R a1, a2;
                        if (pass == 0) {
                            a1 = proj_on_det(x - (R)g.dx / 2, y, dso, cosD, sinD);
                            a2 = proj_on_det(x + (R)g.dx / 2, y, dso, cosD, sinD);
                        } else {
                            a1 = proj_on_det(x, y - (R)g.dy / 2, dso, cosD, sinD);
                            a2 = proj_on_det(x, y + (R)g.dy / 2, dso, cosD, sinD);
                        }
                        a1 = -(a1 / (R)g.da + (R)g.off_a) + nu / R(2);
                        a2 = -(a2 / (R)g.da + (R)g.off_a) + nu / R(2);
                        if (a1 > a2) std::swap(a1, a2);
                        // texAccU: W = nv (border), H = nview (border), D = nu+1 (clamp)
                        R val = (tex3<R>(accU.data(), nv, nview, nu + 1, ivc + R(0.5), iview + R(0.5), a2 + R(0.5), 0, 0, 1)
                               - tex3<R>(accU.data(), nv, nview, nu + 1, ivc + R(0.5), iview + R(0.5), a1 + R(0.5), 0, 0, 1)) / (a2 - a1);
                        if (isFBP) val *= fbp_weight(x, y, dso, cosD, sinD);
                        acc += val;
                    }
                }
                imgOut[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Siddon ray tracing: Siddon.cu:11-153 (InitializeSiddon and helpers),
// :155-238 (gather), :270-351 (transposed scatter).
// ---------------------------------------------------------------------------
struct Grid3 { int nx, ny, nz; double dx, dy, dz, cx, cy, cz; };
struct P3 { double x, y, z; };

template <class R> struct SidVars {
    R ax, ay, az, dax, day, daz; int gx, gy, gz; int px, py, pz; R now, prev, len; bool valid;
};

template <class R>
inline bool inbound_alpha(R& a0, R& aN, R d, R s, int n)
{
    const R eps = R(1e-6);
    if (std::fabs(d) > eps) { a0 = (0 - s) / d; aN = (n - s) / d; }
    else if (s < 0 || s > n) return false;
    return true;
}
template <class R> inline int inbound_first(R p, R d, int n)
{
    const R eps = R(1e-6);
    int g = (int)p;
    if (std::fabs(d) > eps) { if (g < 0) g = 0; else if (g >= n) g = n - 1; }
    return g;
}
template <class R> inline R outbound_first(int g, R d)
{
    const R eps = R(1e-6);
    if (d > eps) return (R)(g + 1);
    return (R)g;
}

template <class R>
SidVars<R> init_siddon(P3 s3, P3 t3, const Grid3& gr)
{
    SidVars<R> v; v.valid = false;
    R sx = (R)s3.x, sy = (R)s3.y, sz = (R)s3.z, tx = (R)t3.x, ty = (R)t3.y, tz = (R)t3.z;
    R dist = std::sqrt((sx - tx) * (sx - tx) + (sy - ty) * (sy - ty) + (sz - tz) * (sz - tz));
    sx = (sx - (R)gr.cx) / (R)gr.dx + gr.nx / R(2); sy = (sy - (R)gr.cy) / (R)gr.dy + gr.ny / R(2); sz = (sz - (R)gr.cz) / (R)gr.dz + gr.nz / R(2);
    tx = (tx - (R)gr.cx) / (R)gr.dx + gr.nx / R(2); ty = (ty - (R)gr.cy) / (R)gr.dy + gr.ny / R(2); tz = (tz - (R)gr.cz) / (R)gr.dz + gr.nz / R(2);
    R ddx = tx - sx, ddy = ty - sy, ddz = tz - sz;
    R a0x = 1, a0y = 1, a0z = 1, a1x = 0, a1y = 0, a1z = 0;
    if (!inbound_alpha(a0x, a1x, ddx, sx, gr.nx) || !inbound_alpha(a0y, a1y, ddy, sy, gr.ny) ||
        !inbound_alpha(a0z, a1z, ddz, sz, gr.nz)) return v;
    R amin = std::max(R(0), std::max(std::min(a0x, a1x), std::max(std::min(a0y, a1y), std::min(a0z, a1z))));
    R amax = std::min(R(1), std::min(std::max(a0x, a1x), std::min(std::max(a0y, a1y), std::max(a0z, a1z))));
    if (amax <= amin) return v;
    R p0x = sx + amin * ddx, p0y = sy + amin * ddy, p0z = sz + amin * ddz;
    v.gx = inbound_first(p0x, ddx, gr.nx); v.gy = inbound_first(p0y, ddy, gr.ny); v.gz = inbound_first(p0z, ddz, gr.nz);
    R p1x = outbound_first(v.gx, ddx), p1y = outbound_first(v.gy, ddy), p1z = outbound_first(v.gz, ddz);
    const R eps = R(1e-6);
    R ax = 2, ay = 2, az = 2, dax = 0, day = 0, daz = 0;
    if (std::fabs(ddx) > eps) { ax = (p1x - sx) / ddx; dax = 1 / std::fabs(ddx); }
    if (std::fabs(ddy) > eps) { ay = (p1y - sy) / ddy; day = 1 / std::fabs(ddy); }
    if (std::fabs(ddz) > eps) { az = (p1z - sz) / ddz; daz = 1 / std::fabs(ddz); }
    v.ax = ax * dist; v.ay = ay * dist; v.az = az * dist;
    v.dax = dax * dist; v.day = day * dist; v.daz = daz * dist;
    v.px = ddx > 0; v.py = ddy > 0; v.pz = ddz > 0;
    v.now = v.prev = amin * dist;
    v.len = dist * (amax - amin);
    v.valid = !(v.gx < 0 || v.gx >= gr.nx || v.gy < 0 || v.gy >= gr.ny || v.gz < 0 || v.gz >= gr.nz);
    return v;
}

// Walks the ray; calls f(linear voxel index (x,y,z order, no channel), length).
template <class R, class F>
void siddon_walk(P3 s3, P3 t3, const Grid3& gr, F f)
{
    SidVars<R> v = init_siddon<R>(s3, t3, gr);
    if (!v.valid) return;
    bool tracing = true;
    int gx = v.gx, gy = v.gy, gz = v.gz;
    while (tracing) {
        int cx = gx, cy = gy, cz = gz;
        if (v.ax < v.ay && v.ax < v.az) {
            v.now = v.ax; v.ax += v.dax;
            if (v.px) { gx++; if (gx >= gr.nx) tracing = false; } else { gx--; if (gx < 0) tracing = false; }
        } else if (v.ay < v.az) {
            v.now = v.ay; v.ay += v.day;
            if (v.py) { gy++; if (gy >= gr.ny) tracing = false; } else { gy--; if (gy < 0) tracing = false; }
        } else {
            v.now = v.az; v.az += v.daz;
            if (v.pz) { gz++; if (gz >= gr.nz) tracing = false; } else { gz--; if (gz < 0) tracing = false; }
        }
        f(((size_t)cx * gr.ny + cy) * gr.nz + cz, v.now - v.prev);
        v.prev = v.now;
    }
}

// Fan ray end points: Siddon.cu:253-263
template <class R>
inline void fan_ray(const Fan& g, float degf, int iu, int iv, P3& s, P3& t)
{
    R d = (R)degf;
    R cosD = rcos(d), sinD = rsin(d);
    R a = (-(iu - (g.nu - 1) / R(2)) - (R)g.off_a) * (R)g.da;
    R z = (iv - (g.nv - 1) / R(2) + (R)g.off_v) * (R)g.dv;
    R dso = (R)g.dso, dsd = (R)g.dsd;
    R rx = dsd * rsin(a), ry = -dso + dsd * rcos(a);
    s = { (double)(dso * sinD), (double)(-dso * cosD), (double)z };
    t = { (double)(rx * cosD - ry * sinD), (double)(rx * sinD + ry * cosD), (double)z };
}

template <class R>
void siddon_fan_fp(R* prj, const float* img, const float* deg, const Fan& g)
{
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 0, 0, 0 };
    const int nc = g.nC;
    for (int b = 0; b < g.nB; b++)
    for (int c = 0; c < nc; c++) {
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int iview = 0; iview < g.nview; iview++)
        for (int iu = 0; iu < g.nu; iu++)
        for (int iv = 0; iv < g.nv; iv++) {
            P3 s, t; fan_ray<R>(g, deg[iview], iu, iv, s, t);
            R val = 0;
            const float* base = img + (size_t)b * g.nx * g.ny * g.nz * nc + c;
            siddon_walk<R>(s, t, gr, [&](size_t vox, R len) { val += (R)base[vox * nc] * len; });
            prj[prj_idx(b, iu, iview, iv, c, g.nu, g.nview, g.nv, nc)] = val;
        }
    }
}

template <class R>
void siddon_fan_bp(R* imgOut, const float* prj, const float* deg, const Fan& g)
{
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 0, 0, 0 };
    const int nc = g.nC;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    for (int b = 0; b < g.nB; b++)
    for (int c = 0; c < nc; c++) {
        vector<R> total(nvox, R(0));
        #pragma omp parallel
        {
            vector<R> part(nvox, R(0));
            #pragma omp for collapse(2) schedule(dynamic, 4)
            for (int iview = 0; iview < g.nview; iview++)
            for (int iu = 0; iu < g.nu; iu++)
            for (int iv = 0; iv < g.nv; iv++) {
                P3 s, t; fan_ray<R>(g, deg[iview], iu, iv, s, t);
                R val = (R)prj[prj_idx(b, iu, iview, iv, c, g.nu, g.nview, g.nv, nc)];
                siddon_walk<R>(s, t, gr, [&](size_t vox, R len) { part[vox] += val * len; });
            }
            #pragma omp critical
            for (size_t i = 0; i < nvox; i++) total[i] += part[i];
        }
        for (size_t i = 0; i < nvox; i++) imgOut[((size_t)b * nvox + i) * nc + c] = total[i];
    }
}

// ---------------------------------------------------------------------------
// Pixel-driven fan BP: PixelDriven.cu:11-97 (always FBP-weighted, :93).
// ---------------------------------------------------------------------------
template <class R>
void pixel_fan_bp(R* imgOut, const float* prj, const float* deg, const Fan& g)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, nc = g.nC, nu = g.nu, nv = g.nv, nview = g.nview;
    vector<R> tex((size_t)nu * nview * nv);
    for (int b = 0; b < g.nB; b++)
    for (int c = 0; c < nc; c++) {
        for (int iu = 0; iu < nu; iu++) for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++)
            tex[((size_t)iu * nview + iview) * nv + iv] = (R)prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nc)];
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const R dso = (R)g.dso;
            R x = (ix - (nx - 1) / R(2)) * (R)g.dx, y = (iy - (ny - 1) / R(2)) * (R)g.dy;
            vector<R> val(nz, R(0));
            for (int i = 0; i < nview; i++) {
                R d = (R)deg[i];
                R cosD = rcos(d), sinD = rsin(d);
                R rx = x * cosD + y * sinD, ry = -x * sinD + y * cosD;
                R a = std::atan(rx / (ry + dso));
                R dist = dso * dso / (rx * rx + (dso + ry) * (dso + ry));
                R pu = -(a / (R)g.da + (R)g.off_a) + (nu - R(1)) / R(2);
                for (int iz = 0; iz < nz; iz++) {
                    R pv = (iz - (nz - 1) / R(2)) * (R)g.dz / (R)g.dv + (R)g.off_v + (nv - R(1)) / R(2);
                    // texPrj: W = nv, H = nview, D = nu, all border
                    val[iz] += tex3<R>(tex.data(), nv, nview, nu, pv + R(0.5), i + R(0.5), pu + R(0.5), 0, 0, 0) * dist;
                }
            }
            for (int iz = 0; iz < nz; iz++) imgOut[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)] = val[iz];
        }
    }
}

// ---------------------------------------------------------------------------
// Ramp filter: Filter.cu:66-214 (getRamp), :216-342 (fan), :344-430 (parallel).
// The FFT-domain product of cuFFT is an exact linear convolution of each
// detector row with the (2nu-1)-tap windowed kernel (no wrap-around because
// the row is zero-padded to 2nu-1).  Here: exact DFTs for the window step and
// a direct convolution for the filtering step.
// ---------------------------------------------------------------------------
template <class R>
void ramp_kernel(vector<R>& kr, int nu, R da, int filterType, bool equalSpace)
{
    const int L = 2 * nu - 1;
    vector<double> h(L);
    for (int i = 0; i < L; i++) {
        int k = i - (nu - 1);
        if (k == 0) h[i] = 1.0 / (4.0 * (double)da * (double)da);
        else if (k % 2 != 0) {
            if (equalSpace) h[i] = -1.0 / (kPI * kPI * (double)k * k * (double)da * (double)da);
            else { double s = std::sin(k * (double)da); h[i] = -1.0 / (kPI * kPI * s * s); }
        } else h[i] = 0;
    }
    vector<double> win(L, 1.0);
    for (int i = 0; i < L; i++) {
        double t = (double)i / (double)(L - 1);
        switch (filterType) {
        case 1: win[i] = 0.54 + 0.46 * std::cos(2 * kPI * t); break;
        case 2: win[i] = 0.5 + 0.5 * std::cos(2 * kPI * t); break;
        case 3: win[i] = std::fabs(std::cos(kPI * t)); break;
        default: win[i] = 1.0;
        }
    }
    // forward DFT (cuFFT sign convention: exp(-2 pi i f n / L)), window, inverse, Re / L
    vector<double> cs(L), sn(L);
    for (int i = 0; i < L; i++) { cs[i] = std::cos(2 * M_PI * i / L); sn[i] = std::sin(2 * M_PI * i / L); }
    vector<double> Hr(L), Hi(L);
    for (int f = 0; f < L; f++) {
        double re = 0, im = 0;
        for (int n = 0; n < L; n++) {
            int m = (int)(((long)f * n) % L);
            re += h[n] * cs[m]; im -= h[n] * sn[m];
        }
        Hr[f] = re * win[f]; Hi[f] = im * win[f];
    }
    kr.assign(L, R(0));
    for (int n = 0; n < L; n++) {
        double re = 0;
        for (int f = 0; f < L; f++) {
            int m = (int)(((long)f * n) % L);
            re += Hr[f] * cs[m] - Hi[f] * sn[m];
        }
        kr[n] = (R)(re / L);
    }
}

template <class R>
void ramp_filter(R* out, const float* prj, int nB, int nC, int nu, int nview, int nv,
                 float da, float off_a, float dsd, float dso, int filterType, bool fan)
{
    vector<R> kr; ramp_kernel<R>(kr, nu, (R)da, filterType, !fan);
    vector<R> pw(nu, R(1));
    if (fan) for (int i = 0; i < nu; i++) pw[i] = std::cos(((i - (nu - 1) / R(2)) - (R)off_a) * (R)da);
    (void)dsd;
    // post scale after C2R (which multiplies by L) and the /L folded into the
    // reference's scale: Filter.cu:301 (fan) and :393 (parallel)
    R scale = fan ? (R)kPI / nview * (R)da / (R)dso : (R)kPI / nview * (R)da;
    #pragma omp parallel for collapse(2)
    for (int b = 0; b < nB; b++)
    for (int iview = 0; iview < nview; iview++)
    for (int iv = 0; iv < nv; iv++)
    for (int c = 0; c < nC; c++) {
        vector<R> x(nu);
        for (int j = 0; j < nu; j++) x[j] = (R)prj[prj_idx(b, j, iview, iv, c, nu, nview, nv, nC)] * pw[j];
        for (int iu = 0; iu < nu; iu++) {
            R s = 0;
            for (int j = 0; j < nu; j++) s += x[j] * kr[iu - j + nu - 1];
            out[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nC)] = s * scale;
        }
    }
}

// ---------------------------------------------------------------------------
// TV surrogate: TotalVariation.cu:9-68 (kernel) and :124-134 (host sum).
// ---------------------------------------------------------------------------
template <class R>
R tv_sqs(R* grad, R* norm, const float* img, int nx, int ny)
{
    R tv = 0;
    for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
        size_t ind = (size_t)ix * ny + iy;
        R x = img[ind];
        R dx0 = ix == 0 ? x : x - (R)img[ind - ny];
        R dx1 = ix == nx - 1 ? x : x - (R)img[ind + ny];
        R dy0 = iy == 0 ? x : x - (R)img[ind - 1];
        R dy1 = iy == ny - 1 ? x : x - (R)img[ind + 1];
        const R eps = R(1e-6f);
        R t0 = std::sqrt(dx0 * dx0 + eps), t1 = std::sqrt(dx1 * dx1 + eps);
        R t2 = std::sqrt(dy0 * dy0 + eps), t3 = std::sqrt(dy1 * dy1 + eps);
        grad[ind] = (dx0 / t0 + dx1 / t1 + dy0 / t2 + dy1 / t3) / 2;
        norm[ind] = 1 / t0 + 1 / t1 + 1 / t2 + 1 / t3;
        tv += (t0 + t1 + t2 + t3) / 2;
    }
    return tv;
}

// ---------------------------------------------------------------------------
// OS-SQS data pass: ossqs.cu:12-161 with the SiddonFan projector.
// ---------------------------------------------------------------------------
template <class R>
void os_sqs(R* res, const float* img, const float* prj, const float* normImg, const float* wprj,
            const float* deg, float lb, float ub, int nviewPerSubset, int nSubsets, const Fan& g, float lambda)
{
    const size_t nimg = (size_t)g.nB * g.nx * g.ny * g.nz * g.nC;
    vector<R> x(nimg);
    for (size_t i = 0; i < nimg; i++) x[i] = img[i];
    vector<float> xf(nimg);
    int startView = 0;
    for (int s = 0; s < nSubsets && startView < g.nview; s++) {
        int cur = std::min(nviewPerSubset, g.nview - startView);
        Fan gs = g; gs.nview = cur;
        // subset sinogram slices (reference keeps them strided inside the full sinogram)
        size_t nsub = (size_t)g.nB * g.nu * cur * g.nv * g.nC;
        vector<float> degs(deg + startView, deg + startView + cur);
        // forward projection of the current x (in R); the gather is linear so
        // it is evaluated on R-valued images by splitting into an exact float
        // image plus nothing: for R = double we project in double directly.
        vector<R> fp(nsub);
        {
            Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 0, 0, 0 };
            for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
                #pragma omp parallel for collapse(2) schedule(dynamic, 4)
                for (int iview = 0; iview < cur; iview++)
                for (int iu = 0; iu < g.nu; iu++)
                for (int iv = 0; iv < g.nv; iv++) {
                    P3 sp, tp; fan_ray<R>(gs, degs[iview], iu, iv, sp, tp);
                    R val = 0;
                    const R* base = x.data() + (size_t)b * g.nx * g.ny * g.nz * g.nC + c;
                    siddon_walk<R>(sp, tp, gr, [&](size_t vox, R len) { val += base[vox * g.nC] * len; });
                    fp[prj_idx(b, iu, iview, iv, c, g.nu, cur, g.nv, g.nC)] = val;
                }
            }
        }
        // r = w (Ax - b)
        vector<float> r(nsub);
        vector<R> rr(nsub);
        for (int b = 0; b < g.nB; b++) for (int iu = 0; iu < g.nu; iu++) for (int iview = 0; iview < cur; iview++)
        for (int iv = 0; iv < g.nv; iv++) for (int c = 0; c < g.nC; c++) {
            size_t i = prj_idx(b, iu, iview, iv, c, g.nu, cur, g.nv, g.nC);
            size_t j = prj_idx(b, iu, startView + iview, iv, c, g.nu, g.nview, g.nv, g.nC);
            rr[i] = (fp[i] - (R)prj[j]) * (R)wprj[j];
        }
        // A' r (scatter, R accumulation)
        vector<R> bp(nimg, R(0));
        {
            Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 0, 0, 0 };
            const size_t nvox = (size_t)g.nx * g.ny * g.nz;
            for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
                vector<R> total(nvox, R(0));
                #pragma omp parallel
                {
                    vector<R> part(nvox, R(0));
                    #pragma omp for collapse(2) schedule(dynamic, 4)
                    for (int iview = 0; iview < cur; iview++)
                    for (int iu = 0; iu < g.nu; iu++)
                    for (int iv = 0; iv < g.nv; iv++) {
                        P3 sp, tp; fan_ray<R>(gs, degs[iview], iu, iv, sp, tp);
                        R val = rr[prj_idx(b, iu, iview, iv, c, g.nu, cur, g.nv, g.nC)];
                        siddon_walk<R>(sp, tp, gr, [&](size_t vox, R len) { part[vox] += val * len; });
                    }
                    #pragma omp critical
                    for (size_t i = 0; i < nvox; i++) total[i] += part[i];
                }
                for (size_t i = 0; i < nvox; i++) bp[((size_t)b * nvox + i) * g.nC + c] = total[i];
            }
        }
        for (size_t i = 0; i < nimg; i++) {
            R u = bp[i] / ((R)normImg[i] + R(1e-6f));
            x[i] = x[i] - (R)lambda * nSubsets * u;
            if (lb < ub) { if (x[i] < (R)lb) x[i] = lb; else if (x[i] > (R)ub) x[i] = ub; }
        }
        startView += cur;
    }
    for (size_t i = 0; i < nimg; i++) res[i] = x[i];
}

// ---------------------------------------------------------------------------
// Siddon parallel / cone with arbitrary geometry:
// SiddonParallel.cu:11-105, SiddonCone.cu:11-68, cudaMath.h:25-60.
// ---------------------------------------------------------------------------
template <class R> struct Vec3 { R x, y, z; };
template <class R> inline R vlen(Vec3<R> v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
template <class R> inline Vec3<R> vnorm(Vec3<R> v)
{ R l = vlen(v); if (l < R(1e-6f)) return v; return { v.x / l, v.y / l, v.z / l }; }

template <class R>
void move_near_grid(Vec3<R>& s, Vec3<R>& t, const Grid3& gr)
{
    R r = vlen<R>({ (R)(gr.nx * gr.dx), (R)(gr.ny * gr.dy), (R)(gr.nz * gr.dz) });
    Vec3<R> d = vnorm<R>({ t.x - s.x, t.y - s.y, t.z - s.z });
    R t0 = ((R)gr.cx - s.x) * d.x + ((R)gr.cy - s.y) * d.y + ((R)gr.cz - s.z) * d.z;
    R t1 = (t.x - (R)gr.cx) * d.x + (t.y - (R)gr.cy) * d.y + (t.z - (R)gr.cz) * d.z;
    if (t0 > r) s = { s.x + (t0 - r) * d.x, s.y + (t0 - r) * d.y, s.z + (t0 - r) * d.z };
    if (t1 > r) t = { t.x - (t1 - r) * d.x, t.y - (t1 - r) * d.y, t.z - (t1 - r) * d.z };
}

struct Arb {
    int nB, nC, nx, ny, nz; double dx, dy, dz, cx, cy, cz;
    int nu, nview, nv; double du, dv, off_u, off_v;
};

template <class R>
inline void arb_ray(const Arb& g, const float* detCenter, const float* detU, const float* detV,
                    const float* srcOrDir, bool parallel, int iview, int iu, int iv, P3& s, P3& t)
{
    const float* C = detCenter + 3 * iview; const float* U = detU + 3 * iview; const float* V = detV + 3 * iview;
    const float* S = srcOrDir + 3 * iview;
    R u = (iu - (R)g.off_u - (g.nu - 1) / R(2)) * (R)g.du;
    R v = (iv - (R)g.off_v - (g.nv - 1) / R(2)) * (R)g.dv;
    Vec3<R> dst{ (R)C[0] + (R)U[0] * u + (R)V[0] * v, (R)C[1] + (R)U[1] * u + (R)V[1] * v, (R)C[2] + (R)U[2] * u + (R)V[2] * v };
    Vec3<R> src;
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.cx, g.cy, g.cz };
    if (parallel) {
        R m1 = std::sqrt(((R)g.cx - (R)C[0]) * ((R)g.cx - (R)C[0]) + ((R)g.cy - (R)C[1]) * ((R)g.cy - (R)C[1]) + ((R)g.cz - (R)C[2]) * ((R)g.cz - (R)C[2]));
        R m2 = std::sqrt((R)(g.nx * g.dx) * (R)(g.nx * g.dx) + (R)(g.ny * g.dy) * (R)(g.ny * g.dy) + (R)(g.nz * g.dz) * (R)(g.nz * g.dz)) / 2;
        R m = std::max(m1, m2) * R(1.1f);
        src = { (R)C[0] + (R)S[0] * m + (R)U[0] * u + (R)V[0] * v,
                (R)C[1] + (R)S[1] * m + (R)U[1] * u + (R)V[1] * v,
                (R)C[2] + (R)S[2] * m + (R)U[2] * u + (R)V[2] * v };
    } else {
        src = { (R)S[0], (R)S[1], (R)S[2] };
    }
    move_near_grid<R>(src, dst, gr);
    s = { (double)src.x, (double)src.y, (double)src.z };
    t = { (double)dst.x, (double)dst.y, (double)dst.z };
}

template <class R>
void siddon_arb_fp(R* prj, const float* img, const float* detCenter, const float* detU, const float* detV,
                   const float* srcOrDir, bool parallel, const Arb& g)
{
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.cx, g.cy, g.cz };
    for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int iview = 0; iview < g.nview; iview++)
        for (int iu = 0; iu < g.nu; iu++)
        for (int iv = 0; iv < g.nv; iv++) {
            P3 s, t; arb_ray<R>(g, detCenter, detU, detV, srcOrDir, parallel, iview, iu, iv, s, t);
            R val = 0;
            const float* base = img + (size_t)b * g.nx * g.ny * g.nz * g.nC + c;
            siddon_walk<R>(s, t, gr, [&](size_t vox, R len) { val += (R)base[vox * g.nC] * len; });
            prj[prj_idx(b, iu, iview, iv, c, g.nu, g.nview, g.nv, g.nC)] = val;
        }
    }
}

template <class R>
void siddon_arb_bp(R* imgOut, const float* prj, const float* detCenter, const float* detU, const float* detV,
                   const float* srcOrDir, bool parallel, const Arb& g)
{
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.cx, g.cy, g.cz };
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
        vector<R> total(nvox, R(0));
        #pragma omp parallel
        {
            vector<R> part(nvox, R(0));
            #pragma omp for collapse(2) schedule(dynamic, 4)
            for (int iview = 0; iview < g.nview; iview++)
            for (int iu = 0; iu < g.nu; iu++)
            for (int iv = 0; iv < g.nv; iv++) {
                P3 s, t; arb_ray<R>(g, detCenter, detU, detV, srcOrDir, parallel, iview, iu, iv, s, t);
                R val = (R)prj[prj_idx(b, iu, iview, iv, c, g.nu, g.nview, g.nv, g.nC)];
                siddon_walk<R>(s, t, gr, [&](size_t vox, R len) { part[vox] += val * len; });
            }
            #pragma omp critical
            for (size_t i = 0; i < nvox; i++) total[i] += part[i];
        }
        for (size_t i = 0; i < nvox; i++) imgOut[((size_t)b * nvox + i) * g.nC + c] = total[i];
    }
}

// OS-SQS, arbitrary cone: ossqs.cu:163-333 (no weights, mask after the update)
template <class R>
void os_sqs_cone(R* res, const float* img, const float* prj, const float* normImg, const float* mask,
                 const float* detCenter, const float* detU, const float* detV, const float* src,
                 float lb, float ub, int nviewPerSubset, int nSubsets, const Arb& g, float lambda)
{
    const size_t nimg = (size_t)g.nB * g.nx * g.ny * g.nz * g.nC;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    Grid3 gr{ g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.cx, g.cy, g.cz };
    vector<R> x(img, img + nimg);
    int startView = 0;
    for (int s = 0; s < nSubsets && startView < g.nview; s++) {
        int cur = std::min(nviewPerSubset, g.nview - startView);
        Arb gs = g; gs.nview = cur;
        const float* C = detCenter + 3 * startView; const float* U = detU + 3 * startView;
        const float* V = detV + 3 * startView; const float* S = src + 3 * startView;
        size_t nsub = (size_t)g.nB * g.nu * cur * g.nv * g.nC;
        vector<R> rr(nsub);
        for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
            #pragma omp parallel for collapse(2) schedule(dynamic, 4)
            for (int iview = 0; iview < cur; iview++)
            for (int iu = 0; iu < g.nu; iu++)
            for (int iv = 0; iv < g.nv; iv++) {
                P3 sp, tp; arb_ray<R>(gs, C, U, V, S, false, iview, iu, iv, sp, tp);
                R val = 0;
                const R* base = x.data() + (size_t)b * nvox * g.nC + c;
                siddon_walk<R>(sp, tp, gr, [&](size_t vox, R len) { val += base[vox * g.nC] * len; });
                size_t i = prj_idx(b, iu, iview, iv, c, g.nu, cur, g.nv, g.nC);
                size_t j = prj_idx(b, iu, startView + iview, iv, c, g.nu, g.nview, g.nv, g.nC);
                rr[i] = val - (R)prj[j];
            }
        }
        vector<R> bp(nimg, R(0));
        for (int b = 0; b < g.nB; b++) for (int c = 0; c < g.nC; c++) {
            vector<R> total(nvox, R(0));
            #pragma omp parallel
            {
                vector<R> part(nvox, R(0));
                #pragma omp for collapse(2) schedule(dynamic, 4)
                for (int iview = 0; iview < cur; iview++)
                for (int iu = 0; iu < g.nu; iu++)
                for (int iv = 0; iv < g.nv; iv++) {
                    P3 sp, tp; arb_ray<R>(gs, C, U, V, S, false, iview, iu, iv, sp, tp);
                    R val = rr[prj_idx(b, iu, iview, iv, c, g.nu, cur, g.nv, g.nC)];
                    siddon_walk<R>(sp, tp, gr, [&](size_t vox, R len) { part[vox] += val * len; });
                }
                #pragma omp critical
                for (size_t i = 0; i < nvox; i++) total[i] += part[i];
            }
            for (size_t i = 0; i < nvox; i++) bp[((size_t)b * nvox + i) * g.nC + c] = total[i];
        }
        for (size_t i = 0; i < nimg; i++) {
            R u = bp[i] / ((R)normImg[i] + R(1e-6f));
            x[i] = (x[i] - (R)lambda * nSubsets * u) * (R)mask[i];
            if (lb < ub) { if (x[i] < (R)lb) x[i] = lb; else if (x[i] > (R)ub) x[i] = ub; }
        }
        startView += cur;
    }
    for (size_t i = 0; i < nimg; i++) res[i] = x[i];
}

// ---------------------------------------------------------------------------
// DD tomosynthesis: DistanceDriven3D.cu:37-227 (FP), :409-559 (BP).
// Integral images in R (the reference already uses FP64 there).
// ---------------------------------------------------------------------------
template <class R> inline R clamp_f(R x, R s, R e) { return x < s ? s : (x >= e ? e - 1 : x); }
inline int clamp_i(int x, int s, int e) { return x < s ? s : (x >= e ? e - 1 : x); }

// DistanceDriven3D.cu:110-135; the float casts of x,y and of the weights are kept
template <class R>
R interp_xy(const R* acc, R x, R y, int iz, int nx, int ny, int nz)
{
    x = clamp_f<R>(x, 0, (R)nx); y = clamp_f<R>(y, 0, (R)ny);
    int ix = (int)x, iy = (int)y;
    int ix1 = clamp_i(ix + 1, 0, nx), iy1 = clamp_i(iy + 1, 0, ny);
    R wx = x - ix, wy = y - iy;
    return acc[((size_t)ix * ny + iy) * nz + iz] * (1 - wx) * (1 - wy) + acc[((size_t)ix1 * ny + iy) * nz + iz] * wx * (1 - wy)
         + acc[((size_t)ix * ny + iy1) * nz + iz] * (1 - wx) * wy + acc[((size_t)ix1 * ny + iy1) * nz + iz] * wx * wy;
}

template <class R>
void dd_tomo_fp(R* prj, const float* img, const float* detCenter, const float* src, const Arb& g)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, nc = g.nC, nu = g.nu, nv = g.nv, nview = g.nview;
    vector<R> acc((size_t)(nx + 1) * (ny + 1) * nz);
    for (int b = 0; b < g.nB; b++) for (int c = 0; c < nc; c++) {
        std::fill(acc.begin(), acc.end(), R(0));
        for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++)
            acc[((size_t)(ix + 1) * (ny + 1) + iy + 1) * nz + iz] = acc[((size_t)ix * (ny + 1) + iy + 1) * nz + iz]
                + (R)img[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)];
        for (int ix = 0; ix < nx; ix++) for (int iz = 0; iz < nz; iz++) for (int iy = 0; iy < ny; iy++)
            acc[((size_t)(ix + 1) * (ny + 1) + iy + 1) * nz + iz] += acc[((size_t)(ix + 1) * (ny + 1) + iy) * nz + iz];
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int iview = 0; iview < nview; iview++)
        for (int iu = 0; iu < nu; iu++)
        for (int iv = 0; iv < nv; iv++) {
            const float* Cc = detCenter + 3 * iview; const float* S = src + 3 * iview;
            R u = (iu - (R)g.off_u - (nu - 1) / R(2)) * (R)g.du;
            R v = (iv - (R)g.off_v - (nv - 1) / R(2)) * (R)g.dv;
            // detU = (1,0,0), detV = (0,1,0) (DistanceDriven3D.cu:283-299)
            auto dstp = [&](R uu, R vv) { return Vec3<R>{ (R)Cc[0] + uu, (R)Cc[1] + vv, (R)Cc[2] }; };
            auto toimg = [&](Vec3<R> p) {
                return Vec3<R>{ (p.x - (R)g.cx) / (R)g.dx + nx / R(2), (p.y - (R)g.cy) / (R)g.dy + ny / R(2), (p.z - (R)g.cz) / (R)g.dz + nz / R(2) };
            };
            Vec3<R> dx1 = toimg(dstp(u - (R)g.du / 2, v)), dx2 = toimg(dstp(u + (R)g.du / 2, v));
            Vec3<R> dy1 = toimg(dstp(u, v - (R)g.dv / 2)), dy2 = toimg(dstp(u, v + (R)g.dv / 2));
            Vec3<R> s = toimg(Vec3<R>{ (R)S[0], (R)S[1], (R)S[2] });
            if (dx1.x > dx2.x) std::swap(dx1, dx2);
            if (dy1.y > dy2.y) std::swap(dy1, dy2);
            R rx1 = (dx1.x - s.x) / (dx1.z - s.z), rx2 = (dx2.x - s.x) / (dx2.z - s.z);
            R ry1 = (dy1.y - s.y) / (dy1.z - s.z), ry2 = (dy2.y - s.y) / (dy2.z - s.z);
            R val = 0;
            for (int iz = 0; iz < nz; iz++) {
                R x1 = s.x + rx1 * (iz - s.z), x2 = s.x + rx2 * (iz - s.z);
                R y1 = s.y + ry1 * (iz - s.z), y2 = s.y + ry2 * (iz - s.z);
                val += (interp_xy(acc.data(), x2, y2, iz, nx + 1, ny + 1, nz) + interp_xy(acc.data(), x1, y1, iz, nx + 1, ny + 1, nz)
                      - interp_xy(acc.data(), x2, y1, iz, nx + 1, ny + 1, nz) - interp_xy(acc.data(), x1, y2, iz, nx + 1, ny + 1, nz))
                      / ((x2 - x1) * (y2 - y1));
            }
            Vec3<R> d = dstp(u, v);
            R ddx = (R)S[0] - d.x, ddy = (R)S[1] - d.y, ddz = (R)S[2] - d.z;
            val *= (R)g.dz / std::fabs(ddz) * std::sqrt(ddz * ddz + ddy * ddy + ddx * ddx);
            prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nc)] = val;
        }
    }
}

template <class R>
R interp_uv(const R* acc, R u, R v, int iview, int nu, int nv, int nview)
{
    u = clamp_f<R>(u, 0, (R)nu); v = clamp_f<R>(v, 0, (R)nv);
    int iu = (int)u, iv = (int)v;
    int iu1 = clamp_i(iu + 1, 0, nu), iv1 = clamp_i(iv + 1, 0, nv);
    R wu = u - iu, wv = v - iv;
    return acc[((size_t)iu * nview + iview) * nv + iv] * (1 - wu) * (1 - wv) + acc[((size_t)iu1 * nview + iview) * nv + iv] * wu * (1 - wv)
         + acc[((size_t)iu * nview + iview) * nv + iv1] * (1 - wu) * wv + acc[((size_t)iu1 * nview + iview) * nv + iv1] * wu * wv;
}

template <class R>
void dd_tomo_bp(R* imgOut, const float* prj, const float* detCenter, const float* src, const Arb& g)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, nc = g.nC, nu = g.nu, nv = g.nv, nview = g.nview;
    vector<R> w((size_t)nu * nview * nv), acc((size_t)(nu + 1) * nview * (nv + 1));
    for (int b = 0; b < g.nB; b++) for (int c = 0; c < nc; c++) {
        for (int iu = 0; iu < nu; iu++) for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++) {
            const float* Cc = detCenter + 3 * iview; const float* S = src + 3 * iview;
            R u = (iu - (R)g.off_u - (nu - 1) / R(2)) * (R)g.du;
            R v = (iv - (R)g.off_v - (nv - 1) / R(2)) * (R)g.dv;
            R dxp = (R)Cc[0] + u, dyp = (R)Cc[1] + v, dzp = (R)Cc[2];
            R ddx = (R)S[0] - dxp, ddy = (R)S[1] - dyp, ddz = (R)S[2] - dzp;
            w[((size_t)iu * nview + iview) * nv + iv] = (R)prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nc)]
                * ((R)g.dz / std::fabs(ddz) * std::sqrt(ddz * ddz + ddy * ddy + ddx * ddx));
        }
        std::fill(acc.begin(), acc.end(), R(0));
        for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++) for (int iu = 0; iu < nu; iu++)
            acc[((size_t)(iu + 1) * nview + iview) * (nv + 1) + iv + 1] = acc[((size_t)iu * nview + iview) * (nv + 1) + iv + 1]
                + w[((size_t)iu * nview + iview) * nv + iv];
        for (int iu = 0; iu < nu; iu++) for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++)
            acc[((size_t)(iu + 1) * nview + iview) * (nv + 1) + iv + 1] += acc[((size_t)(iu + 1) * nview + iview) * (nv + 1) + iv];
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++) {
            R x = (ix - (nx - 1) / R(2)) * (R)g.dx + (R)g.cx;
            R y = (iy - (ny - 1) / R(2)) * (R)g.dy + (R)g.cy;
            R z = (iz - (nz - 1) / R(2)) * (R)g.dz + (R)g.cz;
            R val = 0;
            for (int iview = 0; iview < nview; iview++) {
                const float* Cc = detCenter + 3 * iview; const float* S = src + 3 * iview;
                // ProjectConeToDetCart, DistanceDriven3D.cu:478-489
                auto proj = [&](R px, R py, R pz, R& uo, R& vo) {
                    R r = ((R)Cc[2] - (R)S[2]) / (pz - (R)S[2]);
                    R uu = (R)S[0] - (R)Cc[0] + r * (px - (R)S[0]);
                    R vv = (R)S[1] - (R)Cc[1] + r * (py - (R)S[1]);
                    uo = uu / (R)g.du + (R)g.off_u + (nu - 1) / R(2);
                    vo = vv / (R)g.dv + (R)g.off_v + (nv - 1) / R(2);
                };
                R u1, u2, v1, v2, t;
                proj(x - (R)g.dx / 2, y, z, u1, t); proj(x + (R)g.dx / 2, y, z, u2, t);
                proj(x, y - (R)g.dy / 2, z, t, v1); proj(x, y + (R)g.dy / 2, z, t, v2);
                val += (interp_uv(acc.data(), u2, v2, iview, nu + 1, nv + 1, nview) - interp_uv(acc.data(), u2, v1, iview, nu + 1, nv + 1, nview)
                      + interp_uv(acc.data(), u1, v1, iview, nu + 1, nv + 1, nview) - interp_uv(acc.data(), u1, v2, iview, nu + 1, nv + 1, nview))
                      / ((u2 - u1) * (v2 - v1));
            }
            imgOut[img_idx(b, ix, iy, iz, c, nx, ny, nz, nc)] = val;
        }
    }
}

// ---------------------------------------------------------------------------
// Helical BP from rebinned parallel-cone data: Recon3D/HelicalBP.cu:10-83.
// ---------------------------------------------------------------------------
// The voxel's pi-segment start as the reference computes it, in float
// (HelicalBP.cu:26-29: z = iz dz + volZ0; int((z / zrot 2 PI - PI / 4)
// nviewPerPI / PI)), with the multiply-adds contracted to FMAs as nvcc does by
// default.  The value is often an integer in exact arithmetic (for instance
// 8 z - 12 at zrot = 12, 48 views per pi), so its truncation follows the last
// bit of the float computation: computed in FP64 it differs by one view at
// such heights.
inline int helical_istart(int iz, float dz, float volZ0, float zrot, int nviewPerPI)
{
    const float PIf = (float)kPI;
    const float z = std::fma((float)iz, dz, volZ0);
    return (int)(std::fma(z / zrot * 2.0f, PIf, -PIf / 4.0f) * (float)nviewPerPI / PIf);
}

template <class R>
void helical_bp(R* imgOut, const float* prj, int nB, int nC, int nx, int ny, int nz, float dx, float dy, float dz,
                int nu, int nview, int nv, float du, float dv, float off_u, float off_v, float dsd, float dso,
                int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q)
{
    vector<R> tex((size_t)nu * nview * nv);
    for (int b = 0; b < nB; b++) for (int c = 0; c < nC; c++) {
        for (int iu = 0; iu < nu; iu++) for (int iview = 0; iview < nview; iview++) for (int iv = 0; iv < nv; iv++)
            tex[((size_t)iu * nview + iview) * nv + iv] = (R)prj[prj_idx(b, iu, iview, iv, c, nu, nview, nv, nC)];
        #pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++) {
            R x = (ix - (nx - 1) / R(2)) * (R)dx, y = (iy - (ny - 1) / R(2)) * (R)dy, z = iz * (R)dz + (R)volZ0;
            const R PI = (R)kPI;
            const int iStart = helical_istart(iz, dz, volZ0, zrot, nviewPerPI);
            R imgVal = 0;
            for (int i = iStart; i < iStart + nviewPerPI; i++) {
                R totalW = 0, val = 0;
                for (int k = i - mPI * nviewPerPI; k <= i + mPI * nviewPerPI; k += nviewPerPI) {
                    if (k < 0 || k >= nview) continue;
                    R theta = (R)theta0 + k * PI / nviewPerPI;
                    R ct = rcos(theta), st = rsin(theta);
                    R u = x * st - y * ct;
                    R v = (z - (R)zrot * (theta - std::asin(u / (R)dso) - (R)theta0) / 2 / PI) * (R)dsd
                          / (std::sqrt((R)dso * (R)dso - u * u) - (x * ct + y * st));
                    R q = std::fabs(v / (R)dv / nv * R(2));
                    R w;
                    if (q <= (R)Q) w = 1;
                    else if (q < 1) { w = rcos((q - (R)Q) / (1 - (R)Q) * PI / 2); w *= w; }
                    else w = 0;
                    u = u / (R)du + (nu - R(1)) / R(2) + (R)off_u;
                    v = v / (R)dv + (nv - R(1)) / R(2) + (R)off_v;
                    val += tex3<R>(tex.data(), nv, nview, nu, v + R(0.5), k + R(0.5), u + R(0.5), 0, 0, 0) * w;
                    totalW += w;
                }
                imgVal += val / (totalW + R(1e-6f));
            }
            imgOut[(((size_t)b * nx + ix) * ny * nz + iy * nz + iz) * nC + c] = imgVal;
        }
    }
}

} // namespace

// ===========================================================================
// C exports (f64 = parity oracle, f32 = CPU timing baseline)
// ---------------------------------------------------------------------------
// Circular cone-beam DD pair with a per-view main axis (no reference
// counterpart: the reference's DD 3D pair, DistanceDriven3D.cu:139-227, is
// for tomosynthesis with the detector normal to z).  Definition shared with
// the GPU kernels:
//  * a view is an X view iff |C.x - S.x| >= |C.y - S.y| (float32), C the
//    detector centre and S the source; X views integrate over the planes
//    x = voxel-centre x of every voxel layer, Y views over y planes;
//  * detector cell (iu, iv): u = (iu - off_u - (nu-1)/2) du, v likewise, point
//    P(u, v) = C + u U + v V; on plane x = xp the cell's footprint is the
//    rectangle [y(u-), y(u+)] x [z(v-), z(v+)] where y(u+-) are the plane
//    crossings of the rays S -> P(u+-, v_c) and z(v+-) those of S -> P(u_c, v+-);
//  * ray value = sum over planes of the rectangle average of the plane (zero
//    outside the volume) times dx |P_c - S| / |P_c.x - S.x| (dy ... for Y views).
// The adjoint is the exact transpose (a scatter here; a gather on the GPU).
// Parity of this definition against the line integral itself is checked with
// the Siddon cone projector on smooth volumes.
// ---------------------------------------------------------------------------
template <class R, class F>
void dd_cone_ray(const Arb& g, const float* C, const float* U, const float* V, const float* S, int view, int iu,
                 int iv, F f)
{
    const double cx0 = C[3 * view], cy0 = C[3 * view + 1], cz0 = C[3 * view + 2];
    const double sx = S[3 * view], sy = S[3 * view + 1], sz = S[3 * view + 2];
    const bool xv = std::fabs((float)C[3 * view] - (float)S[3 * view]) >= std::fabs((float)C[3 * view + 1] - (float)S[3 * view + 1]);
    const double uc = (iu - g.off_u - (g.nu - 1) / 2.0) * g.du, vc = (iv - g.off_v - (g.nv - 1) / 2.0) * g.dv;
    auto P = [&](double u, double v, double out[3]) {
        for (int k = 0; k < 3; k++) out[k] = C[3 * view + k] + u * U[3 * view + k] + v * V[3 * view + k];
    };
    double Pc[3], Ea[3], Eb[3], Fa[3], Fb[3];
    P(uc, vc, Pc);
    P(uc - g.du / 2, vc, Ea); P(uc + g.du / 2, vc, Eb);
    P(uc, vc - g.dv / 2, Fa); P(uc, vc + g.dv / 2, Fb);
    (void)cx0; (void)cy0; (void)cz0;
    const double Sv[3] = {sx, sy, sz};
    const int ax = xv ? 0 : 1, bx = xv ? 1 : 0;       // main axis, in-plane axis
    const int nmain = xv ? g.nx : g.ny, nin = xv ? g.ny : g.nx;
    const double dmain = xv ? g.dx : g.dy, din = xv ? g.dy : g.dx, cmain = xv ? g.cx : g.cy, cin = xv ? g.cy : g.cx;
    const double len = std::sqrt((Pc[0] - sx) * (Pc[0] - sx) + (Pc[1] - sy) * (Pc[1] - sy) + (Pc[2] - sz) * (Pc[2] - sz));
    const double dpc = Pc[ax] - Sv[ax];
    if (std::fabs(dpc) < 1e-9) return;
    const double L = dmain * len / std::fabs(dpc);
    for (int i = 0; i < nmain; i++) {
        const double xp = (i + 0.5 - nmain / 2.0) * dmain + cmain;
        auto cross = [&](const double E[3], int axis, double c0, double d0, int n0) {
            const double t = (xp - Sv[ax]) / (E[ax] - Sv[ax]);
            return (Sv[axis] + t * (E[axis] - Sv[axis]) - c0) / d0 + n0 / 2.0;
        };
        double y1 = cross(Ea, bx, cin, din, nin), y2 = cross(Eb, bx, cin, din, nin);
        double z1 = cross(Fa, 2, g.cz, g.dz, g.nz), z2 = cross(Fb, 2, g.cz, g.dz, g.nz);
        if (y1 > y2) std::swap(y1, y2);
        if (z1 > z2) std::swap(z1, z2);
        const double area = (y2 - y1) * (z2 - z1);
        if (!(area > 0)) continue;
        const double ly = std::max(y1, 0.0), hy = std::min(y2, (double)nin);
        const double lz = std::max(z1, 0.0), hz = std::min(z2, (double)g.nz);
        if (!(hy > ly) || !(hz > lz)) continue;
        for (int j = (int)ly; j <= std::min((int)hy, nin - 1); j++) {
            const double oy = std::min(hy, j + 1.0) - std::max(ly, (double)j);
            if (oy <= 0) continue;
            for (int k = (int)lz; k <= std::min((int)hz, g.nz - 1); k++) {
                const double oz = std::min(hz, k + 1.0) - std::max(lz, (double)k);
                if (oz <= 0) continue;
                const long ix = xv ? i : j, iy = xv ? j : i;
                f(((long)ix * g.ny + iy) * g.nz + k, (R)(oy * oz / area * L));
            }
        }
    }
}

template <class R>
void dd_cone_fp(R* prj, const float* img, const float* C, const float* U, const float* V, const float* S, const Arb& g)
{
    const long nvox = (long)g.nx * g.ny * g.nz;
    for (int b = 0; b < g.nB; b++)
        for (int c = 0; c < g.nC; c++) {
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
            for (int iu = 0; iu < g.nu; iu++)
                for (int view = 0; view < g.nview; view++)
                    for (int iv = 0; iv < g.nv; iv++) {
                        R acc = 0;
                        dd_cone_ray<R>(g, C, U, V, S, view, iu, iv, [&](long vox, R w) {
                            acc += w * (R)img[((long)b * nvox + vox) * g.nC + c];
                        });
                        prj[((((long)b * g.nu + iu) * g.nview + view) * g.nv + iv) * g.nC + c] = acc;
                    }
        }
}

template <class R>
void dd_cone_adj(R* img, const float* prj, const float* C, const float* U, const float* V, const float* S, const Arb& g)
{
    const long nvox = (long)g.nx * g.ny * g.nz;
    std::fill(img, img + nvox * g.nB * g.nC, R(0));
    for (int b = 0; b < g.nB; b++)
        for (int c = 0; c < g.nC; c++)
            for (int view = 0; view < g.nview; view++)
                for (int iu = 0; iu < g.nu; iu++)
                    for (int iv = 0; iv < g.nv; iv++) {
                        const R p = (R)prj[((((long)b * g.nu + iu) * g.nview + view) * g.nv + iv) * g.nC + c];
                        if (p == 0) continue;
                        dd_cone_ray<R>(g, C, U, V, S, view, iu, iv, [&](long vox, R w) {
                            img[((long)b * nvox + vox) * g.nC + c] += w * p;
                        });
                    }
}

// ---------------------------------------------------------------------------
// Isotropic TV separable surrogate: PriorFunctionSolver/TVSQS.cu:8-100
// (2D: Variation2DKernel, TVSQS2DKernel) and :165-288 (3D).  Layout
// [nb][nx][ny][nz][nc]; backward differences are 0 at the low edge, forward
// ones 0 at the high edge where the neighbour's variation is replaced by the
// voxel's own (TVSQS.cu:64-80,223-257); s2 = dims / var + sum 1 / var_next.
// ---------------------------------------------------------------------------
template <class R>
void tv_iso_sqs(R* s1, R* s2, R* var, const float* img, int nb, int nx, int ny, int nz, int nc, int dims, R eps)
{
    if (dims == 2) nz = 1;
    const long sx = (long)ny * nz * nc, sy = (long)nz * nc, sz = nc;
    const long n = (long)nb * nx * ny * nz * nc;
    auto at = [&](long i) { return (R)img[i]; };
    for (long i = 0; i < n; i++) {
        const long r = i % ((long)nx * ny * nz * nc);
        const int iz = (int)((r / nc) % nz), iy = (int)((r / sy) % ny), ix = (int)(r / sx);
        const R x = at(i);
        const R dx = ix > 0 ? x - at(i - sx) : R(0);
        const R dy = iy > 0 ? x - at(i - sy) : R(0);
        const R dz = (dims == 3 && iz > 0) ? x - at(i - sz) : R(0);
        var[i] = std::sqrt(dx * dx + dy * dy + dz * dz + eps);
    }
    for (long i = 0; i < n; i++) {
        const long r = i % ((long)nx * ny * nz * nc);
        const int iz = (int)((r / nc) % nz), iy = (int)((r / sy) % ny), ix = (int)(r / sx);
        const R x = at(i), v = var[i];
        R d0 = 0, g1 = 0, c1 = 0;
        auto axis = [&](int idx, int nn, long st) {
            if (idx > 0) d0 += x - at(i - st);
            const bool up = idx < nn - 1;
            const R va = up ? var[i + st] : v;
            if (up) g1 += (x - at(i + st)) / va;
            c1 += R(1) / va;
        };
        axis(ix, nx, sx);
        axis(iy, ny, sy);
        if (dims == 3) axis(iz, nz, sz);
        s1[i] = d0 / v + g1;
        s2[i] = R(dims) / v + c1;
    }
}

// ===========================================================================
#define FAN_ARGS int nB, int nC, int nx, int ny, int nz, float dx, float dy, float dz, \
    int nu, int nview, int nv, float da, float dv, float off_a, float off_v, float dsd, float dso, int type
#define FAN_MAKE Fan{ nB, nC, nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso, type }
#define ARB_ARGS int nB, int nC, int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz, \
    int nu, int nview, int nv, float du, float dv, float off_u, float off_v
#define ARB_MAKE Arb{ nB, nC, nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v }

#define EXPORT_BOTH(NAME, BODY_T) \
    extern "C" void NAME##_f64 BODY_T(double) \
    extern "C" void NAME##_f32 BODY_T(float)

extern "C" {

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_split_views(int* cls, const float* deg, int nview)
{
    vector<int> vx, vy; split_views(deg, nview, vx, vy);
    for (int i = 0; i < nview; i++) cls[i] = 0;
    for (int i : vy) cls[i] = 1;
    return (int)vx.size();
}

#define DEF_FAN(NAME, FN, OUTT) \
    void NAME(OUTT* out, const float* in, const float* deg, FAN_ARGS) { FN<OUTT>(out, in, deg, FAN_MAKE); }
DEF_FAN(orc_dd_fan_fp_f64, dd_fan_fp, double)
DEF_FAN(orc_dd_fan_fp_f32, dd_fan_fp, float)
DEF_FAN(orc_dd_fan_bp_f64, dd_fan_bp, double)
DEF_FAN(orc_dd_fan_bp_f32, dd_fan_bp, float)
DEF_FAN(orc_siddon_fan_fp_f64, siddon_fan_fp, double)
DEF_FAN(orc_siddon_fan_fp_f32, siddon_fan_fp, float)
DEF_FAN(orc_siddon_fan_bp_f64, siddon_fan_bp, double)
DEF_FAN(orc_siddon_fan_bp_f32, siddon_fan_bp, float)
DEF_FAN(orc_pixel_fan_bp_f64, pixel_fan_bp, double)
DEF_FAN(orc_pixel_fan_bp_f32, pixel_fan_bp, float)

void orc_ramp_kernel_f64(double* kr, int nu, float da, int filterType, int equalSpace)
{
    vector<double> k; ramp_kernel<double>(k, nu, (double)da, filterType, equalSpace != 0);
    std::copy(k.begin(), k.end(), kr);
}

void orc_filter_f64(double* out, const float* prj, int nB, int nC, int nu, int nview, int nv, float da, float dv,
                    float off_a, float off_v, float dsd, float dso, int typeFilter, int fan)
{ (void)dv; (void)off_v; ramp_filter<double>(out, prj, nB, nC, nu, nview, nv, da, off_a, dsd, dso, typeFilter, fan != 0); }
void orc_filter_f32(float* out, const float* prj, int nB, int nC, int nu, int nview, int nv, float da, float dv,
                    float off_a, float off_v, float dsd, float dso, int typeFilter, int fan)
{ (void)dv; (void)off_v; ramp_filter<float>(out, prj, nB, nC, nu, nview, nv, da, off_a, dsd, dso, typeFilter, fan != 0); }

double orc_tv_sqs_f64(double* grad, double* norm, const float* img, int nx, int ny)
{ return tv_sqs<double>(grad, norm, img, nx, ny); }
float orc_tv_sqs_f32(float* grad, float* norm, const float* img, int nx, int ny)
{ return tv_sqs<float>(grad, norm, img, nx, ny); }

void orc_os_sqs_f64(double* res, const float* img, const float* prj, const float* normImg, const float* wprj,
                    const float* deg, float lb, float ub, int nviewPerSubset, int nSubsets, FAN_ARGS, float lambda)
{ os_sqs<double>(res, img, prj, normImg, wprj, deg, lb, ub, nviewPerSubset, nSubsets, FAN_MAKE, lambda); }
void orc_os_sqs_f32(float* res, const float* img, const float* prj, const float* normImg, const float* wprj,
                    const float* deg, float lb, float ub, int nviewPerSubset, int nSubsets, FAN_ARGS, float lambda)
{ os_sqs<float>(res, img, prj, normImg, wprj, deg, lb, ub, nviewPerSubset, nSubsets, FAN_MAKE, lambda); }

void orc_siddon_par_fp_f64(double* prj, const float* img, const float* C, const float* U, const float* V,
                           const float* dir, ARB_ARGS)
{ siddon_arb_fp<double>(prj, img, C, U, V, dir, true, ARB_MAKE); }
void orc_siddon_par_bp_f64(double* img, const float* prj, const float* C, const float* U, const float* V,
                           const float* dir, ARB_ARGS)
{ siddon_arb_bp<double>(img, prj, C, U, V, dir, true, ARB_MAKE); }
void orc_siddon_cone_fp_f64(double* prj, const float* img, const float* C, const float* U, const float* V,
                            const float* src, ARB_ARGS)
{ siddon_arb_fp<double>(prj, img, C, U, V, src, false, ARB_MAKE); }
void orc_siddon_cone_bp_f64(double* img, const float* prj, const float* C, const float* U, const float* V,
                            const float* src, ARB_ARGS)
{ siddon_arb_bp<double>(img, prj, C, U, V, src, false, ARB_MAKE); }
void orc_siddon_cone_fp_f32(float* prj, const float* img, const float* C, const float* U, const float* V,
                            const float* src, ARB_ARGS)
{ siddon_arb_fp<float>(prj, img, C, U, V, src, false, ARB_MAKE); }
void orc_siddon_cone_bp_f32(float* img, const float* prj, const float* C, const float* U, const float* V,
                            const float* src, ARB_ARGS)
{ siddon_arb_bp<float>(img, prj, C, U, V, src, false, ARB_MAKE); }

void orc_os_sqs_cone_f64(double* res, const float* img, const float* prj, const float* normImg, const float* mask,
                         const float* C, const float* U, const float* V, const float* src, float lb, float ub,
                         int nviewPerSubset, int nSubsets, ARB_ARGS, float lambda)
{ os_sqs_cone<double>(res, img, prj, normImg, mask, C, U, V, src, lb, ub, nviewPerSubset, nSubsets, ARB_MAKE, lambda); }

void orc_dd_tomo_fp_f64(double* prj, const float* img, const float* C, const float* src, ARB_ARGS)
{ dd_tomo_fp<double>(prj, img, C, src, ARB_MAKE); }
void orc_dd_tomo_bp_f64(double* img, const float* prj, const float* C, const float* src, ARB_ARGS)
{ dd_tomo_bp<double>(img, prj, C, src, ARB_MAKE); }
void orc_dd_tomo_fp_f32(float* prj, const float* img, const float* C, const float* src, ARB_ARGS)
{ dd_tomo_fp<float>(prj, img, C, src, ARB_MAKE); }
void orc_dd_tomo_bp_f32(float* img, const float* prj, const float* C, const float* src, ARB_ARGS)
{ dd_tomo_bp<float>(img, prj, C, src, ARB_MAKE); }

void orc_dd_cone_fp_f64(double* prj, const float* img, const float* C, const float* U, const float* V,
                        const float* src, ARB_ARGS)
{ dd_cone_fp<double>(prj, img, C, U, V, src, ARB_MAKE); }
void orc_dd_cone_adj_f64(double* img, const float* prj, const float* C, const float* U, const float* V,
                         const float* src, ARB_ARGS)
{ dd_cone_adj<double>(img, prj, C, U, V, src, ARB_MAKE); }
void orc_dd_cone_fp_f32(float* prj, const float* img, const float* C, const float* U, const float* V,
                        const float* src, ARB_ARGS)
{ dd_cone_fp<float>(prj, img, C, U, V, src, ARB_MAKE); }
void orc_dd_cone_adj_f32(float* img, const float* prj, const float* C, const float* U, const float* V,
                         const float* src, ARB_ARGS)
{ dd_cone_adj<float>(img, prj, C, U, V, src, ARB_MAKE); }

void orc_tv_iso_sqs_f64(double* s1, double* s2, double* var, const float* img, int nb, int nx, int ny, int nz, int nc,
                        int dims, double eps)
{ tv_iso_sqs<double>(s1, s2, var, img, nb, nx, ny, nz, nc, dims, eps); }

void orc_helical_bp_f64(double* img, const float* prj, int nB, int nC, int nx, int ny, int nz, float dx, float dy, float dz,
                        int nu, int nview, int nv, float du, float dv, float off_u, float off_v, float dsd, float dso,
                        int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q)
{ helical_bp<double>(img, prj, nB, nC, nx, ny, nz, dx, dy, dz, nu, nview, nv, du, dv, off_u, off_v, dsd, dso,
                     nviewPerPI, theta0, volZ0, zrot, mPI, Q); }

} // extern "C"
