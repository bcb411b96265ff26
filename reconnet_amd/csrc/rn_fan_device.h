// Device helpers shared by the fan-beam kernels (line parameters of a ray,
// exact footprint averages and their transposes).
#pragma once

#include "rn_fan.h"

namespace rn {
namespace fan {

// Element at byte offset `boff` (32-bit) from a wave-uniform base: written so
// that the load is addressed as scalar base + 32-bit lane offset, with one
// offset computation shared by loads from several bases.  Callers keep every
// byte offset below 2^31 (checked on the host).
template <class T>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, unsigned boff)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

// Values a lane carries: one slice (float, or double for the FP64
// normaliser) or four interleaved slices (float4, each component with the
// arithmetic of the float case).
template <class T>
__device__ __forceinline__ T vzero() { return T(0); }
template <>
__device__ __forceinline__ float4 vzero<float4>() { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
__device__ __forceinline__ void vadd(float& a, float b) { a += b; }
__device__ __forceinline__ void vadd(double& a, double b) { a += b; }
__device__ __forceinline__ void vadd(float4& a, float4 b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; }
// a += v * w (w converted to the value type first)
__device__ __forceinline__ void vmadd(float& a, float v, float w) { a += v * w; }
__device__ __forceinline__ void vmadd(double& a, double v, float w) { a += v * (double)w; }
__device__ __forceinline__ void vmadd(float4& a, float4 v, float w)
{
    a.x += v.x * w; a.y += v.y * w; a.z += v.z * w; a.w += v.w * w;
}
__device__ __forceinline__ float vscale(float v, float f) { return v * f; }
__device__ __forceinline__ double vscale(double v, float f) { return v * (double)f; }
__device__ __forceinline__ float4 vscale(float4 v, float f) { return make_float4(v.x * f, v.y * f, v.z * f, v.w * f); }

// ---------------------------------------------------------------------------
// geometry helpers
// ---------------------------------------------------------------------------

// Ray of detector angle a (sin sa, cos ca) in view t (cos ct, sin st):
// X views:  X(Y) = A + B (Y - ny/2)   (absolute image coordinates)
__device__ __forceinline__ void xline(const FanK& g, float ct, float st, float sa, float ca,
                                      float& A, float& B, float& cu)
{
    cu = ca * ct + sa * st;            // cos(a - t)
    const float su = sa * ct - ca * st;  // sin(a - t)
    const float inv = 1.0f / cu;
    A = 0.5f * g.nx + (g.dso / g.dx) * sa * inv;
    B = (g.dy / g.dx) * su * inv;
}

// Y views:  Y(X) = A + B (X - nx/2)
__device__ __forceinline__ void yline(const FanK& g, float ct, float st, float sa, float ca,
                                      float& A, float& B, float& su)
{
    const float cu = ca * ct + sa * st;
    su = sa * ct - ca * st;
    const float inv = 1.0f / su;
    A = 0.5f * g.ny - (g.dso / g.dy) * sa * inv;
    B = (g.dx / g.dy) * cu * inv;
}

// acc[q] += scale * the average of row q (`rboff` bytes after the
// wave-uniform base rows[q]) over one interval [lo, hi] of its n
// piecewise-constant cells (zero outside [0, n)): the cells, their overlaps and the
// combined weight overlap * scale / width are computed once for the S rows,
// and each row adds one FMA per cell (a point interval, width 0, adds its
// cell times scale).
template <class T, int S>
__device__ __forceinline__ void seg_acc_rows(const T* const (&rows)[S], unsigned rboff, int n, float lo, float hi,
                                             float scale, T (&acc)[S])
{
    const float w = hi - lo;
    const float l = fmaxf(lo, 0.0f), h = fminf(hi, (float)n);
    if (!(h > l)) {
        if (w > 0.0f || !(lo >= 0.0f && lo < (float)n)) return;
        const unsigned boff = rboff + (unsigned)(int)lo * (unsigned)sizeof(T);
        #pragma unroll
        for (int q = 0; q < S; ++q) vmadd(acc[q], ld_at(rows[q], boff), scale);
        return;
    }
    const float k = scale / w;
    const int m0 = (int)l;
    const int m1 = min((int)h, n - 1);
    for (int m = m0; m <= m1; ++m) {
        const float pl = fmaxf(l, (float)m), ph = fminf(h, (float)(m + 1));
        const unsigned boff = rboff + (unsigned)m * (unsigned)sizeof(T);
        #pragma unroll
        for (int q = 0; q < S; ++q) vmadd(acc[q], ld_at(rows[q], boff), (ph - pl) * k);
    }
}

// seg_acc_rows for G rows staged in LDS: row q of `st` holds bins wlo ..
// wlo + CAP - 1 (CAP float4 entries).  The first NB bins of the interval are
// taken without branches (a bin past its end with weight 0, its index
// clamped into the interval), further bins in a loop.
template <int G, int CAP, int NB>
__device__ __forceinline__ void seg_acc_staged(const float4* st, int wlo, int n, float lo, float hi, float scale,
                                               float4 (&acc)[G])
{
    const float w = hi - lo;
    const float l = fmaxf(lo, 0.0f), h = fminf(hi, (float)n);
    const bool normal = h > l;
    const bool point = !normal && !(w > 0.0f) && lo >= 0.0f && lo < (float)n;
    // (an interval that reads nothing: m0 = wlo, m1 = wlo - 1, so that no
    // float outside the int range is converted)
    const int m0 = normal ? (int)l : (point ? (int)lo : wlo);
    const int m1 = normal ? min((int)h, n - 1) : (point ? m0 : m0 - 1);
    const int t0 = m0 - wlo, tl = max(m1 - wlo, 0);
    const float k = normal ? scale / w : scale;
    #pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int m = m0 + j;
        const bool on = m <= m1;
        const int t = on ? t0 + j : tl;
        const float ov = normal ? fminf(h, (float)(m + 1)) - fmaxf(l, (float)m) : 1.0f;
        const float c = on ? ov * k : 0.0f;
        #pragma unroll
        for (int q = 0; q < G; ++q) vmadd(acc[q], st[q * CAP + t], c);
    }
    for (int m = m0 + NB; m <= m1; ++m) {
        const float pl = fmaxf(l, (float)m), ph = fminf(h, (float)(m + 1));
        const float c = (ph - pl) * k;
        #pragma unroll
        for (int q = 0; q < G; ++q) vmadd(acc[q], st[q * CAP + m - wlo], c);
    }
}

// The bins seg_acc_rows reads for the interval [lo, hi] of a row of n bins:
// m0 .. m1, or m0 = n, m1 = -1 when it reads none.
__device__ __forceinline__ void seg_avg_bins(int n, float lo, float hi, int& m0, int& m1)
{
    const float w = hi - lo;
    const float l = fmaxf(lo, 0.0f), h = fminf(hi, (float)n);
    if (h > l) {
        m0 = (int)l;
        m1 = min((int)h, n - 1);
    } else if (!(w > 0.0f) && lo >= 0.0f && lo < (float)n) {
        m0 = m1 = (int)lo;
    } else {
        m0 = n;
        m1 = -1;
    }
}

// conservative half width (in bins) of the set of rays through a disc of
// radius r at distance dist from the source:  asin(r/d) <= r/(d-r)
__device__ __forceinline__ float half_width_bins(float r, float dist, float inv_da)
{
    return r * inv_da / fmaxf(dist - r, 1e-3f) + 0.02f;
}

// ---------------------------------------------------------------------------
// per-ray forward projections and per-pixel gathers
// ---------------------------------------------------------------------------

// Sum over the cells of a line of n cells (zero outside [0, n)) of value times
// overlap with [lo, hi] (no division: the caller scales by L / |B|).
template <class T>
__device__ __forceinline__ T seg_sum(const T* __restrict__ line, int n, float lo, float hi)
{
    const float l = fmaxf(lo, 0.0f), h = fminf(hi, (float)n);
    T s = vzero<T>();
    if (!(h > l)) return s;
    const int m0 = (int)l;
    const int m1 = min((int)h, n - 1);
    for (int m = m0; m <= m1; ++m) {
        const float pl = fmaxf(l, (float)m), ph = fminf(h, (float)(m + 1));
        vmadd(s, line[m], ph - pl);
    }
    return s;
}

// seg_sum for hi - lo < 2 (at most three cells), without branches: the three
// loads are independent and every weight max(0, min(h, m+1) - max(l, m)) is
// zero for cells outside [l, h] (indices are clamped into the line).
// Siddon FP (exact ray / pixel intersection lengths), one line per image row
// (X views) or column (Y views): the ray crosses line j over [j, j+1] of the
// line axis, i.e. over [lo, hi] = [A + B (j - n/2), A + B (j + 1 - n/2)] of the
// cell axis, with path length L per line; the length inside a cell is L times
// the cell's share of [lo, hi], overlap / (hi - lo) with hi - lo taken from the
// same rounded end points (so a line's shares add up to one).  When the end
// points round to the same value (a ray along the line axis) the cell holding
// lo gets the whole length.
template <class T>
__device__ __forceinline__ T siddon_line_sum(const T* __restrict__ line, int n, float lo, float hi)
{
    const float w = hi - lo;
    if (w > 0.0f) return vscale(seg_sum<T>(line, n, lo, hi), __builtin_amdgcn_rcpf(w));
    return (lo >= 0.0f && lo < (float)n) ? line[(int)lo] : vzero<T>();
}

// A line of at most three cells for hi - lo < 2: the cells m0 .. m0 + 2 (indices
// clamped into the line) and their shares max(0, min(h, m + 1) - max(l, m))
// / (hi - lo), which vanish outside [l, h]; a zero-width interval gives the
// cell holding lo the whole share.  The cells and shares depend on the
// geometry only, so one Line3 serves every slice of a batch.
struct Line3 {
    unsigned i0, i1, i2;
    float w0, w1, w2;
};

__device__ __forceinline__ Line3 line3(int n, float lo, float hi)
{
    const float fn = (float)n;
    const float w = hi - lo;
    const float l = fminf(fmaxf(lo, 0.0f), fn), h = fminf(hi, fn);
    const int m0 = (int)l;                 // l in [m0, m0 + 1) unless l == n (then every weight is 0)
    const float f0 = (float)m0;
    const bool pos = w > 0.0f;
    const float inv = __builtin_amdgcn_rcpf(pos ? w : 1.0f);
    Line3 c;
    const float o0 = fmaxf(fminf(h, f0 + 1.0f) - l, 0.0f);
    c.w0 = (pos ? o0 : ((lo >= 0.0f && lo < fn) ? 1.0f : 0.0f)) * inv;
    c.w1 = fmaxf(fminf(h, f0 + 2.0f) - (f0 + 1.0f), 0.0f) * inv;
    c.w2 = fmaxf(fminf(h, f0 + 3.0f) - (f0 + 2.0f), 0.0f) * inv;
    const unsigned last = (unsigned)(n - 1);
    c.i0 = min((unsigned)m0, last);
    c.i1 = min((unsigned)m0 + 1u, last);
    c.i2 = min((unsigned)m0 + 2u, last);
    return c;
}

// acc += v0 w0 + v1 w1 [+ v2 w2] per value component (fused multiply-adds)
__device__ __forceinline__ void lmadd(float& a, float v0, float v1, float v2, const Line3& c, bool three)
{
    a = fmaf(v0, c.w0, a);
    a = fmaf(v1, c.w1, a);
    if (three) a = fmaf(v2, c.w2, a);
}
__device__ __forceinline__ void lmadd(double& a, double v0, double v1, double v2, const Line3& c, bool three)
{
    a = fma(v0, (double)c.w0, a);
    a = fma(v1, (double)c.w1, a);
    if (three) a = fma(v2, (double)c.w2, a);
}
__device__ __forceinline__ void lmadd(float4& a, float4 v0, float4 v1, float4 v2, const Line3& c, bool three)
{
    lmadd(a.x, v0.x, v1.x, v2.x, c, three);
    lmadd(a.y, v0.y, v1.y, v2.y, c, three);
    lmadd(a.z, v0.z, v1.z, v2.z, c, three);
    lmadd(a.w, v0.w, v1.w, v2.w, c, three);
}

// acc += the line starting `lboff` bytes after a wave-uniform base, over the
// first NC cells (NC = 2 when the interval is known to be narrower than one
// cell: the third share is then exactly 0 and its term is dropped)
template <class T, int NC>
__device__ __forceinline__ void line_madd_at(T& acc, const T* __restrict__ base, unsigned lboff, const Line3& c)
{
    const T v0 = ld_at(base, lboff + c.i0 * (unsigned)sizeof(T));
    const T v1 = ld_at(base, lboff + c.i1 * (unsigned)sizeof(T));
    const T v2 = NC == 2 ? v1 : ld_at(base, lboff + c.i2 * (unsigned)sizeof(T));
    lmadd(acc, v0, v1, v2, c, NC == 3);
}

// The narrow-ray line loop of siddon_fp_ray_slices: lines wj0 .. wj1 of the S
// slice bases; line j ends where line j + 1 starts (base + j + 1 is exact, so
// the shared end point is the same FMA result for both lines)
template <class T, int S, int NC>
__device__ __forceinline__ void fp_narrow_lines(const T* const (&Pq)[S], int n, float4 r, float base, int wj0,
                                                int wj1, T (&acc)[S])
{
    float ea = fmaf(r.y, base + (float)wj0, r.x);
    constexpr int kUnroll = S > 4 ? 1 : (S > 2 ? 2 : 4);
    #pragma unroll kUnroll
    for (int j = wj0; j <= wj1; ++j) {
        const float eb = fmaf(r.y, base + (float)(j + 1), r.x);
        const Line3 c = line3(n, fminf(ea, eb), fmaxf(ea, eb));
        const unsigned lboff = (unsigned)(j * n) * (unsigned)sizeof(T);
        #pragma unroll
        for (int q = 0; q < S; ++q) line_madd_at<T, NC>(acc[q], Pq[q], lboff, c);
        ea = eb;
    }
}

// Minimum and maximum of v over the 64 lanes of a wave (every lane active):
// pair, quad, half-row and row exchanges by DPP, then the four rows' values
// read into scalar registers, so the result is wave-uniform.  The exchanges'
// "old" operand is the operation's identity, so that each one compiles to a
// single min / max with a DPP source.
__device__ __forceinline__ int wave_min(int v)
{
    constexpr int id = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xf, 0xf, false));    // quad_perm [1, 0, 3, 2]
    v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xf, 0xf, false));    // quad_perm [2, 3, 0, 1]
    v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));   // row_half_mirror
    v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));   // row_mirror
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_max(int v)
{
    constexpr int id = (int)0x80000000;
    v = max(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// One ray (bin iu of `view`) through S slices of one geometry: slice q reads
// planes Pyx + q * plane / Pxy + q * plane (slices q >= nsl read slice nsl - 1;
// the caller discards their sums).  Every lane of the wavefront calls it
// (inactive lanes with active = false and any valid iu).  The lines are walked
// over the union of the lanes' line ranges, so the loop and the line
// addresses are wave-uniform; lanes outside their own range add exact zeros.
// The cells met and their weights are computed once per line and applied to
// every slice, each slice with the arithmetic of a single-slice call.
template <class T, int S>
__device__ __forceinline__ void siddon_fp_ray_slices(const FanK& g, const T* __restrict__ Pyx,
                                                     const T* __restrict__ Pxy, long plane, int nsl, int view,
                                                     int iu, bool active, T (&acc)[S])
{
    const float4 r = g.sray[(long)view * g.nu + iu];
    const bool xv = g.vcls[view] == 0;
    const T* __restrict__ P = xv ? Pyx : Pxy;
    const int nl = xv ? g.ny : g.nx;     // lines
    const int n = xv ? g.nx : g.ny;      // cells per line
    const float base = -0.5f * (float)nl;
    // lines whose interval can meet [0, n]: A + B (j - nl/2) in [-|B| - 1, n + 1]
    // (one line of margin on each side; the other lines add exact zeros)
    int j0 = 0, j1 = nl - 1;
    if (r.y != 0.0f) {
        const float ib = 1.0f / r.y;
        const float ta = (-fabsf(r.y) - 1.0f - r.x) * ib - base, tb = ((float)n + 1.0f - r.x) * ib - base;
        const float tlo = fminf(fmaxf(fminf(ta, tb) - 1.0f, 0.0f), (float)nl);
        const float thi = fminf(fmaxf(ta, tb) + 1.0f, (float)(nl - 1));
        j0 = (int)tlo;
        j1 = thi >= tlo ? (int)thi : -1;
    } else if (!(r.x >= 0.0f && r.x < (float)n)) {
        j1 = -1;
    }
    if (!active) { j0 = nl; j1 = -1; }
    const int wj0 = wave_min(j0);
    const int wj1 = wave_max(j1);
    const T* Pq[S];                      // wave-uniform slice bases
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        Pq[q] = P + (long)min(q, nsl - 1) * plane;
        acc[q] = vzero<T>();
    }
    // [lo, hi] (rounded) spans at most three cells on every lane (two when
    // |B| < 1: the rounding error of the end points is far below 1e-3)
    const bool narrow = __all(!active || fabsf(r.y) <= 1.999f);
    if (narrow) {
        if (__all(!active || fabsf(r.y) <= 0.999f)) fp_narrow_lines<T, S, 2>(Pq, n, r, base, wj0, wj1, acc);
        else fp_narrow_lines<T, S, 3>(Pq, n, r, base, wj0, wj1, acc);
    } else {
        for (int j = wj0; j <= wj1; ++j) {
            const float lr = base + (float)j;
            const float ea = fmaf(r.y, lr, r.x), eb = fmaf(r.y, lr + 1.0f, r.x);
            #pragma unroll
            for (int q = 0; q < S; ++q)
                vadd(acc[q], siddon_line_sum<T>(Pq[q] + (long)j * n, n, fminf(ea, eb), fmaxf(ea, eb)));
        }
    }
    #pragma unroll
    for (int q = 0; q < S; ++q) acc[q] = vscale(acc[q], r.z);
}

// The padded-line variant of the narrow line loop for G interleaved slice
// groups (float4 cells; padded line j of group k starts at Pq[k] + j * pitch
// with kQPadL zero cells, so cell m is at + j * pitch + kQPadL + m, and ends
// with kQPadR zero cells; every offset from Pq[k] is non-negative).
// The crossing interval [lo, hi] of a line spans the cells m0 = floor(lo) ..
// m0 + NC - 1; the cell index is only clamped into the padded line, so an
// interval outside the image reads pad zeros, and the shares are
// (min(hi, m0 + 1) - lo, ..) / (hi - lo) without clamping to [0, n].  A
// zero-width interval gives cell m0 the whole share.
template <int G, int NC, bool WIDE>
__device__ __forceinline__ void fp_quad_lines_padded(const float4* const (&Pq)[G], int n, int pitch, float4 r,
                                                     float base, int wj0, int wj1, float4 (&acc)[G])
{
    float ea = fmaf(r.y, base + (float)wj0, r.x);
    float fj = base + (float)(wj0 + 1);
    #pragma unroll 2
    for (int j = wj0; j <= wj1; ++j) {
        const float eb = fmaf(r.y, fj, r.x);
        fj += 1.0f;
        const float lo = fminf(ea, eb), hi = fmaxf(ea, eb);
        const float w = hi - lo;
        const float f0 = floorf(lo);
        // clamped into the padded line: a clamped index only reads pad cells
        // (kQPadL, kQPadR >= NC), whatever the shares
        const int m0 = (int)__builtin_amdgcn_fmed3f(f0, -(float)kQPadL, (float)(n + kQPadR - NC));
        const bool pos = WIDE || w > 0.0f;   // WIDE: every interval of the view has w > 0
        const float inv = __builtin_amdgcn_rcpf(pos ? w : 1.0f);
        const float o0 = fminf(hi, f0 + 1.0f) - lo;
        float w0, w1, w2 = 0.0f;
        if (NC == 2) {
            w0 = (pos ? o0 : 1.0f) * inv;
            w1 = (w - o0) * inv;
        } else {
            const float o1 = fmaxf(fminf(hi, f0 + 2.0f) - (f0 + 1.0f), 0.0f);
            w0 = (pos ? o0 : 1.0f) * inv;
            w1 = o1 * inv;
            w2 = (w - o0 - o1) * inv;
        }
        const unsigned boff = (unsigned)(j * pitch + kQPadL + m0) * 16u;   // m0 >= -kQPadL
        #pragma unroll
        for (int k = 0; k < G; ++k) {
            const float4 v0 = ld_at(Pq[k], boff), v1 = ld_at(Pq[k], boff + 16u);
            float4& a = acc[k];
            a.x = fmaf(v0.x, w0, a.x); a.y = fmaf(v0.y, w0, a.y); a.z = fmaf(v0.z, w0, a.z); a.w = fmaf(v0.w, w0, a.w);
            a.x = fmaf(v1.x, w1, a.x); a.y = fmaf(v1.y, w1, a.y); a.z = fmaf(v1.z, w1, a.z); a.w = fmaf(v1.w, w1, a.w);
            if (NC == 3) {
                const float4 v2 = ld_at(Pq[k], boff + 32u);
                a.x = fmaf(v2.x, w2, a.x); a.y = fmaf(v2.y, w2, a.y);
                a.z = fmaf(v2.z, w2, a.z); a.w = fmaf(v2.w, w2, a.w);
            }
        }
        ea = eb;
    }
}

// One ray through G interleaved slice groups of padded quad images (Qyx for
// X views, Qxy for Y views; groups qstride cells apart, groups k >= nq read
// group nq - 1): the line range, wave-uniform loop and final scaling of
// siddon_fp_ray_slices with the padded line loop above (rays with |B| >= 2,
// none in ordinary geometries, take the clamped general path).  H = 2 or 4:
// the H parts of 64 / H lanes of the wave carry the same rays and walk
// consecutive parts of the wave's line range; their sums are added across the
// parts (every lane then holds the ray's total).
template <int G, int H = 1>
__device__ __forceinline__ void siddon_fp_ray_quads(const FanK& g, const float4* __restrict__ Qyx,
                                                    const float4* __restrict__ Qxy, int nq, int view, int iu,
                                                    bool active, float4 (&acc)[G])
{
    const float4 r = g.sray[(long)view * g.nu + iu];
    const bool xv = g.vcls[view] == 0;
    const float4* __restrict__ P = xv ? Qyx : Qxy;
    const int nl = xv ? g.ny : g.nx;     // lines
    const int n = xv ? g.nx : g.ny;      // cells per line
    const int pitch = qpitch(n);
    const long qstride = (long)nl * pitch;
    const float base = -0.5f * (float)nl;
    int j0 = 0, j1 = nl - 1;
    if (r.y != 0.0f) {
        const float ib = 1.0f / r.y;
        const float ta = (-fabsf(r.y) - 1.0f - r.x) * ib - base, tb = ((float)n + 1.0f - r.x) * ib - base;
        const float tlo = fminf(fmaxf(fminf(ta, tb) - 1.0f, 0.0f), (float)nl);
        const float thi = fminf(fmaxf(ta, tb) + 1.0f, (float)(nl - 1));
        j0 = (int)tlo;
        j1 = thi >= tlo ? (int)thi : -1;
    } else if (!(r.x >= 0.0f && r.x < (float)n)) {
        j1 = -1;
    }
    if (!active) { j0 = nl; j1 = -1; }
    int wj0 = wave_min(j0);
    int wj1 = wave_max(j1);
    static_assert(H == 1 || H == 2 || H == 4, "parts of a wave");
    const int len = wj1 - wj0 + 1;
    if (H > 1 && len > 0) {
        const int part = (int)(threadIdx.x & 63) / (64 / H);
        const int a = wj0 + len * part / H, b = wj0 + len * (part + 1) / H - 1;
        wj0 = a;
        wj1 = b;
    }
    const float4* Pq[G];                 // wave-uniform group bases at the start of padded line 0
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        Pq[k] = P + (long)min(k, nq - 1) * qstride;
        acc[k] = vzero<float4>();
    }
    if (__all(!active || fabsf(r.y) <= 0.999f)) {
        if (g.vwide[view]) fp_quad_lines_padded<G, 2, true>(Pq, n, pitch, r, base, wj0, wj1, acc);
        else fp_quad_lines_padded<G, 2, false>(Pq, n, pitch, r, base, wj0, wj1, acc);
    } else if (__all(!active || fabsf(r.y) <= 1.999f)) {
        fp_quad_lines_padded<G, 3, false>(Pq, n, pitch, r, base, wj0, wj1, acc);
    } else {
        for (int j = wj0; j <= wj1; ++j) {
            const float lr = base + (float)j;
            const float ea = fmaf(r.y, lr, r.x), eb = fmaf(r.y, lr + 1.0f, r.x);
            #pragma unroll
            for (int k = 0; k < G; ++k)
                vadd(acc[k], siddon_line_sum<float4>(Pq[k] + (long)j * pitch + kQPadL, n, fminf(ea, eb),
                                                     fmaxf(ea, eb)));
        }
    }
    #pragma unroll
    for (int o = 32; o >= 64 / H; o >>= 1) {
        #pragma unroll
        for (int k = 0; k < G; ++k) {
            acc[k].x += __shfl_xor(acc[k].x, o);
            acc[k].y += __shfl_xor(acc[k].y, o);
            acc[k].z += __shfl_xor(acc[k].z, o);
            acc[k].w += __shfl_xor(acc[k].w, o);
        }
    }
    #pragma unroll
    for (int k = 0; k < G; ++k) acc[k] = vscale(acc[k], r.z);
}

// ---- distance-driven forward projection on the line machinery ----
//
// DDFPFanKernelX/Y (DistanceDriven.cu:64-196): on line l (image row iy for X
// views, column ix for Y views; line coordinate lr = l - nl/2) the footprint
// of bin iu lies between its two edge rays, [e0, e1] with e_k = A_k + B_k lr
// (the FP64 table g.dray), and the ray's value adds the average of the
// line's row-averaged plane over that footprint; the sum is scaled by the
// path length per line L = dy/|cos(a - t)| (X) or dx/|sin(a - t)| (Y), the
// Siddon table's L.  The footprint average is the share arithmetic of the
// Siddon line intervals (overlap / width, the width unclipped), so the same
// branch-free two- or three-cell line code serves both projectors.

// Lines whose footprint can meet [0, n]: the union over the two edges of the
// lines where the edge lies in [-2, n + 2] (one line of margin on each side;
// the other lines add exact zeros while the footprint is narrower than 2).
__device__ __forceinline__ void dd_edge_lines(float A, float B, int n, int nl, float base, int& j0, int& j1)
{
    j0 = 0;
    j1 = nl - 1;
    if (B != 0.0f) {
        const float ib = 1.0f / B;
        const float ta = (-2.0f - A) * ib - base, tb = ((float)n + 2.0f - A) * ib - base;
        const float tlo = fminf(fmaxf(fminf(ta, tb) - 1.0f, 0.0f), (float)nl);
        const float thi = fminf(fmaxf(ta, tb) + 1.0f, (float)(nl - 1));
        j0 = (int)tlo;
        j1 = thi >= tlo ? (int)thi : -1;
    } else if (!(A >= -2.0f && A <= (float)n + 2.0f)) {
        j0 = nl;
        j1 = -1;
    }
}

// Line range of bin ray d = (A0, B0, A1, B1) and the largest footprint width
// over it (0 if the range is empty).
__device__ __forceinline__ void dd_ray_lines(float4 d, int n, int nl, float base, int& j0, int& j1, float& wmax)
{
    int a0, a1, b0, b1;
    dd_edge_lines(d.x, d.y, n, nl, base, a0, a1);
    dd_edge_lines(d.z, d.w, n, nl, base, b0, b1);
    j0 = min(a0, b0);
    j1 = max(a1, b1);
    wmax = 0.0f;
    if (j1 >= j0) {
        const float dA = d.z - d.x, dB = d.w - d.y;
        wmax = fmaxf(fabsf(fmaf(dB, base + (float)j0, dA)), fabsf(fmaf(dB, base + (float)j1, dA)));
    }
}

// Plain planes: lines wj0 .. wj1 of S slices, footprints at most NC cells wide
template <class T, int S, int NC>
__device__ __forceinline__ void dd_narrow_lines(const T* const (&Pq)[S], int n, float4 d, float base, int wj0,
                                                int wj1, T (&acc)[S])
{
    constexpr int kUnroll = S > 4 ? 1 : (S > 2 ? 2 : 4);
    #pragma unroll kUnroll
    for (int j = wj0; j <= wj1; ++j) {
        const float lr = base + (float)j;
        const float ea = fmaf(d.y, lr, d.x), eb = fmaf(d.w, lr, d.z);
        const Line3 c = line3(n, fminf(ea, eb), fmaxf(ea, eb));
        const unsigned lboff = (unsigned)(j * n) * (unsigned)sizeof(T);
        #pragma unroll
        for (int q = 0; q < S; ++q) line_madd_at<T, NC>(acc[q], Pq[q], lboff, c);
    }
}

// One bin of one view through S consecutive slices of the row-averaged planes
// RX [s][ny][nx] (X views) and RY [s][nx][ny] (Y views); slices q >= nsl read
// slice nsl - 1.  Wave-uniform line loop as in siddon_fp_ray_slices.
template <class T, int S>
__device__ __forceinline__ void dd_fp_ray_slices(const FanK& g, const T* __restrict__ RX, const T* __restrict__ RY,
                                                 long plane, int nsl, int view, int iu, bool active, T (&acc)[S])
{
    const float4 d = g.dray[(long)view * g.nu + iu];
    const float L = g.sray[(long)view * g.nu + iu].z;
    const bool xv = g.vcls[view] == 0;
    const T* __restrict__ P = xv ? RX : RY;
    const int nl = xv ? g.ny : g.nx;
    const int n = xv ? g.nx : g.ny;
    const float base = -0.5f * (float)nl;
    int j0, j1;
    float wmax;
    dd_ray_lines(d, n, nl, base, j0, j1, wmax);
    if (!active) { j0 = nl; j1 = -1; wmax = 0.0f; }
    const int wj0 = wave_min(j0);
    const int wj1 = wave_max(j1);
    const T* Pq[S];
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        Pq[q] = P + (long)min(q, nsl - 1) * plane;
        acc[q] = vzero<T>();
    }
    if (__all(wmax <= 0.999f)) {
        dd_narrow_lines<T, S, 2>(Pq, n, d, base, wj0, wj1, acc);
    } else if (__all(wmax <= 1.999f)) {
        dd_narrow_lines<T, S, 3>(Pq, n, d, base, wj0, wj1, acc);
    } else {
        for (int j = wj0; j <= wj1; ++j) {
            const float lr = base + (float)j;
            const float ea = fmaf(d.y, lr, d.x), eb = fmaf(d.w, lr, d.z);
            #pragma unroll
            for (int q = 0; q < S; ++q)
                vadd(acc[q], vscale(seg_sum<T>(Pq[q] + (long)j * n, n, fminf(ea, eb), fmaxf(ea, eb)),
                                    1.0f / fabsf(eb - ea)));
        }
    }
    #pragma unroll
    for (int q = 0; q < S; ++q) acc[q] = vscale(acc[q], L);
}

// Padded quad planes (see fp_quad_lines_padded): G groups, footprints at most
// NC cells wide
template <int G, int NC>
__device__ __forceinline__ void dd_quad_lines_padded(const float4* const (&Pq)[G], int n, int pitch, float4 d,
                                                     float base, int wj0, int wj1, float4 (&acc)[G])
{
    #pragma unroll 2
    for (int j = wj0; j <= wj1; ++j) {
        const float lr = base + (float)j;
        const float ea = fmaf(d.y, lr, d.x), eb = fmaf(d.w, lr, d.z);
        const float lo = fminf(ea, eb), hi = fmaxf(ea, eb);
        const float w = hi - lo;
        const float f0 = floorf(lo);
        const int m0 = (int)__builtin_amdgcn_fmed3f(f0, -(float)kQPadL, (float)(n + kQPadR - NC));
        const bool pos = w > 0.0f;
        const float inv = __builtin_amdgcn_rcpf(pos ? w : 1.0f);
        const float o0 = fminf(hi, f0 + 1.0f) - lo;
        float w0, w1, w2 = 0.0f;
        if (NC == 2) {
            w0 = (pos ? o0 : 1.0f) * inv;
            w1 = (w - o0) * inv;
        } else {
            const float o1 = fmaxf(fminf(hi, f0 + 2.0f) - (f0 + 1.0f), 0.0f);
            w0 = (pos ? o0 : 1.0f) * inv;
            w1 = o1 * inv;
            w2 = (w - o0 - o1) * inv;
        }
        const unsigned boff = (unsigned)(j * pitch + kQPadL + m0) * 16u;
        #pragma unroll
        for (int k = 0; k < G; ++k) {
            const float4 v0 = ld_at(Pq[k], boff), v1 = ld_at(Pq[k], boff + 16u);
            float4& a = acc[k];
            a.x = fmaf(v0.x, w0, a.x); a.y = fmaf(v0.y, w0, a.y); a.z = fmaf(v0.z, w0, a.z); a.w = fmaf(v0.w, w0, a.w);
            a.x = fmaf(v1.x, w1, a.x); a.y = fmaf(v1.y, w1, a.y); a.z = fmaf(v1.z, w1, a.z); a.w = fmaf(v1.w, w1, a.w);
            if (NC == 3) {
                const float4 v2 = ld_at(Pq[k], boff + 32u);
                a.x = fmaf(v2.x, w2, a.x); a.y = fmaf(v2.y, w2, a.y);
                a.z = fmaf(v2.z, w2, a.z); a.w = fmaf(v2.w, w2, a.w);
            }
        }
    }
}

// One bin ray through G interleaved slice groups of the padded quad averaged
// planes QX [q][ny][qpitch(nx)] (X views) and QY [q][nx][qpitch(ny)] (Y views);
// H parts of 64 / H lanes split the wave's line range as in
// siddon_fp_ray_quads.
template <int G, int H = 1>
__device__ __forceinline__ void dd_fp_ray_quads(const FanK& g, const float4* __restrict__ QX,
                                                const float4* __restrict__ QY, int nq, int view, int iu, bool active,
                                                float4 (&acc)[G])
{
    const float4 d = g.dray[(long)view * g.nu + iu];
    const float L = g.sray[(long)view * g.nu + iu].z;
    const bool xv = g.vcls[view] == 0;
    const float4* __restrict__ P = xv ? QX : QY;
    const int nl = xv ? g.ny : g.nx;
    const int n = xv ? g.nx : g.ny;
    const int pitch = qpitch(n);
    const long qstride = (long)nl * pitch;
    const float base = -0.5f * (float)nl;
    int j0, j1;
    float wmax;
    dd_ray_lines(d, n, nl, base, j0, j1, wmax);
    if (!active) { j0 = nl; j1 = -1; wmax = 0.0f; }
    int wj0 = wave_min(j0);
    int wj1 = wave_max(j1);
    static_assert(H == 1 || H == 2 || H == 4, "parts of a wave");
    const int len = wj1 - wj0 + 1;
    if (H > 1 && len > 0) {
        const int part = (int)(threadIdx.x & 63) / (64 / H);
        const int a = wj0 + len * part / H, b = wj0 + len * (part + 1) / H - 1;
        wj0 = a;
        wj1 = b;
    }
    const float4* Pq[G];
    #pragma unroll
    for (int k = 0; k < G; ++k) {
        Pq[k] = P + (long)min(k, nq - 1) * qstride;
        acc[k] = vzero<float4>();
    }
    if (__all(wmax <= 0.999f)) {
        dd_quad_lines_padded<G, 2>(Pq, n, pitch, d, base, wj0, wj1, acc);
    } else if (__all(wmax <= 1.999f)) {
        dd_quad_lines_padded<G, 3>(Pq, n, pitch, d, base, wj0, wj1, acc);
    } else {
        for (int j = wj0; j <= wj1; ++j) {
            const float lr = base + (float)j;
            const float ea = fmaf(d.y, lr, d.x), eb = fmaf(d.w, lr, d.z);
            #pragma unroll
            for (int k = 0; k < G; ++k)
                vadd(acc[k], vscale(seg_sum<float4>(Pq[k] + (long)j * pitch + kQPadL, n, fminf(ea, eb),
                                                    fmaxf(ea, eb)), 1.0f / fabsf(eb - ea)));
        }
    }
    #pragma unroll
    for (int o = 32; o >= 64 / H; o >>= 1) {
        #pragma unroll
        for (int k = 0; k < G; ++k) {
            acc[k].x += __shfl_xor(acc[k].x, o);
            acc[k].y += __shfl_xor(acc[k].y, o);
            acc[k].z += __shfl_xor(acc[k].z, o);
            acc[k].w += __shfl_xor(acc[k].w, o);
        }
    }
    #pragma unroll
    for (int k = 0; k < G; ++k) acc[k] = vscale(acc[k], L);
}

// One ray of one slice (siddon_fp_ray_slices with S = 1).
template <class T>
__device__ __forceinline__ T siddon_fp_ray(const FanK& g, const T* __restrict__ Pyx, const T* __restrict__ Pxy,
                                           int view, int iu, bool active = true)
{
    T acc[1];
    siddon_fp_ray_slices<T, 1>(g, Pyx, Pxy, 0, 1, view, iu, active, acc);
    return acc[0];
}

// atan from an odd polynomial on [-1, 1] (coefficients fitted here by
// iteratively reweighted least squares on Chebyshev nodes; maximum error
// 2.5e-7 rad) and, unless UNIT (|z| <= 1 known), atan(z) = sign(z) pi/2 -
// atan(1/z) outside.  Used only to locate the bins a pixel can meet;
// RN_ATAN_FAST_ERR bounds its error with FP32 evaluation included.
#define RN_ATAN_FAST_ERR 1.0e-6f
template <bool UNIT>
__device__ __forceinline__ float atan_fast(float z)
{
    const float az = fabsf(z);
    const bool inv = !UNIT && az > 1.0f;
    const float t = inv ? __builtin_amdgcn_rcpf(az) : az;
    const float s = t * t;
    float p = 0.006811782025349205f;
    p = fmaf(p, s, -0.033604191884230736f);
    p = fmaf(p, s, 0.07962364933877517f);
    p = fmaf(p, s, -0.13233341755196557f);
    p = fmaf(p, s, 0.1980781590519013f);
    p = fmaf(p, s, -0.3331736818605284f);
    p = fmaf(p, s, 0.9999961116602261f);
    float r = p * t;
    if (inv) r = 1.5707963267948966f - r;
    return copysignf(r, z);
}

// First and last bin whose centre ray can meet pixel (ix, iy) of view `view`:
// the pixel lies in the disc of radius rpix around its centre, whose rays span
// the detector angles atan(rx / d0) +- asin(s), s = rpix / dist, and
// asin(s) <= s (1 + 2 s) for 0 <= s <= 1.  UNIT: |rx / d0| <= 1 for every pixel
// and view of the geometry (FanK::unit_angle).
template <bool UNIT>
__device__ __forceinline__ void bin_range(const FanK& g, float xc, float yc, float rpix, float inv_da, int view,
                                          int& k0, int& k1)
{
    const float ct = g.vcos[view], st = g.vsin[view];
    const float rx = xc * ct + yc * st, ry = -xc * st + yc * ct;
    const float d0 = ry + g.dso;
    const float uc = g.cu - g.off_a - atan_fast<UNIT>(rx * __builtin_amdgcn_rcpf(d0)) * inv_da;
    const float sr = fminf(rpix * __builtin_amdgcn_rsqf(fmaf(rx, rx, d0 * d0)), 1.0f);
    const float hw = sr * fmaf(2.0f, sr, 1.0f) * fabsf(inv_da) + 0.02f + RN_ATAN_FAST_ERR * fabsf(inv_da);
    k0 = max((int)ceilf(uc - hw), 0);
    k1 = min((int)floorf(uc + hw), g.nu - 1);
}

// Length of ray r in cell `cell` of line lr (the weight siddon_fp_ray gives
// it), computed without branches.  WIDE: the interval's FP32 width is known to
// be positive (no zero-width case to select; the same value as the general
// form then).
template <bool WIDE>
__device__ __forceinline__ float siddon_cell_weight(const float4 r, float lr, float cell)
{
    const float ea = fmaf(r.y, lr, r.x), eb = fmaf(r.y, lr + 1.0f, r.x);
    const float lo = fminf(ea, eb), hi = fmaxf(ea, eb), w = hi - lo;
    const float ov = fmaxf(fminf(hi, cell + 1.0f) - fmaxf(lo, cell), 0.0f);
    if (WIDE) return ov * __builtin_amdgcn_rcpf(w) * r.z;
    const bool pos = w > 0.0f;
    const float pt = (lo >= cell && lo < cell + 1.0f) ? 1.0f : 0.0f;
    const float inv = __builtin_amdgcn_rcpf(pos ? w : 1.0f);
    return (pos ? ov * inv : pt) * r.z;
}

// the bins k0 .. k1 of one view for S slices (slice q's row `rboff` bytes
// after the wave-uniform base pq[q]; the view's ray table row `rowoff` bytes
// after g.sray): the first NB unrolled (each under its own lane mask, so an
// empty range k1 < k0 adds nothing), the rest in a loop; each weight is
// computed once and applied to every slice.
template <class T, int NB, int S>
__device__ __forceinline__ void siddon_gather_bins(const float4* __restrict__ sray, unsigned rowoff,
                                                   const T* const (&pq)[S], unsigned rboff, int nu, int k0, int k1,
                                                   float lr, float cell, T (&acc)[S])
{
    #pragma unroll
    for (int j = 0; j < NB; ++j) {
        // lanes past k1 take no part (their loads are masked off, so they
        // cost no vector-memory bandwidth); the others add weight x value
        if (k0 + j <= k1) {
            const unsigned kk = (unsigned)(k0 + j);
            const float wt = siddon_cell_weight<false>(ld_at(sray, rowoff + kk * 16u), lr, cell);
            const unsigned boff = rboff + kk * (unsigned)sizeof(T);
            #pragma unroll
            for (int q = 0; q < S; ++q) vmadd(acc[q], ld_at(pq[q], boff), wt);
        }
    }
    for (int k = k0 + NB; k <= k1; ++k) {
        const float wt = siddon_cell_weight<false>(ld_at(sray, rowoff + (unsigned)k * 16u), lr, cell);
        const unsigned boff = rboff + (unsigned)k * (unsigned)sizeof(T);
        #pragma unroll
        for (int q = 0; q < S; ++q) vmadd(acc[q], ld_at(pq[q], boff), wt);
    }
}

// siddon_gather_bins for G slice groups from a wave's staged window in LDS:
// row q of `st` holds group q's values and row G the view's ray table, CAP
// bins per row from the window's first bin (every slot holds written values:
// the caller clears the rows before its first window); t0 .. t1 are the pixel's bins
// relative to that first bin.  The first NB bins are taken by every lane with
// no branch around the accumulators: a lane past t1 reads slot 0 (a bin of
// this view's window) with weight 0, which keeps the kernel in registers.
// The rest in a loop.
template <int NB, int G, int CAP, bool WIDE>
__device__ __forceinline__ void siddon_gather_staged(const float4* st, int t0, int t1, float lr, float cell,
                                                     float4 (&acc)[G])
{
    #pragma unroll
    for (int j = 0; j < NB; ++j) {
        const bool on = t0 + j <= t1;
        const int t = on ? t0 + j : 0;
        const float wt = on ? siddon_cell_weight<WIDE>(st[G * CAP + t], lr, cell) : 0.0f;
        #pragma unroll
        for (int q = 0; q < G; ++q) vmadd(acc[q], st[q * CAP + t], wt);
    }
    for (int t = t0 + NB; t <= t1; ++t) {
        const float wt = siddon_cell_weight<WIDE>(st[G * CAP + t], lr, cell);
        #pragma unroll
        for (int q = 0; q < G; ++q) vmadd(acc[q], st[q * CAP + t], wt);
    }
}

// Exact transpose of siddon_fp_ray: pixel (ix, iy) gathers, over views
// view0 .. view0+nviews-1 of prj ([nviews][nu] per slice, slices sstride
// elements apart; slices q >= nsl read slice nsl - 1 and are discarded by the
// caller), the rays that cross it, with the weights siddon_fp_ray gives it.
// The bin range and the weights are computed once per view for all S slices.
// The first NB bins of a pixel's range are taken without branches, the rest
// in a loop (NB changes only the speed).
template <class T, int NB, bool UNIT, int S>
__device__ __forceinline__ void siddon_bp_views(const FanK& g, const T* __restrict__ prj, long sstride, int nsl,
                                                int view0, int nviews, int ix, int iy, T (&acc)[S])
{
    const float xc = (ix - 0.5f * (g.nx - 1)) * g.dx, yc = (iy - 0.5f * (g.ny - 1)) * g.dy;
    const float rpix = 0.5f * sqrtf(g.dx * g.dx + g.dy * g.dy);
    const float inv_da = 1.0f / g.da;
    const float yr = (float)iy - 0.5f * g.ny;
    const float xr = (float)ix - 0.5f * g.nx;
    const T* pq[S];                             // wave-uniform slice bases
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        pq[q] = prj + (long)min(q, nsl - 1) * sstride;
        acc[q] = vzero<T>();
    }
    for (int vi = 0; vi < nviews; ++vi) {
        const int view = view0 + vi;
        int k0, k1;
        bin_range<UNIT>(g, xc, yc, rpix, inv_da, view, k0, k1);
        const bool xv = g.vcls[view] == 0;
        const float lr = xv ? yr : xr;          // this pixel's line
        const float cell = xv ? (float)ix : (float)iy;
        siddon_gather_bins<T, NB, S>(g.sray, (unsigned)(view * g.nu) * 16u, pq,
                                     (unsigned)(vi * g.nu) * (unsigned)sizeof(T), g.nu, k0, k1, lr, cell, acc);
    }
}

template <class T, int NB, int S>
__device__ __forceinline__ void siddon_bp_pixel_slices(const FanK& g, const T* __restrict__ prj, long sstride,
                                                       int nsl, int view0, int nviews, int ix, int iy,
                                                       T (&acc)[S])
{
    if (g.unit_angle) siddon_bp_views<T, NB, true, S>(g, prj, sstride, nsl, view0, nviews, ix, iy, acc);
    else siddon_bp_views<T, NB, false, S>(g, prj, sstride, nsl, view0, nviews, ix, iy, acc);
}

// one pixel of one slice
template <class T, int NB = 4>
__device__ __forceinline__ T siddon_bp_pixel(const FanK& g, const T* __restrict__ prj, int view0, int nviews,
                                             int ix, int iy)
{
    T acc[1];
    siddon_bp_pixel_slices<T, NB, 1>(g, prj, 0, 1, view0, nviews, ix, iy, acc);
    return acc[0];
}

// Exact transpose of the DD FP (dd_fp_ray_slices; no reference counterpart:
// the reference DD BP is not the adjoint of its FP).  The pixel enters line
// iy (or ix), its lower edge, and line iy + 1 (or ix + 1), its upper edge,
// with weight 1/2 each (each line averages two rows); on each it takes, from
// every bin whose footprint meets its segment, the FP's weight
// L_k overlap / width from the same edge-ray table.  One pass over the
// candidate bins serves both lines: the bins whose centre lies within half a
// bin of the angles the pixel's disc spans (bin_range plus half a bin; the
// others have zero overlap).  S slices share the weights (slice q's
// view-major sinogram `sstride` elements after prj; q >= nsl re-reads the last).
template <bool UNIT, int S>
__device__ __forceinline__ void dd_adj_pixel_slices(const FanK& g, const float* __restrict__ prj, long sstride,
                                                    int nsl, int view0, int nviews, int ix, int iy,
                                                    float (&acc)[S])
{
    const float inv_da = 1.0f / g.da;
    const float ainv = fabsf(inv_da);
    #pragma unroll
    for (int q = 0; q < S; ++q) acc[q] = 0.0f;
    const float* pq[S];
    #pragma unroll
    for (int q = 0; q < S; ++q) pq[q] = prj + (long)min(q, nsl - 1) * sstride;
    const float xc = ((float)ix - 0.5f * (g.nx - 1)) * g.dx, yc = ((float)iy - 0.5f * (g.ny - 1)) * g.dy;
    const float rpix = 0.5f * sqrtf(g.dx * g.dx + g.dy * g.dy);
    for (int vi = 0; vi < nviews; ++vi) {
        const int view = view0 + vi;
        const float ct = g.vcos[view], st = g.vsin[view];
        const float4* __restrict__ dr = g.dray + (long)view * g.nu;
        const float4* __restrict__ sr = g.sray + (long)view * g.nu;
        const long row = (long)vi * g.nu;
        const bool isx = g.vcls[view] == 0;
        const int l0 = isx ? iy : ix;
        const int nl = isx ? g.ny : g.nx;
        const float fcell = (float)(isx ? ix : iy);
        const float up = l0 + 1 < nl ? 0.5f : 0.0f;     // the upper edge's line exists
        const float rx = xc * ct + yc * st, ry = -xc * st + yc * ct;
        const float d0 = ry + g.dso;
        const float uc = g.cu - g.off_a - atan_fast<UNIT>(rx * __builtin_amdgcn_rcpf(d0)) * inv_da;
        const float sn = fminf(rpix * __builtin_amdgcn_rsqf(fmaf(rx, rx, d0 * d0)), 1.0f);
        const float hw = sn * fmaf(2.0f, sn, 1.0f) * ainv + 0.52f + RN_ATAN_FAST_ERR * ainv;
        const int k0 = max((int)ceilf(uc - hw), 0), k1 = min((int)floorf(uc + hw), g.nu - 1);
        const float lr = (float)l0 - 0.5f * (float)nl;
        for (int k = k0; k <= k1; ++k) {
            const float4 d = dr[k];
            float sh[2];
            #pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float le = lr + (float)e;
                const float ea = fmaf(d.y, le, d.x), eb = fmaf(d.w, le, d.z);
                const float lo = fminf(ea, eb), hi = fmaxf(ea, eb), w = hi - lo;
                const float ov = fmaxf(fminf(hi, fcell + 1.0f) - fmaxf(lo, fcell), 0.0f);
                const bool pos = w > 0.0f;
                sh[e] = pos ? ov * __builtin_amdgcn_rcpf(w) : ((lo >= fcell && lo < fcell + 1.0f) ? 1.0f : 0.0f);
            }
            const float wk = fmaf(up, sh[1], 0.5f * sh[0]) * sr[k].z;
            #pragma unroll
            for (int q = 0; q < S; ++q) acc[q] = fmaf(pq[q][row + k], wk, acc[q]);
        }
    }
}

} // namespace fan
} // namespace rn

namespace rn {

// The OS-SQS(-TV) update of one pixel of one slice (ossqs.cu:124-135 with the
// TV terms of TotalVariation.cu:9-68): x - (step g + beta2 gTV) / (D + beta2 nTV + 1e-6),
// cropped to [lb, ub] when crop is set.  dx0, dx1, dy0, dy1 are the differences
// to the 4-neighbours (the pixel value itself at the image border); gTV =
// sum d / psi / 2 and nTV = sum 1 / psi with psi = sqrt(d^2 + eps), here with
// reciprocal square roots (about 1 ulp each), and the quotient as a product
// with a 1-ulp reciprocal.  g: the backprojected weighted residual, d: the
// normaliser.
__device__ __forceinline__ float sqs_update_core(float x, float dx0, float dx1, float dy0, float dy1, float g,
                                                 float d, float step, float beta2, float eps, float lb, float ub,
                                                 int crop)
{
    float num = step * g;
    float den = d;
    if (beta2 != 0.0f) {
        const float r0 = __builtin_amdgcn_rsqf(fmaf(dx0, dx0, eps)), r1 = __builtin_amdgcn_rsqf(fmaf(dx1, dx1, eps));
        const float r2 = __builtin_amdgcn_rsqf(fmaf(dy0, dy0, eps)), r3 = __builtin_amdgcn_rsqf(fmaf(dy1, dy1, eps));
        const float gt = 0.5f * (dx0 * r0 + dx1 * r1 + dy0 * r2 + dy1 * r3);
        const float nt = r0 + r1 + r2 + r3;
        num = fmaf(beta2, gt, num);
        den = fmaf(beta2, nt, den);
    }
    float xn = x - num * __builtin_amdgcn_rcpf(den + 1e-6f);     // 1-ulp reciprocal
    if (crop) xn = xn < lb ? lb : (xn > ub ? ub : xn);
    return xn;
}

// sqs_update_core for pixel (ix, iy) of one slice Xs [nx][ny], pix = ix * ny + iy
__device__ __forceinline__ float sqs_pixel_update(const float* __restrict__ Xs, long pix, int ix, int iy, int nx,
                                                  int ny, float g, float d, float step, float beta2, float eps,
                                                  float lb, float ub, int crop)
{
    const float x = Xs[pix];
    float dx0 = 0.0f, dx1 = 0.0f, dy0 = 0.0f, dy1 = 0.0f;
    if (beta2 != 0.0f) {
        dx0 = ix == 0 ? x : x - Xs[pix - ny];
        dx1 = ix == nx - 1 ? x : x - Xs[pix + ny];
        dy0 = iy == 0 ? x : x - Xs[pix - 1];
        dy1 = iy == ny - 1 ? x : x - Xs[pix + 1];
    }
    return sqs_update_core(x, dx0, dx1, dy0, dy1, g, d, step, beta2, eps, lb, ub, crop);
}

} // namespace rn
