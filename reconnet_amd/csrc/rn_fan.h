// Fan-beam geometry, device trig tables and launchers for the 2D fan kernels.
//
// Conventions (ReconNet4D_no_tf/DistanceDriven.cu:14-20,81-99,514-518;
// Siddon.cu:253-263; Projector.h:5):
//   image      img[B][nx][ny][nz][C], pixel (ix,iy) centre at ((ix-(nx-1)/2)dx, (iy-(ny-1)/2)dy),
//              image coordinates X = x/dx + nx/2 (origin at the corner of pixel 0);
//   sinogram   prj[B][nu][nview][nv][C];
//   view k     source at (dso sin t, -dso cos t), t = deg[k];
//   bin iu     equiangular arc, detector angle a = (-(iu-(nu-1)/2) - off_a) da,
//              edges at a +- da/2; a ray of angle a has unit direction
//              (sin(a-t), cos(a-t)) and satisfies  X*uy - Y*ux = dso sin a  (physical).
//   X views    |cosf t| > |sinf t| in float32 (DistanceDriven.cu:302): lines are image rows;
//   Y views    the others: lines are image columns.
//
// Internal device layouts used by the kernels (one 2D "slice" per (b, z-or-v, c)):
//   plane xy : [slice][nx][ny]     (the reference layout when nz = nc = 1)
//   plane yx : [slice][ny][nx]     (transposed copy, rows contiguous)
//   sino  vu : [slice][nview][nu]  (view-major, bins contiguous)
#pragma once

#include "rn_common.h"

#include <vector>

namespace rn {

struct FanGeo {
    int nx, ny, nz;
    float dx, dy, dz;
    int nu, nview, nv;
    float da, dv, off_a, off_v, dsd, dso;
};

// Kernel-side view of a geometry: scalars plus device tables.
struct FanK {
    int nx, ny, nu, nview;
    float dx, dy, da, off_a, dso, dsd;
    float cu;        // (nu-1)/2
    const float* vcos;   // [nview] cos t
    const float* vsin;   // [nview] sin t
    const int* vcls;     // [nview] 0 = X view, 1 = Y view
    const float* esin;   // [nu+1] sin of bin-edge angles (edge e = a_e + da/2 of bin e)
    const float* ecos;
    const float* bsin;   // [nu] sin of bin-centre angles
    const float* bcos;
    // [nview][nu] line of the ray through bin centre k, in image coordinates:
    // X views X(Y) = A + B (Y - ny/2), Y views Y(X) = A + B (X - nx/2);
    // (A, B, L, 0) with L the path length per image line (dy / |cos(a - t)| or
    // dx / |sin(a - t)|).  Computed in FP64.
    const float4* sray;
    // [nview][nu] the two edge rays of bin k in the same line form,
    // (A_lo, B_lo, A_hi, B_hi) for the edges a -+ da/2 (the DD footprint on a
    // line lies between them).  Computed in FP64.
    const float4* dray;
    const int* vwide;    // [nview] 1: every Siddon line interval of the view has a positive FP32 width
    int unit_angle;      // 1: |rx / (ry + dso)| < 1 for every pixel centre and view (image well inside the circle)
};

// Device tables for one (geometry, angle set).  Angles are kept on the host
// too: the X/Y split is a float32 host decision.
class FanTables {
public:
    void build(const FanGeo& g, const float* deg_host, hipStream_t st);
    bool matches(const FanGeo& g, const float* deg_host) const;
    FanK k{};
    FanGeo geo{};
    std::vector<float> deg;
    std::vector<int> cls;
private:
    DevBuf buf_, ray_, dray_;
    std::vector<float> host_;
};

// ---- 2D kernels on internal layouts (all pointers are device pointers) ----

// line-driven DD FP: RX = row-averaged planes (yx), RY = column-averaged planes (xy)
void launch_dd_fan_fp(const FanK& g, const float* RX, const float* RY, float* prj_vu,
                      int nslices, int view0, int nviews, hipStream_t st);
// DD FP on padded quads of the averaged planes: QX [q][ny][qpitch(nx)],
// QY [q][nx][qpitch(ny)]; writes the plain view-major sinograms [s][nviews][nu]
void launch_dd_fan_fp_q(const FanK& g, const float4* QX, const float4* QY, float* prj_vu, int nslices, int view0,
                        int nviews, hipStream_t st);
// line-driven Siddon FP: Pyx, Pxy = the image planes in both layouts
void launch_siddon_fan_fp(const FanK& g, const float* Pyx, const float* Pxy, float* prj_vu,
                          int nslices, int view0, int nviews, hipStream_t st);
// Siddon FP whose epilogue writes w * (Ax - meas) (meas, w: full vu sinograms)
void launch_siddon_fan_fp_residual(const FanK& g, const float* Pyx, const float* Pxy, float* res_vu,
                                   int nslices, int view0, int nviews, const float* meas_vu, const float* w_vu,
                                   hipStream_t st);
// FP64 variants (used only for the SQS normaliser A'(w A 1))
void launch_siddon_fan_fp_f64(const FanK& g, const double* Pyx, const double* Pxy, double* prj_vu,
                              int nslices, int view0, int nviews, hipStream_t st);
void launch_siddon_fan_bp_f64(const FanK& g, const double* prj_vu, double* img_xy, int nslices,
                              int view0, int nviews, hipStream_t st);
// pixel-driven gathers (exact transposes of the two FPs above)
void launch_siddon_fan_bp(const FanK& g, const float* prj_vu, float* img_xy, int nslices,
                          int view0, int nviews, bool accumulate, hipStream_t st);
void launch_dd_fan_adj(const FanK& g, const float* prj_vu, float* img_xy, int nslices,
                       int view0, int nviews, bool accumulate, hipStream_t st);
// reference-parity DD BP (DDBPFanKernelX/Y) on a pre-weighted vu sinogram
void launch_dd_fan_bp(const FanK& g, const float* wprj_vu, float* img_xy, int nslices, bool fbp,
                      hipStream_t st);
void launch_dd_preweight(const FanK& g, const float* prj_vu, float* wprj_vu, int nslices, hipStream_t st);
// pixel-driven FBP BP (bpFanKernel3D)
void launch_pixel_fan_bp(const FanK& g, const float* prj_vu, float* img_xy, int nslices, hipStream_t st);

// ---- interleaved ("quad") layout: groups of 4 slices stored as one float4 per
// cell (slice 4 q + c in component c; slices past nslices are zero) ----
// Images in this layout have padded lines: kQPadL zero cells before and
// kQPadR after the n cells of every line (pitch qpitch(n)), so that the
// forward projector reads 2 or 3 cells around any crossing without clamping
// its cell interval.
constexpr int kQPadL = 3, kQPadR = 3;
__host__ __device__ constexpr int qpitch(int n) { return n + kQPadL + kQPadR; }
// Siddon FP + weighted residual on padded quad images Qyx [q][ny][qpitch(nx)],
// Qxy [q][nx][qpitch(ny)];
// meas, w: quad sinograms [q][view][nu]; view view0 + i goes to row res_voff + i
// of res [q][res_vstride][nu]
void launch_siddon_fan_fp_residual_q(const FanK& g, const float4* Qyx, const float4* Qxy, float4* res_q,
                                     int nslices, int view0, int nviews, const float4* meas_q, const float4* w_q,
                                     int res_vstride, int res_voff, hipStream_t st);
// out[b][j][:] = in[b][perm[j]][:] for a batch of (rows x cols) matrices (perm: device, rows entries)
void launch_permute_rows(const float* in, float* out, const int* perm, int rows, int cols, int batch,
                         hipStream_t st);
// Siddon BP gather of a quad subset sinogram [q][view][nu] into plain images [s][nx][ny]
void launch_siddon_fan_bp_q(const FanK& g, const float4* prj_q, float* img_xy, int nslices, int view0,
                            int nviews, hipStream_t st);
// The OS-SQS update applied to a backprojected subset gradient: the current
// image is read from its interleaved copy Qold (padded quads [q][nx][qpitch(ny)])
// and the normaliser from D4 (quads [q][nx][ny]); the new image goes to Xn
// (plain [s][nx][ny]) and to its interleaved copies Qxy, Qyx (padded quads,
// slices past the batch 0), which must not alias Qold
struct SqsUpd {
    const float4* Qold;
    const float4* D4;
    float* Xn;
    float4* Qxy;
    float4* Qyx;
    float step, beta2, eps, lb, ub;
    int crop;
};
// launch_siddon_fan_bp_q with the update fused into the backprojection's epilogue
// when the chosen launch shape is one of the LDS-staged ones (returns true; img_xy
// is then not written), else the plain backprojection into img_xy (returns false:
// the caller applies the update)
bool launch_siddon_fan_bp_q_update(const FanK& g, const float4* prj_q, float* img_xy, int nslices, int view0,
                                   int nviews, const SqsUpd& u, hipStream_t st);
// reference-parity DD BP of a quad pre-weighted sinogram [q][view][nu] into plain images
void launch_dd_fan_bp_q(const FanK& g, const float4* wprj_q, float* img_xy, int nslices, bool fbp,
                        hipStream_t st);
// plain [s][n1][n2] -> quads [q][n1][n2] (if q_same) and (if q_tr) the transposed quads [q][n2][n1];
// padded: lines of the padded layout (pitch qpitch(n2), resp. qpitch(n1); the
// pad cells are left as they are)
void launch_pack_quads(const float* in, float4* q_same, float4* q_tr, int n1, int n2, int nslices, bool padded,
                       hipStream_t st);

// ---- layout helpers ----
// out[b][c][r] = in[b][r][c] for a batch of (rows x cols) matrices
void launch_transpose(const float* in, float* out, int rows, int cols, int batch, hipStream_t st);
// axis 0: out[i][j] = (in[i-1][j] + in[i][j]) / 2 ; axis 1: out[i][j] = (in[i][j-1] + in[i][j]) / 2
void launch_avg_prev(const float* in, float* out, int n0, int n1, int batch, int axis, hipStream_t st);
// dst[(b*nout + o)*nc + c][i] = sum_p M[o][p] * src[((b*n + i)*np + p)*nc + c]
void launch_mix_gather(const float* src, float* dst, const float* M, int nB, int n, int np, int nc, int nout,
                       hipStream_t st);
// dst[((b*n + i)*nout + o)*nc + c] (+)= sum_p M[o][p] * src[(b*np + p)*nc + c][i]
void launch_mix_scatter(const float* src, float* dst, const float* M, int nB, int n, int np, int nc, int nout,
                        bool accumulate, hipStream_t st);
void launch_fill(float* p, float v, size_t n, hipStream_t st);
void launch_axpy(float* y, const float* x, float a, size_t n, hipStream_t st);

} // namespace rn
