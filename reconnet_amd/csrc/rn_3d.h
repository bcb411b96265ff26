// Arbitrary-geometry (parallel / cone) Siddon, DD tomosynthesis, cone OS-SQS
// and helical BP on reference-layout device buffers.
#pragma once

#include "rn_common.h"

namespace rn {

struct ArbGeo {
    int nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
    int nu, nview, nv;
    float du, dv, off_u, off_v;
    // every view has its detector rows along z and its columns in a
    // horizontal plane (detV = (0, 0, v), detU.z = 0; arb_rows_along_z):
    // the cone backprojection then runs a kernel without the general-view
    // path.  Set by callers that hold the views on the host; 0 is always safe.
    int rows_z = 0;
};

bool arb_rows_along_z(const float* detU, const float* detV, int nview);

// Siddon with per-view detCenter / detU / detV and either invRayDir
// (parallel, SiddonParallel.cu) or the source position (cone, SiddonCone.cu).
// All float3 arrays are device [nview][3].
void siddon_arb_fp(const ArbGeo& g, bool parallel, int nB, int nC, const float* img, float* prj,
                   const float* detCenter, const float* detU, const float* detV, const float* srcOrDir,
                   DevBuf& work, hipStream_t st);
void siddon_arb_bp(const ArbGeo& g, bool parallel, int nB, int nC, const float* prj, float* img,
                   const float* detCenter, const float* detU, const float* detV, const float* srcOrDir,
                   DevBuf& work, hipStream_t st);

// DistanceDriven3D.cu tomosynthesis projector pair (detector axes = x, y)
void dd_tomo_fp(const ArbGeo& g, int nB, int nC, const float* img, float* prj,
                const float* detCenter, const float* src, DevBuf& work, hipStream_t st);
void dd_tomo_bp(const ArbGeo& g, int nB, int nC, const float* prj, float* img,
                const float* detCenter, const float* src, DevBuf& work, hipStream_t st);

// cOSSQSConeAbitrary (ossqs.cu:163-333); x is updated in place
void os_sqs_cone(const ArbGeo& g, int nB, int nC, float* x, const float* prj, const float* norm, const float* mask,
                 const float* detCenter, const float* detU, const float* detV, const float* src,
                 float lb, float ub, int nviewPerSubset, int nSubsets, float lambda, DevBuf& work, hipStream_t st);

// Circular cone-beam plan (no reference counterpart): per-view detector
// centre, axes and source on the device plus the host's float32 choice of each
// view's main axis (x iff |C.x - S.x| >= |C.y - S.y|).  The DD pair needs
// detector columns parallel to z (detV = (0, 0, v)) and rows normal to it
// (detU.z = 0), the geometry of circular CT.
class ConePlan {
public:
    void setup(const ArbGeo& g, const float* C, const float* U, const float* V, const float* S, hipStream_t st);
    // reference layouts: img [B][nx][ny][nz][C], prj [B][nu][nview][nv][C] (device)
    void dd_fp(int nB, int nC, const float* img, float* prj, hipStream_t st) const;
    void dd_adj(int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st) const;
    void siddon_fp(int nB, int nC, const float* img, float* prj, DevBuf& work, hipStream_t st) const;
    void siddon_bp(int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st) const;
    ArbGeo geo{};
private:
    // vec_: C, U, V, S per view; cls_: main axis per view; views_, cols_:
    // per-view and per-(view, column) constants of the table-driven adjoint
    DevBuf vec_, cls_, views_, cols_;
};

struct HelicalGeo {
    int nx, ny, nz;
    float dx, dy, dz;
    int nu, nview, nv;
    float du, dv, off_u, off_v, dsd, dso;
    int nviewPerPI;
    float theta0, volZ0, zrot;
    int mPI;
    float Q;
};

// bpHelicalParallelConebeam (Recon3D/HelicalBP.cu:10-83)
void helical_bp(const HelicalGeo& h, int nB, int nC, const float* prj, float* img, DevBuf& work, hipStream_t st);

// Helical fan -> parallel-cone rebinning (no reference counterpart: the
// helical BP consumes rebinned data).  Raw views: source angle
// b_m = beta0 + m dbeta at (dso cos b, dso sin b, z(b)), a flat detector at
// distance dsd, column s = (iu - off_u - (nu-1)/2) du along (sin b, -cos b, 0),
// row w = (iv - off_v - (nv-1)/2) dv along z (the arguments of
// cSiddonConeProjectionAbitrary).  Parallel views: angle
// t_k = theta0 + k dtheta, rays of direction -(cos t, sin t) at signed
// distance u = (l - (nu-1)/2 - off_u) du from the axis along (sin t, -cos t),
// rows v = (iv - (nv-1)/2 - off_v) dv on the cylinder of radius dsd around
// the source (the coordinates of bpHelicalParallelConebeam).  The ray (t, u)
// comes from the source at b = t - asin(u / dso) through the flat-detector
// point s = dsd tan(asin(u / dso)), w = v / cos(asin(u / dso)).
struct RebinGeo {
    int nu_f, nview_f, nv_f;
    float du_f, dv_f, off_u_f, off_v_f;
    int nu_p, nview_p, nv_p;
    float du_p, dv_p, off_u_p, off_v_p;
    float dso, dsd;
    // fractional raw view of parallel view k with gamma = 0: k * kscale + moff;
    // minus gamma * idbeta (kscale = dtheta / dbeta, moff = (theta0 - beta0) / dbeta)
    double kscale, moff, idbeta;
};
// fan [B][nu_f][nview_f][nv_f][C] -> par [B][nu_p][nview_p][nv_p][C] (device):
// trilinear interpolation in (column, view, row), zero outside the raw data
void helical_rebin(const RebinGeo& r, int nB, int nC, const float* fan, float* par, hipStream_t st);

} // namespace rn
