// OS-SQS data step (ossqs.cu), TV surrogate (TotalVariation.cu) and the
// combined SQS-TV iteration, device resident.
#pragma once

#include "rn_fan.h"

#include <vector>

namespace rn {

// TVSQSKernel over nslices 2D images [s][nx][ny]: grad, norm (FP32) and the
// per-slice sums of var reduced in FP64 (host vector, length nslices).
void tv_sqs(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
            std::vector<double>* tv, DevBuf& scratch, hipStream_t st);

// One cOSSQS pass (ossqs.cu:12-161) on reference layouts for any nB, nC, nz,
// nv (device pointers): per subset of views [start, start + cur), the Siddon
// FP of x, r = w (Ax - b) against that block of the full sinogram, the Siddon
// BP of r and x <- x - lambda nSubsets A'r / (D + 1e-6), cropped to [lb, ub]
// when lb < ub.  x is updated in place; a short last subset has
// nview - start views.
void os_sqs_ref_layout(const FanGeo& g, const float* deg_host, int nB, int nC, int nviewPerSubset, int nSubsets,
                       float* x, const float* prj, const float* w, const float* D, float lambda, float lb, float ub,
                       hipStream_t st);

// tv_sqs without any allocation: the FP64 partials go to a file-scope device
// array of kTvScratchDoubles doubles (2 MiB per device, held until the process
// ends); the per-slice costs go to tv_dev (device, nslices doubles; may be
// null).  Needs tv_sqs_scratch_doubles(nx, ny, nslices) <= kTvScratchDoubles,
// else it throws before launching anything.  Calls that overlap on one device
// share the array and race.
constexpr long kTvScratchDoubles = 1L << 18;
long tv_sqs_scratch_doubles(int nx, int ny, int nslices);
void tv_sqs_noalloc(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
                    double* tv_dev, hipStream_t st);

struct SqsStep {
    float lambda = 1.0f;    // step scale of the data term (ossqs.cu:139)
    float beta = 0.0f;      // TV weight (0 = data term only, as cOSSQS)
    float eps = 1e-6f;      // TV smoothing delta^2 (TotalVariation.cu:59)
    float lb = -3.4028235e38f, ub = 3.4028235e38f;
};

// Device-resident OS-SQS state for a batch of 2D fan slices sharing one
// geometry.  Images are [s][nx][ny]; sinograms are held view-major [s][nview][nu].
class OsSqs {
public:
    void setup(const FanGeo& g, const float* deg_host, int nslices, int nviewPerSubset, int nSubsets,
               hipStream_t st);
    // measured sinogram and weights in reference layout [s][nu][nview] (device); w may be null (= 1)
    void set_data(const float* prj_ref, const float* w_ref, hipStream_t st);
    // D = A'(w A 1) with FP64 accumulation (NormImg of ReconNet_no_tf.py:359-394)
    void compute_norm(hipStream_t st);
    void set_norm(const float* norm_xy, hipStream_t st);
    const float* norm() const { return D_.as<float>(); }
    void set_image(const float* x_xy, hipStream_t st);
    void get_image(float* x_xy, hipStream_t st) const;
    float* image() { return X_[cur_].as<float>(); }
    // one pass over all subsets; if track != nullptr, per-slice FP64 sums of
    // (x_new - x_old)^2 and x_new^2 over the pass are returned in it (2*nslices)
    void pass(const SqsStep& o, hipStream_t st, std::vector<double>* track = nullptr);
    // refresh the transposed copy after the image was changed externally
    void sync_transposed(hipStream_t st);
    // Per-kernel timing with HIP events recorded on the pass's stream around
    // every launch (FP+residual, BP, update+transpose).  Totals in ms and the
    // number of timed subset steps since the last reset.
    void set_profiling(bool on);
    void kernel_times(double* ms_fp, double* ms_bp, double* ms_upd, long* steps) const;
    const FanTables& tables() const { return t_; }
    int nslices() const { return ns_; }
private:
    FanTables t_;
    int ns_ = 0, npers_ = 0, nsub_ = 0;
    DevBuf B_, W_, D_, X_[2], Xyx_, R_, G_, red_;
    // interleaved layout (groups of 4 slices as float4), used for 4 or more
    // slices unless RECONNET_SQS_INTERLEAVE=0: images in both orientations,
    // measurement, weights and the subset residual
    bool quads_ = false;
    // one FP launch over the whole subset (default), or one per view class
    // (RECONNET_SQS_FP_SPLIT=1)
    bool fp_split_ = false;
    // Qxy_[c], Qyx_[c]: the interleaved copies of image X_[c] (the update
    // writes the new image's copies while the current ones are read); D4_: the
    // normaliser as quads
    DevBuf Qxy_[2], Qyx_[2], B4_, W4_, R4_, D4_;
    // Within each subset the engine orders the views X class first (a view
    // with |cos| > |sin| walks image rows, the others columns): in XCD order
    // the interleaved forward projection's workgroups of one XCD then walk one
    // image orientation at a time, and the projection can also run as one
    // launch per class.  perm_[j] = caller's view index of engine view j;
    // sub_nx_[k] = number of X views of subset k.
    std::vector<int> perm_, sub_nx_;
    DevBuf perm_dev_;
    int cur_ = 0;
    bool have_w_ = false;
    bool prof_ = false;
    std::vector<hipEvent_t> ev_;
    double t_fp_ = 0, t_bp_ = 0, t_upd_ = 0;
    long t_steps_ = 0;
public:
    ~OsSqs();
};

} // namespace rn
