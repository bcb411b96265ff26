/*
 * C-ABI of libReconNet4D_no_tf.so for AMD Instinct MI355X (gfx950).
 *
 * Part 1 reproduces, symbol for symbol and argument for argument, the 17
 * extern "C" exports of the reference library (wudufan/ReconNet,
 * ReconNet4D_no_tf/), so that python/ReconNet_no_tf.py keeps working over
 * ctypes.  Their contract is the reference's: host pointers, caller-owned and
 * pre-allocated outputs, C-contiguous float32 arrays
 *   images    [nBatches][nx][ny][nz][nChannels]
 *   sinograms [nBatches][nu][nview][nv][nChannels]
 *   float3    arrays [nview][3]
 * and results on the host when the call returns.  Unlike the reference no C++
 * exception crosses the boundary: a failure prints "<function> failed: <reason>"
 * on stderr and is readable through rn_last_error() (part 2); outputs are then
 * left unspecified.
 *
 * Part 2 is additive: the error query, device-pointer entry points on
 * caller-owned HIP streams (used by the Python package and by torch), a
 * batched device-resident OS-SQS engine, and the BF16 convolution kernels of
 * the learned prior.  No torch types appear in any signature.
 */
#ifndef RECONNET_H
#define RECONNET_H

#ifdef __cplusplus
extern "C" {
#endif

/* ===================== Part 1: reference exports ===================== */

/* Projector.cu:12 */
int SetDevice(int device);

/* DistanceDriven.cu:386-389.  typeProjector is accepted and ignored, as in the reference FP. */
void cDistanceDrivenFanProjection(float* prj, const float* img, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector);

/* DistanceDriven.cu:793-796.  typeProjector 1 = FBP weighting, otherwise length pre-weight. */
void cDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector);

/* DistanceDriven3D.cu:346-350 (detU = (1,0,0), detV = (0,1,0) implied) */
void cDistanceDrivenTomoProjection(float* prj, const float* img,
        const float* detCenter, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* DistanceDriven3D.cu:666-670 */
void cDistanceDrivenTomoBackprojection(float* img, const float* prj,
        const float* detCenter, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* Filter.cu:432-435.  typeFilter 0 = Ram-Lak, 1 = Hamming, 2 = Hann, 3 = cosine. */
void cFilterFanFilter(float* pFPrj, const float* pPrj,
        int nBatches, int nChannels, int nu, int nview,
        int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeFilter, int typeProjector);

/* Filter.cu:476-479 */
void cFilterParallelFilter(float* pFPrj, const float* pPrj,
        int nBatches, int nChannels, int nu, int nview,
        int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeFilter, int typeProjector);

/* PixelDriven.cu:100-103 (always FBP-weighted, PixelDriven.cu:93) */
void cPixelDrivenFanBackprojection(float* pImg, const float* pPrj, const float* pDeg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector);

/* Siddon.cu:424-427 */
void cSiddonFanProjection(float* prj, const float* img, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector);

/* Siddon.cu:481-484 */
void cSiddonFanBackprojection(float* img, const float* prj, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector);

/* SiddonParallel.cu:152-156 (output zeroed first; the reference leaves it uninitialised) */
void cSiddonParallelProjection(float* prj, const float* img,
        const float* detCenter, const float* detU, const float* detV, const float* invRayDir,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* SiddonParallel.cu:230-234 (runs over every ray; the reference's launch grid misses rays) */
void cSiddonParallelBackprojection(float* img, const float* prj,
        const float* detCenter, const float* detU, const float* detV, const float* invRayDir,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* SiddonCone.cu:117-121 */
void cSiddonConeProjectionAbitrary(float* prj, const float* img,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* SiddonCone.cu:196-200 */
void cSiddonConeBackprojectionAbitrary(float* img, const float* prj,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v);

/* ossqs.cu:12-17.  A short last subset uses nview - startView views. */
void cOSSQS(float* res, const float* img, const float* prj,
        const float* normImg, const float* wprj,
        const float* degs, float lb, float ub, int nviewPerSubset, int nSubsets,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector, float lambda);

/* ossqs.cu:163-170 */
void cOSSQSConeAbitrary(float* res, const float* img, const float* prj,
        const float* normImg, const float* mask,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        float lb, float ub,
        int nviewPerSubset, int nSubsets, int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        int typeProjector, float lambda);

/* TotalVariation.cu:70.  The returned cost is reduced in FP64 (the reference sums FP32 on the host). */
float cTVSQS2D(float* grad, float* norm, const float* img, int nx, int ny);

/* Recon3D/HelicalBP.cu:132-135 (exported by the reference's separate libRecon3D).
 * Needs nu >= 2 (each row's two interpolation columns are read as a pair). */
void cHelicalBPFromParallelConebeamBackprojection(float* pImg, const float* pPrj,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v, float dsd, float dso,
        int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q);

/* ===================== Part 2: additive API (no reference counterpart) ===================== */

/* Message of the last failure on the calling thread ("" if the last call succeeded). */
const char* rn_last_error(void);

/* Fan geometry plan: trig tables, the float32 X/Y view split and workspaces
 * for one (geometry, angle set).  deg is a host array of nview angles.  A plan
 * is used by one host thread at a time.  Returns NULL on failure. */
void* rn_fan_plan_create(int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, const float* deg);
void rn_fan_plan_destroy(void* plan);

/* Operator kinds for rn_fan_forward / rn_fan_backward */
enum {
    RN_FP_DD = 0,          /* DDFPFanKernelX/Y */
    RN_FP_SIDDON = 1,      /* SiddonFanProjectionKernel */
    RN_BP_DD = 0,          /* DDBPFanKernelX/Y, iterative (length pre-weight) */
    RN_BP_DD_FBP = 1,      /* DDBPFanKernelX/Y, FBP weighting */
    RN_BP_DD_ADJOINT = 2,  /* exact transpose of RN_FP_DD */
    RN_BP_SIDDON = 3,      /* exact transpose of RN_FP_SIDDON (gather, no atomics) */
    RN_BP_PIXEL = 4        /* bpFanKernel3D (FBP weighting) */
};

/* Device pointers in the reference layouts, enqueued on `stream` (hipStream_t,
 * NULL = default stream).  Outputs are overwritten.  0 on success, -1 on error. */
int rn_fan_forward(void* plan, int kind, const float* img, float* prj, int nBatches, int nChannels, void* stream);
int rn_fan_backward(void* plan, int kind, const float* prj, float* img, int nBatches, int nChannels, void* stream);
/* Ramp filter (cFilterFanFilter when fan != 0, else cFilterParallelFilter) on device sinograms. */
int rn_fan_filter(void* plan, const float* prj, float* fprj, int nBatches, int nChannels,
        int typeFilter, int fan, void* stream);

/* Batched OS-SQS(-TV) engine over nslices 2D fan slices sharing one plan
 * geometry (nz = nv = 1).  Images are [nslices][nx][ny], sinograms
 * [nslices][nu][nview], all device pointers. */
void* rn_sqs_create(void* plan, int nslices, int nviewPerSubset, int nSubsets);
void rn_sqs_destroy(void* sqs);
int rn_sqs_set_data(void* sqs, const float* prj, const float* wprj, void* stream);
int rn_sqs_compute_norm(void* sqs, void* stream);        /* D = A'(w A 1), FP64 accumulation */
int rn_sqs_set_norm(void* sqs, const float* norm, void* stream);
int rn_sqs_get_norm(void* sqs, float* norm, void* stream);
int rn_sqs_set_image(void* sqs, const float* img, void* stream);
int rn_sqs_get_image(void* sqs, float* img, void* stream);
/* One pass over all subsets: x <- clamp(x - (lambda M g + 2 beta gTV) / (D + 2 beta nTV + 1e-6)).
 * beta = 0 reproduces cOSSQS.  If track != NULL it receives, per slice, the FP64 sums
 * |x_new - x_old|^2 and |x_new|^2 over the pass (2 * nslices doubles). */
int rn_sqs_pass(void* sqs, float lambda, float beta, float eps, float lb, float ub, double* track, void* stream);

/* Per-kernel HIP-event timing inside rn_sqs_pass (events recorded on the pass's stream
 * around the FP+residual, BP and update launches of every subset).  Enabling resets the
 * totals; rn_sqs_kernel_times returns total ms per kernel and the number of subset steps. */
int rn_sqs_set_profiling(void* sqs, int on);
int rn_sqs_kernel_times(void* sqs, double* ms_fp, double* ms_bp, double* ms_update, long* steps);

/* TV surrogate on nslices device images [nslices][nx][ny]; tv (host, may be NULL) receives
 * the FP64 per-slice costs. */
int rn_tv_sqs(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
        double* tv, void* stream);

/* No counterpart in the reference: rn_tv_sqs for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels and results as rn_tv_sqs; the FP64
 * per-tile partials live in a file-scope device array of 2^18 doubles (2 MiB
 * per device, held until the process ends), so
 * 2 * ceil(nx/16) * ceil(ny/16) * nslices must be <= 262144 (e.g. 128 slices
 * of 512 x 512), else -1 with an error and nothing launched.  tv_dev (device,
 * nslices doubles, may be NULL) receives the FP64 per-slice costs; nothing is
 * synchronised.  Calls that overlap on one device share the array and race. */
int rn_tv_sqs_noalloc(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
        double* tv_dev, void* stream);

/* Learned-prior 3x3 convolutions (no reference counterpart).  Images fp32
 * [B][H][W]; activations bf16 bit patterns (uint16) NHWC [B][H][W][32]; weights
 * w1, w5 fp32 [9][32], wt bf16 [9][32 out][32 in]; tap t = 3 (dy+1) + (dx+1);
 * zero padding.  All pointers are device pointers; 0 on success, -1 on error. */
int rn_cnn_conv_in(const float* x, const float* w1, const float* b1, unsigned short* a1,
        int B, int H, int W, void* stream);
/* mask == NULL: out = relu?(conv(in) + bias) (MFMA forward);
 * mask != NULL: out = conv(in) * (mask > 0) (data gradient with flipped, transposed weights) */
int rn_cnn_conv_mid(const unsigned short* in, const unsigned short* wt, const float* bias,
        const unsigned short* mask, unsigned short* out, int B, int H, int W, int relu, void* stream);
/* y = xres + b5 + conv(a4; w5) (xres may be NULL) */
int rn_cnn_conv_out(const unsigned short* a4, const float* w5, float b5, const float* xres, float* y,
        int B, int H, int W, void* stream);
/* The whole forward network in one launch, no activations kept: the same
 * values as rn_cnn_conv_in, three rn_cnn_conv_mid (relu) and rn_cnn_conv_out
 * with xres = x, bit for bit.  w2..w4 as wt of rn_cnn_conv_mid. */
int rn_cnn_forward(const float* x, const float* w1, const float* b1, const unsigned short* w2, const float* b2,
        const unsigned short* w3, const float* b3, const unsigned short* w4, const float* b4,
        const float* w5, float b5, float* y, int B, int H, int W, void* stream);
/* dz4 = (a4 > 0) * conv^T(dy; w5) */
int rn_cnn_out_dgrad(const float* dy, const float* w5, const unsigned short* a4, unsigned short* dz4,
        int B, int H, int W, void* stream);
/* mode 0: grad[9*32 + 32] of layer 1 (weights, bias) from x and dz1;
 * mode 1: grad[9*32 + 1] of layer 5 from dy and a4 */
int rn_cnn_wgrad_narrow(const float* img, const unsigned short* act, float* grad, int B, int H, int W,
        int mode, void* stream);
int rn_cnn_bias_grad(const unsigned short* dz, float* grad, long npx, void* stream);
/* grad[9][32 out][32 in] = sum_p dz[p][out] in[p + tap][in]  (MFMA) */
int rn_cnn_wgrad_mid(const unsigned short* in, const unsigned short* dz, float* grad, int B, int H, int W,
        void* stream);

/* Batches of 2D fan slices of uneven size (no reference counterpart).  dims
 * [nslices][4] = (nx, ny, nu, nview); geo [nslices][6] = (dx, dy, da, off_a,
 * dsd, dso); angles: the nview angles of every slice, concatenated (host).
 * Images are concatenated [nx][ny] planes, sinograms [nu][nview] planes (device).
 * Each call is one persistent launch over the whole batch (work queue across
 * slices); results equal independent per-slice calls. */
void* rn_ragged_create(int nslices, const int* dims, const float* geo, const float* angles);
void rn_ragged_destroy(void* batch);
int rn_ragged_forward(void* batch, int kind /* RN_FP_DD | RN_FP_SIDDON */, const float* img, float* prj,
        void* stream);
int rn_ragged_backward(void* batch, int kind /* RN_BP_DD_ADJOINT | RN_BP_SIDDON */, const float* prj, float* img,
        void* stream);

/* Isotropic TV separable surrogate of PriorFunctionSolver/TVSQS.cu:8-344
 * (device pointers, layout [nb][nx][ny][nz][nc], nz ignored for dims == 2):
 * var = sqrt(|backward differences|^2 + eps), s1 = gradient of sum(var),
 * s2 = dims / var + sum over axes of 1 / var(next voxel).  The same kernels
 * serve the host-pointer cTVSQS2D / cTVSQS3D of libPriorFunctionSolver.so
 * (include/priorfunctionsolver.h). */
int rn_tv_iso_sqs(const float* img, float* s1, float* s2, float* var, int nb, int nx, int ny, int nz, int nc,
        int dims, float eps, void* stream);

/* Circular cone-beam plans (no reference counterpart; the reference's DD 3D
 * pair, cDistanceDrivenTomo*, is tomosynthesis only).  Per-view detector
 * centre, axes and source as in cSiddonConeProjectionAbitrary (host arrays
 * [nview][3]); the DD pair needs detV = (0, 0, v) and detU.z = 0.  The DD
 * projector integrates, per view, over the voxel planes normal to the main
 * axis (x iff |C.x - S.x| >= |C.y - S.y|) the rectangle footprints of the
 * detector cells; its backward is the exact adjoint.  The Siddon kind is the
 * SiddonCone pair (SiddonCone.cu:117,196).  Device pointers, reference
 * layouts, caller stream.  A plan keeps one scratch buffer (the DD
 * backward's FP64 row sums, 8·nBatches·nChannels·nu·nview·(nv + 1) bytes):
 * calls on one plan must not overlap. */
enum { RN_CONE_DD = 0, RN_CONE_SIDDON = 1 };
void* rn_cone_plan_create(int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        const float* detCenter, const float* detU, const float* detV, const float* src);
void rn_cone_plan_destroy(void* plan);
int rn_cone_forward(void* plan, int kind, const float* img, float* prj, int nBatches, int nChannels, void* stream);
int rn_cone_backward(void* plan, int kind, const float* prj, float* img, int nBatches, int nChannels, void* stream);

/* cHelicalBPFromParallelConebeamBackprojection (Recon3D/HelicalBP.cu:132) on
 * device pointers and a caller stream: prj [B][nu][nview][nv][C] parallel-cone
 * data, img [B][nx][ny][nz][C], nu >= 2.  Its scratch (the projection
 * transposes) is kept per host thread: calls from one thread on different
 * streams must not overlap. */
int rn_helical_bp(const float* prj, float* img, int nBatches, int nChannels, int nx, int ny, int nz,
        float dx, float dy, float dz, int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        float dsd, float dso, int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q, void* stream);
/* Helical fan -> parallel-cone rebinning (no reference counterpart: the
 * reference's helical BP, Recon3D/HelicalBP.cu:10-83, consumes rebinned data).
 * fan: [B][nu_f][nview_f][nv_f][C], flat-detector views of a helical scan,
 * source angle beta0 + m dbeta at (dso cos b, dso sin b, z), detector at
 * distance dsd with columns along (sin b, -cos b, 0) and rows along z, cell
 * (iu, iv) at ((iu - off_u_f - (nu_f-1)/2) du_f, (iv - off_v_f - (nv_f-1)/2) dv_f)
 * (cSiddonConeProjectionAbitrary's detector).  par: [B][nu_p][nview_p][nv_p][C],
 * parallel view theta0 + k dtheta, u = (l - (nu_p-1)/2 - off_u_p) du_p,
 * v = (iv - (nv_p-1)/2 - off_v_p) dv_p on the cylinder of radius dsd around the
 * source: the input of cHelicalBPFromParallelConebeamBackprojection.
 * Trilinear interpolation, zero outside the raw data.  Device pointers. */
int rn_helical_rebin(const float* fan, float* par, int nBatches, int nChannels,
        int nu_f, int nview_f, int nv_f, float du_f, float dv_f, float off_u_f, float off_v_f,
        float beta0, float dbeta,
        int nu_p, int nview_p, int nv_p, float du_p, float dv_p, float off_u_p, float off_v_p,
        float theta0, float dtheta, float dsd, float dso, void* stream);

/* Launch-shape autotuner (no reference counterpart).  On the first launch of
 * a shape key the projector launchers time their candidate block / tile shapes
 * on the caller's stream and keep the fastest for the process; the choice
 * never changes results.  Enabled unless RECONNET_AUTOTUNE=0 or switched off
 * here.  rn_autotune_report writes "kernel a b c d e -> choice (ms ...)"
 * lines into buf (NUL-terminated, truncated to len) and returns the full
 * length. */
int rn_autotune_enable(int on);
int rn_autotune_report(char* buf, int len);
int rn_autotune_reset(void);
/* Force candidate `candidate` of kernel `kernel` (1 Siddon FP, 2 Siddon BP,
 * 3 DD FP, 4 DD adjoint, 5 DD BP, 7 DD FBP BP; 8 / 9 Siddon FP / BP,
 * 10 / 11 DD BP / FBP BP and 12 DD FP on the interleaved 4-slice layout used
 * for 4 or more slices; 13 the CNN's 32 -> 32 convolution, 14 the one-launch CNN forward, 15 the
 * helical backprojection, 16 the cone DD forward projector, 17 the pixel-driven BP, 18 the Siddon cone
 * backprojection of upright detectors, 19 the cone DD adjoint) for every shape, bypassing the
 * timing; candidate < 0 removes the pin.  A launch of a kernel with fewer
 * candidates fails with an error.  Lets tests run every launch shape. */
int rn_autotune_pin(int kernel, int candidate);

#ifdef __cplusplus
}
#endif

#endif /* RECONNET_H */
