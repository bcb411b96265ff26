// OS-SQS with the Siddon fan projector (ossqs.cu:12-161) and the TV surrogate
// (TotalVariation.cu:9-142), device resident.
//
// One subset step is three launches over the whole slice batch:
//   1. Siddon FP of the subset views with the residual w (Ax - b) fused into
//      the epilogue (replaces Minus2D + Multiply2D, ossqs.cu:105-122);
//   2. pixel-driven gather BP of the residual (no atomics);
//   3. a fused update: TV gradient / curvature stencil, the SQS step
//      x - (lambda M g + 2 beta gTV) / (D + 2 beta nTV + 1e-6), the crop to
//      [lb, ub], and the transposed copy that the next FP's X views read.
// With the interleaved layout and an LDS-staged backprojection shape, 2 and 3
// are one launch: the backprojection's epilogue applies the update to the
// gradient it holds in registers (the gradient image is not written).
// The normaliser D = A'(w A 1) is accumulated in FP64; so are the TV cost and
// the convergence sums.

#include "rn_sqs.h"
#include "rn_fan_device.h"
#include "rn_fan_ops.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>

#include <cstring>

namespace rn {

namespace {

constexpr int TB = 16;   // 16 x 16 pixel tiles

__device__ __forceinline__ void tv_terms(const float* __restrict__ X, int ix, int iy, int nx, int ny, float eps,
                                         float& gtv, float& ntv, float& var)
{
    const long ind = (long)ix * ny + iy;
    const float x = X[ind];
    const float dx0 = ix == 0 ? x : x - X[ind - ny];
    const float dx1 = ix == nx - 1 ? x : x - X[ind + ny];
    const float dy0 = iy == 0 ? x : x - X[ind - 1];
    const float dy1 = iy == ny - 1 ? x : x - X[ind + 1];
    const float t0 = sqrtf(dx0 * dx0 + eps), t1 = sqrtf(dx1 * dx1 + eps);
    const float t2 = sqrtf(dy0 * dy0 + eps), t3 = sqrtf(dy1 * dy1 + eps);
    gtv = (dx0 / t0 + dx1 / t1 + dy0 / t2 + dy1 / t3) / 2;
    ntv = 1 / t0 + 1 / t1 + 1 / t2 + 1 / t3;
    var = (t0 + t1 + t2 + t3) / 2;
}

// block-wide FP64 sum of up to two values; thread 0 writes the partials
__device__ __forceinline__ void block_sum2(double a, double b, double* out)
{
    __shared__ double sa[TB * TB], sb[TB * TB];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    sa[t] = a; sb[t] = b;
    __syncthreads();
    for (int k = (blockDim.x * blockDim.y) / 2; k > 0; k >>= 1) {
        if (t < k) { sa[t] += sa[t + k]; sb[t] += sb[t + k]; }
        __syncthreads();
    }
    if (t == 0) { out[0] = sa[0]; out[1] = sb[0]; }
}

__global__ __launch_bounds__(256) void k_tv_sqs(const float* __restrict__ img, float* __restrict__ grad,
                                                float* __restrict__ norm, int nx, int ny, float eps,
                                                double* __restrict__ partial)
{
    const int iy = blockIdx.x * TB + threadIdx.x;
    const int ix = blockIdx.y * TB + threadIdx.y;
    const int s = blockIdx.z;
    const long plane = (long)nx * ny;
    double v = 0.0;
    if (ix < nx && iy < ny) {
        float g, n, var;
        tv_terms(img + s * plane, ix, iy, nx, ny, eps, g, n, var);
        grad[s * plane + (long)ix * ny + iy] = g;
        norm[s * plane + (long)ix * ny + iy] = n;
        v = (double)var;
    }
    const long nb = (long)gridDim.x * gridDim.y;
    block_sum2(v, 0.0, partial + 2 * (s * nb + (long)blockIdx.y * gridDim.x + blockIdx.x));
}

__global__ __launch_bounds__(256) void k_sqs_update(const float* __restrict__ X, float* __restrict__ Xn,
                                                    const float* __restrict__ G, const float* __restrict__ D,
                                                    int nx, int ny, float step, float beta2, float eps,
                                                    float lb, float ub, int crop)
{
    const int iy = blockIdx.x * TB + threadIdx.x;
    const int ix = blockIdx.y * TB + threadIdx.y;
    const int s = blockIdx.z;
    const long plane = (long)nx * ny;
    if (ix >= nx || iy >= ny) return;
    const long pix = (long)ix * ny + iy, i = s * plane + pix;
    Xn[i] = sqs_pixel_update(X + s * plane, pix, ix, iy, nx, ny, G[i], D[i], step, beta2, eps, lb, ub, crop);
}

// k_sqs_update for the interleaved layout, fused with the layout copies: a
// 32 x 32 pixel tile of one group of 4 slices; each slice's new image goes to
// Xn and, as the float4 of the 4 slices, to the padded Qxy [q][nx][qpitch(ny)]
// and (through LDS) to the transposed Qyx [q][ny][qpitch(nx)].  Slices past
// ns hold 0.
__global__ __launch_bounds__(256) void k_sqs_update_q(const float* __restrict__ X, float* __restrict__ Xn,
                                                      const float* __restrict__ G, const float* __restrict__ D,
                                                      float4* __restrict__ Qxy, float4* __restrict__ Qyx, int nx,
                                                      int ny, int ns, float step, float beta2, float eps, float lb,
                                                      float ub, int crop)
{
    __shared__ float4 tile[32][33];
    const int iy0 = blockIdx.x * 32, ix0 = blockIdx.y * 32, q = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const long plane = (long)nx * ny;
    const int pxy = qpitch(ny), pyx = qpitch(nx);
    for (int r = ty; r < 32; r += 8) {
        const int ix = ix0 + r, iy = iy0 + tx;
        if (ix >= nx || iy >= ny) continue;
        const long pix = (long)ix * ny + iy;
        float v[4];
        #pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int s = 4 * q + c;
            v[c] = 0.0f;
            if (s >= ns) continue;
            const long i = s * plane + pix;
            const float xn = sqs_pixel_update(X + s * plane, pix, ix, iy, nx, ny, G[i], D[i], step, beta2, eps, lb,
                                              ub, crop);
            Xn[i] = xn;
            v[c] = xn;
        }
        const float4 f = make_float4(v[0], v[1], v[2], v[3]);
        tile[r][tx] = f;
        Qxy[(long)q * nx * pxy + (long)ix * pxy + kQPadL + iy] = f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int iy = iy0 + r, ix = ix0 + tx;
        if (ix < nx && iy < ny) Qyx[(long)q * ny * pyx + (long)iy * pyx + kQPadL + ix] = tile[tx][r];
    }
}

// per-block FP64 partials of sum (b - a)^2 and sum b^2
__global__ __launch_bounds__(256) void k_diff_sums(const float* __restrict__ a, const float* __restrict__ b,
                                                   int nx, int ny, double* __restrict__ partial)
{
    const int iy = blockIdx.x * TB + threadIdx.x;
    const int ix = blockIdx.y * TB + threadIdx.y;
    const int s = blockIdx.z;
    const long plane = (long)nx * ny;
    double dd = 0.0, bb = 0.0;
    if (ix < nx && iy < ny) {
        const long i = s * plane + (long)ix * ny + iy;
        const double d = (double)b[i] - (double)a[i];
        dd = d * d;
        bb = (double)b[i] * (double)b[i];
    }
    const long nb = (long)gridDim.x * gridDim.y;
    block_sum2(dd, bb, partial + 2 * (s * nb + (long)blockIdx.y * gridDim.x + blockIdx.x));
}

// sum the per-block partials of each slice (FP64), accumulate into out[2*s..];
// first_only: out[s] = sum of the first values only
__global__ void k_reduce_partials(const double* __restrict__ partial, long nb, double* __restrict__ out, int accumulate,
                                  int first_only = 0)
{
    __shared__ double sa[256], sb[256];
    const int s = blockIdx.x;
    double a = 0, b = 0;
    for (long i = threadIdx.x; i < nb; i += blockDim.x) {
        a += partial[2 * (s * nb + i)];
        b += partial[2 * (s * nb + i) + 1];
    }
    sa[threadIdx.x] = a; sb[threadIdx.x] = b;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) { sa[threadIdx.x] += sa[threadIdx.x + k]; sb[threadIdx.x] += sb[threadIdx.x + k]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (first_only) out[s] = sa[0];
        else if (accumulate) { out[2 * s] += sa[0]; out[2 * s + 1] += sb[0]; }
        else { out[2 * s] = sa[0]; out[2 * s + 1] = sb[0]; }
    }
}

__global__ void k_fill_f64(double* p, double v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_weight_rows_f64(const double* __restrict__ fp1, const float* __restrict__ w, double* __restrict__ y,
                                  size_t per_slice)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s = blockIdx.y;
    if (i < per_slice) y[s * per_slice + i] = fp1[i] * (double)w[s * per_slice + i];
}

__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// fp[(bu, j, q)] = w (fp - meas) with (bu, start + j, q) the element of the
// full reference-layout sinogram [bu = b nu + iu][view][q = iv nC + c]
__global__ __launch_bounds__(256) void k_os_residual_ref(float* __restrict__ fp, const float* __restrict__ meas,
                                                         const float* __restrict__ w, long n, int cur, int nview,
                                                         int inner, int start)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long q = i % inner;
    const long t = i / inner;
    const long j = t % cur;
    const long bu = t / cur;
    const long src = (bu * nview + start + j) * inner + q;
    fp[i] = (fp[i] - meas[src]) * w[src];
}

// ossqs.cu:133-145: u = g / (D + 1e-6); x = x - step u; crop to [lb, ub] if lb < ub
__global__ __launch_bounds__(256) void k_os_update_ref(float* __restrict__ x, const float* __restrict__ g,
                                                       const float* __restrict__ D, long n, float step, float lb,
                                                       float ub, int crop)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = g[i] / (D[i] + 1e-6f);
    float v = x[i] - step * u;
    if (crop) v = fminf(fmaxf(v, lb), ub);
    x[i] = v;
}

} // namespace

void os_sqs_ref_layout(const FanGeo& g, const float* deg_host, int nB, int nC, int nviewPerSubset, int nSubsets,
                       float* x, const float* prj, const float* w, const float* D, float lambda, float lb, float ub,
                       hipStream_t st)
{
    RN_REQUIRE(nviewPerSubset > 0 && nSubsets > 0, "nviewPerSubset and nSubsets must be positive");
    const long nimg = (long)nB * g.nx * g.ny * g.nz * nC;
    const int inner = g.nv * nC;
    DevBuf fp((size_t)nB * g.nu * nviewPerSubset * inner * sizeof(float));
    DevBuf grad((size_t)nimg * sizeof(float));
    FanWork work;
    // one table set per subset: a table is not rebuilt while launches that
    // read it may still be queued
    std::vector<FanTables> tabs(nSubsets);
    const float step = lambda * (float)nSubsets;
    for (int s = 0, start = 0; s < nSubsets && start < g.nview; s++) {
        const int cur = std::min(nviewPerSubset, g.nview - start);
        FanGeo gs = g;
        gs.nview = cur;
        FanTables& tab = tabs[s];
        tab.build(gs, deg_host + start, st);
        fan_forward(tab, FanFP::Siddon, nB, nC, x, fp.as<float>(), work, st);
        const long nres = (long)nB * g.nu * cur * inner;
        k_os_residual_ref<<<cdiv(nres, 256), 256, 0, st>>>(fp.as<float>(), prj, w, nres, cur, g.nview, inner, start);
        RN_HIP_CHECK(hipGetLastError());
        fan_backward(tab, FanBP::Siddon, nB, nC, fp.as<float>(), grad.as<float>(), work, st);
        k_os_update_ref<<<cdiv(nimg, 256), 256, 0, st>>>(x, grad.as<float>(), D, nimg, step, lb, ub, lb < ub ? 1 : 0);
        RN_HIP_CHECK(hipGetLastError());
        start += cur;
    }
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

void tv_sqs(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
            std::vector<double>* tv, DevBuf& scratch, hipStream_t st)
{
    RN_REQUIRE(nx > 0 && ny > 0 && nslices > 0, "TV dimensions must be positive");
    dim3 blk(TB, TB), grd(cdiv(ny, TB), cdiv(nx, TB), nslices);
    const long nb = (long)grd.x * grd.y;
    scratch.ensure((2 * nb * nslices + 2 * (size_t)nslices) * sizeof(double));
    double* partial = scratch.as<double>();
    double* out = partial + 2 * nb * nslices;
    hipLaunchKernelGGL(k_tv_sqs, grd, blk, 0, st, img, grad, norm, nx, ny, eps, partial);
    check_launch("tv_sqs");
    hipLaunchKernelGGL(k_reduce_partials, dim3(nslices), dim3(256), 0, st, partial, nb, out, 0, 0);
    check_launch("reduce_partials");
    if (tv) {
        std::vector<double> h(2 * (size_t)nslices);
        RN_HIP_CHECK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        RN_HIP_CHECK(hipStreamSynchronize(st));
        tv->resize(nslices);
        for (int s = 0; s < nslices; s++) (*tv)[s] = h[2 * s];
    }
}

// ---- TV surrogate without allocation ----
// The FP64 per-workgroup cost partials of k_tv_sqs live in this file-scope
// device array instead of a library-allocated buffer.  Host code never names
// it except to resolve its address: each device holds its own copy, at the
// address hipGetSymbolAddress returns with that device current, looked up
// once per device and kept.
namespace {
static __device__ __align__(256) double scratch[kTvScratchDoubles];

double* tv_scratch_ptr()
{
    constexpr int kMaxDevices = 64;
    static std::once_flag once[kMaxDevices];
    static double* ptr[kMaxDevices] = {};
    static hipError_t err[kMaxDevices] = {};
    int dev = 0;
    RN_HIP_CHECK(hipGetDevice(&dev));
    RN_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::call_once(once[dev], [&] {
        void* p = nullptr;
        err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        ptr[dev] = static_cast<double*>(p);
    });
    if (err[dev] != hipSuccess) throw Error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err[dev]));
    RN_REQUIRE(ptr[dev] != nullptr, "no device address for the TV scratch array");
    return ptr[dev];
}
} // namespace

long tv_sqs_scratch_doubles(int nx, int ny, int nslices)
{
    return 2L * cdiv(ny, TB) * cdiv(nx, TB) * (long)nslices;
}

void tv_sqs_noalloc(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
                    double* tv_dev, hipStream_t st)
{
    RN_REQUIRE(nx > 0 && ny > 0 && nslices > 0, "TV dimensions must be positive");
    const long need = tv_sqs_scratch_doubles(nx, ny, nslices);
    RN_REQUIRE(need <= kTvScratchDoubles,
               "TV partials need " + std::to_string(need) + " doubles, beyond the " +
               std::to_string(kTvScratchDoubles) + "-double scratch array (2 per 16 x 16 tile and slice)");
    double* partial = tv_scratch_ptr();
    dim3 blk(TB, TB), grd(cdiv(ny, TB), cdiv(nx, TB), nslices);
    const long nb = (long)grd.x * grd.y;
    // every partial is written by k_tv_sqs before the reduction reads it:
    // nothing in the array has to start at zero
    hipLaunchKernelGGL(k_tv_sqs, grd, blk, 0, st, img, grad, norm, nx, ny, eps, partial);
    check_launch("tv_sqs");
    if (tv_dev) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(nslices), dim3(256), 0, st, partial, nb, tv_dev, 0, 1);
        check_launch("reduce_partials");
    }
}

void OsSqs::setup(const FanGeo& g, const float* deg_host, int nslices, int nviewPerSubset, int nSubsets,
                  hipStream_t st)
{
    RN_REQUIRE(nslices > 0, "number of slices must be positive");
    RN_REQUIRE(nviewPerSubset > 0 && nSubsets > 0, "subset sizes must be positive");
    RN_REQUIRE(g.nz == 1 && g.nv == 1, "OS-SQS engine handles 2D slices (nz = nv = 1)");
    ns_ = nslices; npers_ = nviewPerSubset; nsub_ = nSubsets;
    // X-class views first within each subset (stable); the same float32 class
    // decision as the trig tables
    perm_.resize(g.nview);
    sub_nx_.assign(nSubsets, 0);
    std::vector<float> deg(g.nview);
    for (int k = 0, start = 0; k < nSubsets && start < g.nview; k++) {
        const int end = std::min(start + nviewPerSubset, g.nview);
        int j = start;
        for (int v = start; v < end; v++)
            if (std::fabs(cosf(deg_host[v])) > std::fabs(sinf(deg_host[v]))) perm_[j++] = v;
        sub_nx_[k] = j - start;
        for (int v = start; v < end; v++)
            if (!(std::fabs(cosf(deg_host[v])) > std::fabs(sinf(deg_host[v])))) perm_[j++] = v;
        start = end;
    }
    for (int j = 0; j < g.nview; j++) deg[j] = deg_host[perm_[j]];
    if (!t_.matches(g, deg.data())) t_.build(g, deg.data(), st);
    perm_dev_.ensure(g.nview * sizeof(int));
    RN_HIP_CHECK(hipMemcpyAsync(perm_dev_.get(), perm_.data(), g.nview * sizeof(int), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    const size_t plane = (size_t)g.nx * g.ny, sino = (size_t)g.nu * g.nview;
    B_.ensure(ns_ * sino * sizeof(float));
    W_.ensure(ns_ * sino * sizeof(float));
    D_.ensure(ns_ * plane * sizeof(float));
    X_[0].ensure(ns_ * plane * sizeof(float));
    X_[1].ensure(ns_ * plane * sizeof(float));
    G_.ensure(ns_ * plane * sizeof(float));
    const char* env = std::getenv("RECONNET_SQS_INTERLEAVE");
    quads_ = ns_ >= 4 && !(env && env[0] == '0');
    const char* split = std::getenv("RECONNET_SQS_FP_SPLIT");
    fp_split_ = split && split[0] == '1';
    if (quads_) {
        const size_t nq = (size_t)(ns_ + 3) / 4;
        // padded lines; the pad cells stay zero (only line interiors are written)
        const size_t nxy = nq * g.nx * (size_t)qpitch(g.ny), nyx = nq * g.ny * (size_t)qpitch(g.nx);
        for (int c = 0; c < 2; c++) {
            Qxy_[c].ensure(nxy * sizeof(float4));
            Qyx_[c].ensure(nyx * sizeof(float4));
            launch_fill(Qxy_[c].as<float>(), 0.0f, 4 * nxy, st);
            launch_fill(Qyx_[c].as<float>(), 0.0f, 4 * nyx, st);
        }
        D4_.ensure(nq * plane * sizeof(float4));
        B4_.ensure(nq * sino * sizeof(float4));
        W4_.ensure(nq * sino * sizeof(float4));
        R4_.ensure(nq * (size_t)g.nu * npers_ * sizeof(float4));
    } else {
        Xyx_.ensure(ns_ * plane * sizeof(float));
        R_.ensure(ns_ * (size_t)g.nu * npers_ * sizeof(float));
    }
}

void OsSqs::set_data(const float* prj_ref, const float* w_ref, hipStream_t st)
{
    const FanGeo& g = t_.geo;
    // view-major, then the engine's view order
    DevBuf tmp((size_t)ns_ * g.nu * g.nview * sizeof(float));
    launch_transpose(prj_ref, tmp.as<float>(), g.nu, g.nview, ns_, st);
    launch_permute_rows(tmp.as<float>(), B_.as<float>(), perm_dev_.as<int>(), g.nview, g.nu, ns_, st);
    if (w_ref) {
        launch_transpose(w_ref, tmp.as<float>(), g.nu, g.nview, ns_, st);
        launch_permute_rows(tmp.as<float>(), W_.as<float>(), perm_dev_.as<int>(), g.nview, g.nu, ns_, st);
    } else {
        launch_fill(W_.as<float>(), 1.0f, (size_t)ns_ * g.nu * g.nview, st);
    }
    if (quads_) {
        launch_pack_quads(B_.as<float>(), B4_.as<float4>(), nullptr, g.nview, g.nu, ns_, false, st);
        launch_pack_quads(W_.as<float>(), W4_.as<float4>(), nullptr, g.nview, g.nu, ns_, false, st);
    }
    have_w_ = true;
}

void OsSqs::compute_norm(hipStream_t st)
{
    const FanGeo& g = t_.geo;
    const size_t plane = (size_t)g.nx * g.ny, sino = (size_t)g.nu * g.nview;
    // FP64 scratch: ones image (1 slice), its projection, the weighted
    // projections of every slice and the FP64 normaliser
    DevBuf ones(plane * sizeof(double)), fp1(sino * sizeof(double));
    DevBuf y((size_t)ns_ * sino * sizeof(double)), d64((size_t)ns_ * plane * sizeof(double));
    hipLaunchKernelGGL(k_fill_f64, dim3(cdiv((long)plane, 256)), dim3(256), 0, st, ones.as<double>(), 1.0, plane);
    check_launch("fill_f64");
    launch_siddon_fan_fp_f64(t_.k, ones.as<double>(), ones.as<double>(), fp1.as<double>(), 1, 0, g.nview, st);
    hipLaunchKernelGGL(k_weight_rows_f64, dim3(cdiv((long)sino, 256), ns_), dim3(256), 0, st,
                       fp1.as<double>(), W_.as<float>(), y.as<double>(), sino);
    check_launch("weight_rows_f64");
    launch_siddon_fan_bp_f64(t_.k, y.as<double>(), d64.as<double>(), ns_, 0, g.nview, st);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(cdiv((long)(ns_ * plane), 256)), dim3(256), 0, st,
                       d64.as<double>(), D_.as<float>(), ns_ * plane);
    check_launch("f64_to_f32");
    if (quads_) launch_pack_quads(D_.as<float>(), D4_.as<float4>(), nullptr, g.nx, g.ny, ns_, false, st);
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

void OsSqs::set_norm(const float* norm_xy, hipStream_t st)
{
    const size_t n = (size_t)ns_ * t_.geo.nx * t_.geo.ny;
    RN_HIP_CHECK(hipMemcpyAsync(D_.get(), norm_xy, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (quads_)
        launch_pack_quads(D_.as<float>(), D4_.as<float4>(), nullptr, t_.geo.nx, t_.geo.ny, ns_, false, st);
}

void OsSqs::sync_transposed(hipStream_t st)
{
    if (quads_) launch_pack_quads(X_[cur_].as<float>(), Qxy_[cur_].as<float4>(), Qyx_[cur_].as<float4>(), t_.geo.nx,
                                  t_.geo.ny, ns_, true, st);
    else launch_transpose(X_[cur_].as<float>(), Xyx_.as<float>(), t_.geo.nx, t_.geo.ny, ns_, st);
}

OsSqs::~OsSqs()
{
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
}

void OsSqs::set_profiling(bool on)
{
    prof_ = on;
    t_fp_ = t_bp_ = t_upd_ = 0;
    t_steps_ = 0;
}

void OsSqs::kernel_times(double* ms_fp, double* ms_bp, double* ms_upd, long* steps) const
{
    if (ms_fp) *ms_fp = t_fp_;
    if (ms_bp) *ms_bp = t_bp_;
    if (ms_upd) *ms_upd = t_upd_;
    if (steps) *steps = t_steps_;
}

void OsSqs::set_image(const float* x_xy, hipStream_t st)
{
    const size_t n = (size_t)ns_ * t_.geo.nx * t_.geo.ny;
    RN_HIP_CHECK(hipMemcpyAsync(X_[cur_].get(), x_xy, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    sync_transposed(st);
}

void OsSqs::get_image(float* x_xy, hipStream_t st) const
{
    const size_t n = (size_t)ns_ * t_.geo.nx * t_.geo.ny;
    RN_HIP_CHECK(hipMemcpyAsync(x_xy, X_[cur_].get(), n * sizeof(float), hipMemcpyDeviceToDevice, st));
}

void OsSqs::pass(const SqsStep& o, hipStream_t st, std::vector<double>* track)
{
    RN_REQUIRE(have_w_, "OS-SQS data not set");
    const FanGeo& g = t_.geo;
    const size_t plane = (size_t)g.nx * g.ny;
    dim3 blk(TB, TB), grd(cdiv(g.ny, TB), cdiv(g.nx, TB), ns_);
    const long nb = (long)grd.x * grd.y;
    if (track) {
        red_.ensure((2 * nb * ns_ + 2 * (size_t)ns_) * sizeof(double) + ns_ * plane * sizeof(float));
        float* x0 = reinterpret_cast<float*>(red_.as<double>() + 2 * nb * ns_ + 2 * (size_t)ns_);
        RN_HIP_CHECK(hipMemcpyAsync(x0, X_[cur_].get(), ns_ * plane * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    const bool crop = o.lb < o.ub;
    if (prof_ && ev_.size() < 4 * (size_t)nsub_) {
        for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
        ev_.assign(4 * (size_t)nsub_, nullptr);
        for (auto& e : ev_) RN_HIP_CHECK(hipEventCreate(&e));
    }
    int start = 0, nsteps = 0;
    for (int k = 0; k < nsub_ && start < g.nview; k++) {
        const int cur = std::min(npers_, g.nview - start);
        float* X = X_[cur_].as<float>();
        float* Xn = X_[1 - cur_].as<float>();
        if (prof_) RN_HIP_CHECK(hipEventRecord(ev_[4 * k], st));
        if (quads_) {
            // one launch over the subset (each wave reads the orientation of
            // its view's class; the views are ordered X class first), or one
            // per class
            const int nxv = fp_split_ ? sub_nx_[k] : cur;
            const float4 *qyx = Qyx_[cur_].as<float4>(), *qxy = Qxy_[cur_].as<float4>();
            if (nxv > 0)
                launch_siddon_fan_fp_residual_q(t_.k, qyx, qxy, R4_.as<float4>(), ns_, start, nxv, B4_.as<float4>(),
                                                W4_.as<float4>(), cur, 0, st);
            if (nxv < cur)
                launch_siddon_fan_fp_residual_q(t_.k, qyx, qxy, R4_.as<float4>(), ns_, start + nxv, cur - nxv,
                                                B4_.as<float4>(), W4_.as<float4>(), cur, nxv, st);
        } else {
            launch_siddon_fan_fp_residual(t_.k, Xyx_.as<float>(), X, R_.as<float>(), ns_, start, cur,
                                          B_.as<float>(), W_.as<float>(), st);
        }
        if (prof_) RN_HIP_CHECK(hipEventRecord(ev_[4 * k + 1], st));
        // interleaved layout: the update runs in the backprojection's epilogue
        // when the tuned shape allows it (then G is not written)
        bool fused = false;
        if (quads_) {
            const SqsUpd u{Qxy_[cur_].as<float4>(), D4_.as<float4>(), Xn, Qxy_[1 - cur_].as<float4>(),
                           Qyx_[1 - cur_].as<float4>(), o.lambda * nsub_, 2.0f * o.beta, o.eps, o.lb, o.ub,
                           crop ? 1 : 0};
            fused = launch_siddon_fan_bp_q_update(t_.k, R4_.as<float4>(), G_.as<float>(), ns_, start, cur, u, st);
        } else {
            launch_siddon_fan_bp(t_.k, R_.as<float>(), G_.as<float>(), ns_, start, cur, false, st);
        }
        if (prof_) RN_HIP_CHECK(hipEventRecord(ev_[4 * k + 2], st));
        if (fused) {
            // done in the backprojection
        } else if (quads_) {
            const dim3 gq(cdiv(g.ny, 32), cdiv(g.nx, 32), cdiv(ns_, 4));
            hipLaunchKernelGGL(k_sqs_update_q, gq, dim3(256), 0, st, X, Xn, G_.as<float>(), D_.as<float>(),
                               Qxy_[1 - cur_].as<float4>(), Qyx_[1 - cur_].as<float4>(), g.nx, g.ny, ns_,
                               o.lambda * nsub_, 2.0f * o.beta, o.eps, o.lb, o.ub, crop ? 1 : 0);
            check_launch("sqs_update_q");
        } else {
            hipLaunchKernelGGL(k_sqs_update, grd, blk, 0, st, X, Xn, G_.as<float>(), D_.as<float>(), g.nx, g.ny,
                               o.lambda * nsub_, 2.0f * o.beta, o.eps, o.lb, o.ub, crop ? 1 : 0);
            check_launch("sqs_update");
            launch_transpose(Xn, Xyx_.as<float>(), g.nx, g.ny, ns_, st);
        }
        if (prof_) RN_HIP_CHECK(hipEventRecord(ev_[4 * k + 3], st));
        cur_ = 1 - cur_;
        start += cur;
        nsteps++;
    }
    if (prof_) {
        RN_HIP_CHECK(hipEventSynchronize(ev_[4 * (nsteps - 1) + 3]));
        for (int k = 0; k < nsteps; k++) {
            float a = 0, b = 0, c = 0;
            RN_HIP_CHECK(hipEventElapsedTime(&a, ev_[4 * k], ev_[4 * k + 1]));
            RN_HIP_CHECK(hipEventElapsedTime(&b, ev_[4 * k + 1], ev_[4 * k + 2]));
            RN_HIP_CHECK(hipEventElapsedTime(&c, ev_[4 * k + 2], ev_[4 * k + 3]));
            t_fp_ += a; t_bp_ += b; t_upd_ += c;
        }
        t_steps_ += nsteps;
    }
    if (track) {
        double* partial = red_.as<double>();
        double* sums = partial + 2 * nb * ns_;
        const float* x0 = reinterpret_cast<const float*>(sums + 2 * (size_t)ns_);
        hipLaunchKernelGGL(k_diff_sums, grd, blk, 0, st, x0, X_[cur_].as<float>(), g.nx, g.ny, partial);
        check_launch("diff_sums");
        hipLaunchKernelGGL(k_reduce_partials, dim3(ns_), dim3(256), 0, st, partial, nb, sums, 0, 0);
        check_launch("reduce_partials");
        track->resize(2 * (size_t)ns_);
        RN_HIP_CHECK(hipMemcpyAsync(track->data(), sums, track->size() * sizeof(double), hipMemcpyDeviceToHost, st));
        RN_HIP_CHECK(hipStreamSynchronize(st));
    }
}

} // namespace rn
