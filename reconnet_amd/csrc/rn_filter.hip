// Ramp filtering of fan / parallel sinograms (Filter.cu:66-518).
//
// The reference zero-pads each detector row to L = 2nu-1 samples, multiplies
// by the cosine pre-weight, and applies the windowed ramp through cuFFT
// (R2C, spectrum product, C2R).  Because of the padding that product is an
// exact linear convolution with the (2nu-1)-tap kernel, so here it is computed
// directly: one workgroup per detector row, the weighted row and the taps
// staged in LDS, FP32 accumulation with Kahan compensation over the large
// near taps.  The taps themselves are
// built on the host in FP64 (closed-form Ram-Lak taps, window applied through
// an exact FP64 DFT round trip, Filter.cu:90-193).

#include "rn_filter.h"

#include <cmath>
#include <vector>

namespace rn {

namespace {

__global__ __launch_bounds__(256) void k_ramp_filter(const float* __restrict__ in_vu, float* __restrict__ out_vu,
                                                     const float* __restrict__ taps, const float* __restrict__ pw,
                                                     int nu)
{
    extern __shared__ float sm[];
    float* xs = sm;          // nu weighted samples
    float* hs = sm + nu;     // 2nu-1 taps (scaled)
    const long row = blockIdx.x;
    const float* src = in_vu + row * nu;
    for (int j = threadIdx.x; j < nu; j += blockDim.x) xs[j] = src[j] * pw[j];
    for (int j = threadIdx.x; j < 2 * nu - 1; j += blockDim.x) hs[j] = taps[j];
    __syncthreads();
    // out[iu] = sum_j x[j] h[iu - j + nu - 1].  The taps fall off as 1/k^2:
    // the terms within kNear of the wave's 64 outputs (a band uniform over the
    // wave) are summed with Kahan compensation, the small far terms first with
    // plain FMAs.
    constexpr int kNear = 32;
    const int lane = threadIdx.x & 63;
    const int nu4 = nu & ~3;
    const float4* xs4 = reinterpret_cast<const float4*>(xs);   // 4 samples per broadcast LDS read
    for (int iu = threadIdx.x; iu - lane < nu; iu += blockDim.x) {
        const int iu0 = iu - lane;                   // the wave's first output
        const bool act = iu < nu;
        const float* hp = hs + min(iu, nu - 1) + nu - 1;
        // Kahan band [b0, b1) in whole groups of 4 samples (plus the < 4 samples past nu4)
        const int b0 = max(iu0 - kNear, 0) & ~3;
        const int b1 = min(((iu0 + 63 + kNear) | 3) + 1, nu4);
        float f = 0.0f;
        for (int j = 0; j < b0; j += 4) {
            const float4 x = xs4[j >> 2];
            f = fmaf(x.x, hp[-j], f);
            f = fmaf(x.y, hp[-j - 1], f);
            f = fmaf(x.z, hp[-j - 2], f);
            f = fmaf(x.w, hp[-j - 3], f);
        }
        for (int j = b1; j < nu4; j += 4) {
            const float4 x = xs4[j >> 2];
            f = fmaf(x.x, hp[-j], f);
            f = fmaf(x.y, hp[-j - 1], f);
            f = fmaf(x.z, hp[-j - 2], f);
            f = fmaf(x.w, hp[-j - 3], f);
        }
        float s = f, c = 0.0f;
        auto kahan = [&](float x, float h) {
            const float y = fmaf(x, h, -c);
            const float t = s + y;
            c = (t - s) - y;
            s = t;
        };
        for (int j = b0; j < b1; j += 4) {
            const float4 x = xs4[j >> 2];
            kahan(x.x, hp[-j]);
            kahan(x.y, hp[-j - 1]);
            kahan(x.z, hp[-j - 2]);
            kahan(x.w, hp[-j - 3]);
        }
        for (int j = nu4; j < nu; ++j) kahan(xs[j], hp[-j]);
        if (act) out_vu[row * nu + iu] = s;
    }
}

} // namespace

void ramp_taps_host(std::vector<double>& out, int nu, double da, int filterType, bool equalSpace)
{
    const int L = 2 * nu - 1;
    const double PIf = (double)3.141592657f;   // Projector.h:5
    std::vector<double> h(L);
    for (int i = 0; i < L; i++) {
        const int k = i - (nu - 1);
        if (k == 0) h[i] = 1.0 / (4.0 * da * da);
        else if (k % 2 != 0) {
            if (equalSpace) h[i] = -1.0 / (PIf * PIf * (double)k * k * da * da);
            else { const double s = std::sin(k * da); h[i] = -1.0 / (PIf * PIf * s * s); }
        } else h[i] = 0.0;
    }
    if (filterType < 1 || filterType > 3) { out = h; return; }
    // window applied in the DFT-index order of a length-L transform (Filter.cu:139-178)
    std::vector<double> win(L), cs(L), sn(L);
    for (int i = 0; i < L; i++) {
        const double t = (double)i / (double)(L - 1);
        win[i] = filterType == 1 ? 0.54 + 0.46 * std::cos(2 * PIf * t)
               : filterType == 2 ? 0.5 + 0.5 * std::cos(2 * PIf * t)
               : std::fabs(std::cos(PIf * t));
        cs[i] = std::cos(2 * M_PI * i / L);
        sn[i] = std::sin(2 * M_PI * i / L);
    }
    std::vector<double> Hr(L), Hi(L);
    for (int f = 0; f < L; f++) {
        double re = 0, im = 0;
        for (int n = 0; n < L; n++) {
            const int m = (int)(((long)f * n) % L);
            re += h[n] * cs[m];
            im -= h[n] * sn[m];
        }
        Hr[f] = re * win[f];
        Hi[f] = im * win[f];
    }
    out.assign(L, 0.0);
    for (int n = 0; n < L; n++) {
        double re = 0;
        for (int f = 0; f < L; f++) {
            const int m = (int)(((long)f * n) % L);
            re += Hr[f] * cs[m] - Hi[f] * sn[m];
        }
        out[n] = re / L;
    }
}

void RampFilter::setup(int nu, int nview, float da, float off_a, float dso, int filterType, bool fan, hipStream_t st)
{
    RN_REQUIRE(nu > 0 && nview > 0, "filter dimensions must be positive");
    const bool same = ready_ && nu == nu_ && nview == nview_ && da == da_ && off_a == off_a_ && dso == dso_ &&
                      filterType == type_ && fan == fan_;
    if (same) return;
    nu_ = nu; nview_ = nview; da_ = da; off_a_ = off_a; dso_ = dso; type_ = filterType; fan_ = fan;
    std::vector<double> h;
    ramp_taps_host(h, nu, (double)da, filterType, !fan);
    // post scale of Filter.cu:301 (fan, equiangular) / :393 (parallel); the 1/L of
    // the reference cancels the L of its unnormalised C2R
    const double PIf = (double)3.141592657f;
    const double scale = fan ? PIf / nview * (double)da / (double)dso : PIf / nview * (double)da;
    host_.assign(3 * (size_t)nu - 1, 0.0f);
    for (int i = 0; i < 2 * nu - 1; i++) host_[i] = (float)(h[i] * scale);
    for (int i = 0; i < nu; i++)
        host_[2 * nu - 1 + i] = fan ? (float)std::cos(((i - (nu - 1) / 2.0) - (double)off_a) * (double)da) : 1.0f;
    buf_.alloc(host_.size() * sizeof(float));
    RN_HIP_CHECK(hipMemcpyAsync(buf_.get(), host_.data(), host_.size() * sizeof(float), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    ready_ = true;
}

void RampFilter::apply_vu(const float* in_vu, float* out_vu, long nrows, hipStream_t st) const
{
    if (nrows <= 0) return;
    const size_t lds = (3 * (size_t)nu_ - 1) * sizeof(float);
    RN_REQUIRE(lds <= 160 * 1024, "nu too large for the LDS-resident ramp filter");
    hipLaunchKernelGGL(k_ramp_filter, dim3((unsigned)nrows), dim3(256), lds, st, in_vu, out_vu,
                       buf_.as<float>(), buf_.as<float>() + 2 * nu_ - 1, nu_);
    check_launch("ramp_filter");
}

} // namespace rn
