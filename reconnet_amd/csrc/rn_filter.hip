// Ramp filtering of fan / parallel sinograms (Filter.cu:66-518).
//
// The reference zero-pads each detector row to L = 2nu-1 samples, multiplies
// by the cosine pre-weight, and applies the windowed ramp through cuFFT
// (R2C, spectrum product, C2R).  Because of the padding that product is an
// exact linear convolution with the (2nu-1)-tap kernel, so here it is computed
// directly: a workgroup walks detector rows with the taps staged in LDS once,
// each lane computes 4 adjacent outputs of a row from register windows of
// taps (one 16-byte LDS read per 16 FMAs), FP32 accumulation with Kahan
// compensation over the large near taps.  The taps themselves are
// built on the host in FP64 (closed-form Ram-Lak taps, window applied through
// an exact FP64 DFT round trip, Filter.cu:90-193).

#include "rn_filter.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace rn {

namespace {

// Near band half width (taps |k| <= kNear summed with Kahan compensation last;
// a multiple of 4).  The ramp taps fall off as 1/k^2: the band holds the large
// taps whose products cancel, the far taps are small enough for plain FMAs.
constexpr int kNear = 32;

// Tap and sample tables, all padded to multiples of 4 samples (P = nu rounded up):
//   hf[2P + 8]       far taps, hf[d + P + 4] = h[d] for kNear < |d| <= nu - 1, else 0
//   hr[2kNear + 4]   near taps reversed, hr[n] = h[kNear - n] for n <= 2kNear, else 0
//   xs[P + 2kNear + 8] weighted samples, xs[j + kNear] = x[j], zero outside [0, nu)
inline int ramp_pad(int nu) { return (nu + 3) & ~3; }

// R rows per step of a workgroup: a lane computes 4 adjacent outputs of each of
// the R rows, so that one 16-byte tap read from LDS serves 16 R FMAs (the
// samples of a row are wave-uniform broadcast reads).
template <int R>
__global__ __launch_bounds__(256) void k_ramp_filter(const float* __restrict__ in_vu, float* __restrict__ out_vu,
                                                     const float* __restrict__ tabs, const float* __restrict__ pw,
                                                     int nu, long nrows)
{
    extern __shared__ float4 sm4[];
    const int P = (nu + 3) & ~3;
    const int nhf = 2 * P + 8, nhr = 2 * kNear + 4, nxs = P + 2 * kNear + 8;
    float* sm = reinterpret_cast<float*>(sm4);
    float* hf = sm;
    float* hr = hf + nhf;
    float* xs = hr + nhr;    // R rows of nxs samples
    const int tid = threadIdx.x, bd = blockDim.x, lane = tid & 63;
    for (int j = tid; j < nhf + nhr; j += bd) sm[j] = tabs[j];
    for (int j = tid; j < R * nxs; j += bd) xs[j] = 0.0f;
    for (long row0 = (long)blockIdx.x * R; row0 < nrows; row0 += (long)gridDim.x * R) {
        __syncthreads();    // tables written / the previous rows' readers done
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const float* src = in_vu + min(row0 + rr, nrows - 1) * nu;   // a missing last row repeats, unstored
            for (int j = tid; j < nu; j += bd) xs[rr * nxs + kNear + j] = src[j] * pw[j];
        }
        __syncthreads();
        // lane t computes out[4t .. 4t+3] of each row; the loop bound is the wave's first output.
        // Lanes past the row end compute the last group again (i0 clamped, keeping every table
        // read in range) and store nothing.
        for (int t = tid; 4 * (t - lane) < nu; t += bd) {
            const int it = 4 * t;
            const int i0 = min(it, P - 4);
            // far part: out_r = sum_j x[j] hf[i0 + r - j] in increasing j; the taps of
            // offsets b-4 .. b+3 (b = i0 - j) sit in two aligned float4 registers A, B
            float o[R][4];
#pragma unroll
            for (int rr = 0; rr < R; ++rr) o[rr][0] = o[rr][1] = o[rr][2] = o[rr][3] = 0.0f;
            float4 B = *reinterpret_cast<const float4*>(hf + i0 + P + 4);
            for (int j = 0; j < P; j += 4) {
                const float4 A = *reinterpret_cast<const float4*>(hf + i0 - j + P);
#pragma unroll
                for (int rr = 0; rr < R; ++rr) {
                    const float4 X = *reinterpret_cast<const float4*>(xs + rr * nxs + kNear + j);   // broadcast
                    float* q = o[rr];
                    q[0] = fmaf(X.x, B.x, q[0]); q[0] = fmaf(X.y, A.w, q[0]); q[0] = fmaf(X.z, A.z, q[0]); q[0] = fmaf(X.w, A.y, q[0]);
                    q[1] = fmaf(X.x, B.y, q[1]); q[1] = fmaf(X.y, B.x, q[1]); q[1] = fmaf(X.z, A.w, q[1]); q[1] = fmaf(X.w, A.z, q[1]);
                    q[2] = fmaf(X.x, B.z, q[2]); q[2] = fmaf(X.y, B.y, q[2]); q[2] = fmaf(X.z, B.x, q[2]); q[2] = fmaf(X.w, A.w, q[2]);
                    q[3] = fmaf(X.x, B.w, q[3]); q[3] = fmaf(X.y, B.z, q[3]); q[3] = fmaf(X.z, B.y, q[3]); q[3] = fmaf(X.w, B.x, q[3]);
                }
                B = A;
            }
            // near part with Kahan compensation: out_r += x[i0 + r - kNear + n] h[kNear - n],
            // n = 0 .. 2kNear, samples from two aligned float4 windows of the lane's own row part
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const float* xr = xs + rr * nxs;
                float* sr = o[rr];
                float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                auto kahan = [&](int r, float x, float h) {
                    const float y = fmaf(x, h, -c[r]);
                    const float tt = sr[r] + y;
                    c[r] = (tt - sr[r]) - y;
                    sr[r] = tt;
                };
                float4 C = *reinterpret_cast<const float4*>(xr + i0);
                float4 D = C;
#pragma unroll
                for (int p = 0; p < kNear / 2; ++p) {
                    D = *reinterpret_cast<const float4*>(xr + i0 + 4 * p + 4);
                    const float4 H = *reinterpret_cast<const float4*>(hr + 4 * p);           // broadcast
                    const float w[8] = {C.x, C.y, C.z, C.w, D.x, D.y, D.z, D.w};
                    const float hq[4] = {H.x, H.y, H.z, H.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int r = 0; r < 4; ++r) kahan(r, w[q + r], hq[q]);
                    C = D;
                }
                const float hl = hr[2 * kNear];
                kahan(0, D.x, hl); kahan(1, D.y, hl); kahan(2, D.z, hl); kahan(3, D.w, hl);
                if (it < nu && row0 + rr < nrows) {
                    float* dst = out_vu + (row0 + rr) * nu;
                    if ((nu & 3) == 0 && (reinterpret_cast<uintptr_t>(out_vu) & 15) == 0) {
                        *reinterpret_cast<float4*>(dst + it) = make_float4(sr[0], sr[1], sr[2], sr[3]);
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (it + r < nu) dst[it + r] = sr[r];
                    }
                }
            }
        }
    }
}

constexpr int kRampRows = 4;

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
    const int P = ramp_pad(nu);
    const size_t nhf = 2 * (size_t)P + 8, nhr = 2 * kNear + 4;
    host_.assign(nhf + nhr + nu, 0.0f);
    for (int d = -(nu - 1); d <= nu - 1; d++)
        if (std::abs(d) > kNear) host_[d + P + 4] = (float)(h[d + nu - 1] * scale);
    for (int n = 0; n <= 2 * kNear; n++) {
        const int k = kNear - n;
        if (std::abs(k) <= nu - 1) host_[nhf + n] = (float)(h[k + nu - 1] * scale);
    }
    for (int i = 0; i < nu; i++)
        host_[nhf + nhr + i] = fan ? (float)std::cos(((i - (nu - 1) / 2.0) - (double)off_a) * (double)da) : 1.0f;
    buf_.alloc(host_.size() * sizeof(float));
    RN_HIP_CHECK(hipMemcpyAsync(buf_.get(), host_.data(), host_.size() * sizeof(float), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    ready_ = true;
}

void RampFilter::apply_vu(const float* in_vu, float* out_vu, long nrows, hipStream_t st) const
{
    if (nrows <= 0) return;
    const int P = ramp_pad(nu_);
    const size_t ntab = 2 * (size_t)P + 8 + 2 * kNear + 4;
    const size_t lds = (ntab + (size_t)kRampRows * (P + 2 * kNear + 8)) * sizeof(float);
    RN_REQUIRE(lds <= 64 * 1024, "nu too large for the LDS-resident ramp filter");
    // lanes per row: one per 4 outputs, whole waves, at most 4 waves (larger rows loop)
    const int threads = (int)std::min<long>(256, (((nu_ + 3) / 4 + 63) / 64) * 64);
    // one resident round of workgroups (a second, partial round would idle most CUs)
    int per_cu = 0, cus = 0, dev = 0;
    RN_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ramp_filter<kRampRows>, threads, lds));
    RN_HIP_CHECK(hipGetDevice(&dev));
    RN_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const long steps = (nrows + kRampRows - 1) / kRampRows;
    const unsigned grid = (unsigned)std::max<long>(1, std::min<long>(steps, (long)std::max(per_cu, 1) * cus));
    hipLaunchKernelGGL(k_ramp_filter<kRampRows>, dim3(grid), dim3(threads), lds, st, in_vu, out_vu,
                       buf_.as<float>(), buf_.as<float>() + ntab, nu_, nrows);
    check_launch("ramp_filter");
}

} // namespace rn
