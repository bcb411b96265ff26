// Isotropic TV separable surrogate of the PriorFunctionSolver library
// (PriorFunctionSolver/TVSQS.cu:8-344): per voxel
//   var = sqrt(dx^2 + dy^2 (+ dz^2) + eps)   with backward differences (0 at the low edge),
//   s1  = (dx0 + dy0 (+ dz0)) / var + sum_a d_a1 / var_a,
//   s2  = D / var + sum_a 1 / var_a           (D = 2 or 3 dimensions),
// where d_a1 = x - x[+a] (0 at the high edge) and var_a = var[+a] (var itself at
// the high edge).  s1 is the gradient of sum(var); s2 the surrogate curvature.
// Layout [nb][nx][ny][(nz)][nc], channel innermost, as the reference.
//
// Two passes, each one thread per voxel with every (batch, channel) in the
// grid: var first, then s1 / s2 from var and the image.

#include "rn_prior.h"

namespace rn {

namespace {

struct Dims {
    int nx, ny, nz, nc, nb;
    long sx, sy, sz;   // element strides of x, y, z (z stride = nc)
};

__device__ __forceinline__ bool decode(const Dims& d, long i, int& ix, int& iy, int& iz, long& base)
{
    // i runs over (b, ix, iy, iz, c); base is the element index of the voxel
    const long per = (long)d.nx * d.ny * d.nz * d.nc;
    if (i >= per * d.nb) return false;
    long r = i % per;
    iz = (int)((r / d.nc) % d.nz);
    iy = (int)((r / ((long)d.nc * d.nz)) % d.ny);
    ix = (int)(r / ((long)d.nc * d.nz * d.ny));
    base = i;
    return true;
}

template <int D>
__global__ __launch_bounds__(256) void k_tv_iso_var(Dims d, const float* __restrict__ img, float* __restrict__ var,
                                                    float eps)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int ix, iy, iz;
    long p;
    if (!decode(d, i, ix, iy, iz, p)) return;
    const float x = img[p];
    const float dx = ix > 0 ? x - img[p - d.sx] : 0.0f;
    const float dy = iy > 0 ? x - img[p - d.sy] : 0.0f;
    float s = fmaf(dx, dx, dy * dy);
    if (D == 3) {
        const float dz = iz > 0 ? x - img[p - d.sz] : 0.0f;
        s = fmaf(dz, dz, s);
    }
    var[p] = sqrtf(s + eps);
}

template <int D>
__global__ __launch_bounds__(256) void k_tv_iso_sqs(Dims d, const float* __restrict__ img,
                                                    const float* __restrict__ var, float* __restrict__ s1,
                                                    float* __restrict__ s2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int ix, iy, iz;
    long p;
    if (!decode(d, i, ix, iy, iz, p)) return;
    const float x = img[p], v = var[p];
    float d0 = 0.0f, g1 = 0.0f, c1 = 0.0f;
    auto axis = [&](int idx, int n, long st) {
        if (idx > 0) d0 += x - img[p - st];
        const bool up = idx < n - 1;
        const float va = up ? var[p + st] : v;
        if (up) g1 += (x - img[p + st]) / va;
        c1 += 1.0f / va;
    };
    axis(ix, d.nx, d.sx);
    axis(iy, d.ny, d.sy);
    if (D == 3) axis(iz, d.nz, d.sz);
    s1[p] = d0 / v + g1;
    s2[p] = (float)D / v + c1;
}

} // namespace

void tv_iso_sqs(const float* img, float* s1, float* s2, float* var, int nb, int nx, int ny, int nz, int nc, int dims,
                float eps, hipStream_t st)
{
    RN_REQUIRE(dims == 2 || dims == 3, "isotropic TV-SQS is 2D or 3D");
    RN_REQUIRE(img && s1 && s2 && var, "null pointer");
    if (dims == 2) nz = 1;
    RN_REQUIRE(nb > 0 && nx > 0 && ny > 0 && nz > 0 && nc > 0, "dimensions must be positive");
    Dims d{nx, ny, nz, nc, nb, (long)ny * nz * nc, (long)nz * nc, (long)nc};
    const long n = (long)nb * nx * ny * nz * nc;
    const dim3 blk(256), grd((unsigned)cdiv(n, 256));
    if (dims == 2) {
        hipLaunchKernelGGL(k_tv_iso_var<2>, grd, blk, 0, st, d, img, var, eps);
        check_launch("tv_iso_var");
        hipLaunchKernelGGL(k_tv_iso_sqs<2>, grd, blk, 0, st, d, img, var, s1, s2);
    } else {
        hipLaunchKernelGGL(k_tv_iso_var<3>, grd, blk, 0, st, d, img, var, eps);
        check_launch("tv_iso_var");
        hipLaunchKernelGGL(k_tv_iso_sqs<3>, grd, blk, 0, st, d, img, var, s1, s2);
    }
    check_launch("tv_iso_sqs");
}

} // namespace rn
