// libPriorFunctionSolver.so: the TV-SQS entry points of the reference's
// PriorFunctionSolver library (PriorFunctionSolver/TVSQS.cu:108,290 and the
// device selector it exports), with the same signatures and host-pointer
// buffers, so that python/PriorFunctionSolver.py:161-205 can load this build.
// The OMP dictionary and nuclear-norm solvers of that library are not part of
// this build.  Failures print "<function> failed: <reason>" and do not throw.

#include "priorfunctionsolver.h"
#include "rn_prior.h"

using namespace rn;

namespace {

void tv_host(float* s1, float* s2, float* var, const float* img, int nb, int nx, int ny, int nz, int nc, int dims,
             float eps)
{
    RN_REQUIRE(s1 && s2 && var && img, "null pointer");
    RN_REQUIRE(nb > 0 && nx > 0 && ny > 0 && nz > 0 && nc > 0, "dimensions must be positive");
    const size_t n = (size_t)nb * nx * ny * nz * nc;
    CallStream s;
    DevBuf d(4 * n * sizeof(float));
    float* di = d.as<float>();
    RN_HIP_CHECK(hipMemcpyAsync(di, img, n * sizeof(float), hipMemcpyHostToDevice, s.get()));
    tv_iso_sqs(di, di + n, di + 2 * n, di + 3 * n, nb, nx, ny, nz, nc, dims, eps, s.get());
    RN_HIP_CHECK(hipMemcpyAsync(s1, di + n, n * sizeof(float), hipMemcpyDeviceToHost, s.get()));
    RN_HIP_CHECK(hipMemcpyAsync(s2, di + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, s.get()));
    RN_HIP_CHECK(hipMemcpyAsync(var, di + 3 * n, n * sizeof(float), hipMemcpyDeviceToHost, s.get()));
    s.sync();
}

} // namespace

extern "C" int cSetDevice(int device)
{
    int r = 0;
    guarded("cSetDevice", [&] { RN_HIP_CHECK(hipSetDevice(device)); r = device; });
    return r;
}

extern "C" void cTVSQS2D(float* s1, float* s2, float* var, const float* img, int nb, int nx, int ny, int nc, float eps)
{
    guarded("cTVSQS2D", [&] { tv_host(s1, s2, var, img, nb, nx, ny, 1, nc, 2, eps); });
}

extern "C" void cTVSQS3D(float* s1, float* s2, float* var, const float* img, int nb, int nx, int ny, int nz, int nc,
                         float eps)
{
    guarded("cTVSQS3D", [&] { tv_host(s1, s2, var, img, nb, nx, ny, nz, nc, 3, eps); });
}

extern "C" const char* pfs_last_error(void) { return last_error(); }
