// extern "C" entry points of libReconNet4D_no_tf.so.
//
// Reference exports (host pointers, synchronous) keep the reference's
// argument lists and meaning; each one copies its inputs to the device,
// runs the device-pointer implementation on a per-thread stream, and copies
// the result back.  Geometry tables and workspaces are cached per host thread
// (keyed by device, geometry and angle values), so repeated calls with the same
// geometry neither rebuild tables nor reallocate.  Nothing throws across the
// ABI: failures print "<function> failed: <reason>" and set rn_last_error().

#include "reconnet.h"
#include "rn_3d.h"
#include "rn_fan_ops.h"
#include "rn_filter.h"
#include "rn_prior.h"
#include "rn_sqs.h"
#include "rn_tune.h"

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using namespace rn;

namespace {

struct ThreadCtx {
    int device = -1;
    std::unique_ptr<CallStream> stream;
    FanTables tables;
    FanWork work;
    RampFilter filter;
    DevBuf a, b, c, d, e, f, g;
    hipStream_t st() { return stream->get(); }
};

ThreadCtx& ctx()
{
    thread_local ThreadCtx c;
    int dev = 0;
    RN_HIP_CHECK(hipGetDevice(&dev));
    if (!c.stream || c.device != dev) {
        c = ThreadCtx();
        c.device = dev;
        c.stream.reset(new CallStream());
    }
    return c;
}

void h2d(DevBuf& buf, const float* host, size_t n, hipStream_t st)
{
    RN_REQUIRE(host != nullptr, "null input pointer");
    buf.ensure(n * sizeof(float));
    RN_HIP_CHECK(hipMemcpyAsync(buf.get(), host, n * sizeof(float), hipMemcpyHostToDevice, st));
}

void d2h(float* host, const DevBuf& buf, size_t n, hipStream_t st)
{
    RN_REQUIRE(host != nullptr, "null output pointer");
    RN_HIP_CHECK(hipMemcpyAsync(host, buf.get(), n * sizeof(float), hipMemcpyDeviceToHost, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

FanGeo make_fan(int nx, int ny, int nz, float dx, float dy, float dz, int nu, int nview, int nv,
                float da, float dv, float off_a, float off_v, float dsd, float dso)
{
    FanGeo g{nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso};
    RN_REQUIRE(nx > 0 && ny > 0 && nz > 0, "image dimensions must be positive");
    RN_REQUIRE(nu > 0 && nview > 0 && nv > 0, "sinogram dimensions must be positive");
    return g;
}

void check_bc(int nB, int nC)
{
    RN_REQUIRE(nB > 0 && nC > 0, "nBatches and nChannels must be positive");
}

void fan_fp_host(FanFP kind, float* prj, const float* img, const float* deg, int nB, int nC, const FanGeo& g)
{
    check_bc(nB, nC);
    RN_REQUIRE(deg != nullptr, "null angle pointer");
    ThreadCtx& c = ctx();
    if (!c.tables.matches(g, deg)) c.tables.build(g, deg, c.st());
    const size_t nimg = (size_t)nB * g.nx * g.ny * g.nz * nC, nprj = (size_t)nB * g.nu * g.nview * g.nv * nC;
    h2d(c.a, img, nimg, c.st());
    c.b.ensure(nprj * sizeof(float));
    fan_forward(c.tables, kind, nB, nC, c.a.as<float>(), c.b.as<float>(), c.work, c.st());
    d2h(prj, c.b, nprj, c.st());
}

void fan_bp_host(FanBP kind, float* img, const float* prj, const float* deg, int nB, int nC, const FanGeo& g)
{
    check_bc(nB, nC);
    RN_REQUIRE(deg != nullptr, "null angle pointer");
    ThreadCtx& c = ctx();
    if (!c.tables.matches(g, deg)) c.tables.build(g, deg, c.st());
    const size_t nimg = (size_t)nB * g.nx * g.ny * g.nz * nC, nprj = (size_t)nB * g.nu * g.nview * g.nv * nC;
    h2d(c.a, prj, nprj, c.st());
    c.b.ensure(nimg * sizeof(float));
    fan_backward(c.tables, kind, nB, nC, c.a.as<float>(), c.b.as<float>(), c.work, c.st());
    d2h(img, c.b, nimg, c.st());
}

void filter_device(RampFilter& filt, const float* prj, float* fprj, int nB, int nC, int nu, int nview, int nv,
                   float da, float off_a, float dso, int typeFilter, bool fan, DevBuf& t1, DevBuf& t2, hipStream_t st)
{
    check_bc(nB, nC);
    RN_REQUIRE(nu > 0 && nview > 0 && nv > 0, "sinogram dimensions must be positive");
    filt.setup(nu, nview, da, off_a, dso, typeFilter, fan, st);
    const int ns = nB * nv * nC;
    const size_t n = (size_t)ns * nu * nview;
    const bool direct = (nv * nC == 1);
    t1.ensure(n * sizeof(float));
    t2.ensure(n * sizeof(float));
    std::vector<float> I;
    DevBuf mix;
    const float* Q = prj;
    if (!direct) {
        I.assign((size_t)nv * nv, 0.0f);
        for (int i = 0; i < nv; i++) I[(size_t)i * nv + i] = 1.0f;
        mix.alloc(I.size() * sizeof(float));
        RN_HIP_CHECK(hipMemcpyAsync(mix.get(), I.data(), I.size() * sizeof(float), hipMemcpyHostToDevice, st));
        launch_mix_gather(prj, t2.as<float>(), mix.as<float>(), nB, nu * nview, nv, nC, nv, st);
        Q = t2.as<float>();
    }
    launch_transpose(Q, t1.as<float>(), nu, nview, ns, st);          // [s][nview][nu]
    filt.apply_vu(t1.as<float>(), t2.as<float>(), (long)ns * nview, st);
    if (direct) {
        launch_transpose(t2.as<float>(), fprj, nview, nu, ns, st);
    } else {
        launch_transpose(t2.as<float>(), t1.as<float>(), nview, nu, ns, st);
        launch_mix_scatter(t1.as<float>(), fprj, mix.as<float>(), nB, nu * nview, nv, nC, nv, false, st);
    }
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

void filter_host(float* out, const float* in, int nB, int nC, int nu, int nview, int nv, float da, float off_a,
                 float dso, int typeFilter, bool fan)
{
    ThreadCtx& c = ctx();
    const size_t n = (size_t)nB * nu * nview * nv * nC;
    check_bc(nB, nC);
    RN_REQUIRE(nu > 0 && nview > 0 && nv > 0, "sinogram dimensions must be positive");
    h2d(c.a, in, n, c.st());
    c.b.ensure(n * sizeof(float));
    filter_device(c.filter, c.a.as<float>(), c.b.as<float>(), nB, nC, nu, nview, nv, da, off_a, dso, typeFilter, fan,
                  c.c, c.d, c.st());
    d2h(out, c.b, n, c.st());
}

struct Plan {
    int device = 0;
    FanTables tables;
    FanWork work;
    RampFilter filter;
    DevBuf t1, t2;
};

struct Sqs {
    Plan* plan = nullptr;
    OsSqs engine;
};

} // namespace

// =====================================================================
// Reference exports
// =====================================================================

extern "C" int SetDevice(int device)
{
    int rc = -1;
    guarded("SetDevice", [&] {
        hipError_t e = hipSetDevice(device);
        rc = (int)e;
        if (e != hipSuccess) throw Error(std::string("hipSetDevice: ") + hipGetErrorString(e));
    });
    return rc;
}

extern "C" void cDistanceDrivenFanProjection(float* prj, const float* img, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector)
{
    (void)typeProjector;
    guarded("cDistanceDrivenFanProjection", [&] {
        fan_fp_host(FanFP::DD, prj, img, deg, nBatches, nChannels,
                    make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso));
    });
}

extern "C" void cDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector)
{
    guarded("cDistanceDrivenFanBackprojection", [&] {
        fan_bp_host(typeProjector == 1 ? FanBP::DDFBP : FanBP::DD, img, prj, deg, nBatches, nChannels,
                    make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso));
    });
}

extern "C" void cSiddonFanProjection(float* prj, const float* img, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector)
{
    (void)typeProjector;
    guarded("cSiddonFanProjection", [&] {
        fan_fp_host(FanFP::Siddon, prj, img, deg, nBatches, nChannels,
                    make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso));
    });
}

extern "C" void cSiddonFanBackprojection(float* img, const float* prj, const float* deg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector)
{
    (void)typeProjector;
    guarded("cSiddonFanBackprojection", [&] {
        fan_bp_host(FanBP::Siddon, img, prj, deg, nBatches, nChannels,
                    make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso));
    });
}

extern "C" void cPixelDrivenFanBackprojection(float* pImg, const float* pPrj, const float* pDeg,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector)
{
    (void)typeProjector;
    guarded("cPixelDrivenFanBackprojection", [&] {
        fan_bp_host(FanBP::Pixel, pImg, pPrj, pDeg, nBatches, nChannels,
                    make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso));
    });
}

extern "C" void cFilterFanFilter(float* pFPrj, const float* pPrj,
        int nBatches, int nChannels, int nu, int nview,
        int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeFilter, int typeProjector)
{
    (void)dv; (void)off_v; (void)dsd; (void)typeProjector;
    guarded("cFilterFanFilter", [&] {
        filter_host(pFPrj, pPrj, nBatches, nChannels, nu, nview, nv, da, off_a, dso, typeFilter, true);
    });
}

extern "C" void cFilterParallelFilter(float* pFPrj, const float* pPrj,
        int nBatches, int nChannels, int nu, int nview,
        int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeFilter, int typeProjector)
{
    (void)dv; (void)off_v; (void)dsd; (void)typeProjector;
    guarded("cFilterParallelFilter", [&] {
        filter_host(pFPrj, pPrj, nBatches, nChannels, nu, nview, nv, da, off_a, dso, typeFilter, false);
    });
}

extern "C" float cTVSQS2D(float* grad, float* norm, const float* img, int nx, int ny)
{
    float tv = 0.0f;
    guarded("cTVSQS2D", [&] {
        RN_REQUIRE(nx > 0 && ny > 0, "image dimensions must be positive");
        ThreadCtx& c = ctx();
        const size_t n = (size_t)nx * ny;
        h2d(c.a, img, n, c.st());
        c.b.ensure(n * sizeof(float));
        c.c.ensure(n * sizeof(float));
        std::vector<double> tvs;
        tv_sqs(c.a.as<float>(), c.b.as<float>(), c.c.as<float>(), nx, ny, 1, 1e-6f, &tvs, c.d, c.st());
        d2h(grad, c.b, n, c.st());
        d2h(norm, c.c, n, c.st());
        tv = (float)tvs[0];
    });
    return tv;
}

extern "C" void cOSSQS(float* res, const float* img, const float* prj,
        const float* normImg, const float* wprj,
        const float* degs, float lb, float ub, int nviewPerSubset, int nSubsets,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, int typeProjector, float lambda)
{
    (void)typeProjector;
    guarded("cOSSQS", [&] {
        check_bc(nBatches, nChannels);
        RN_REQUIRE(degs != nullptr && wprj != nullptr && normImg != nullptr, "null input pointer");
        FanGeo g = make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso);
        ThreadCtx& c = ctx();
        if (nz != 1 || nv != 1 || nChannels != 1) {
            // several z planes, detector rows or channels: the pass on the
            // reference layouts with the z/v mixing of the Siddon fan pair
            const size_t nimg = (size_t)nBatches * nx * ny * nz * nChannels;
            const size_t nprj = (size_t)nBatches * nu * nview * nv * nChannels;
            h2d(c.a, prj, nprj, c.st());
            h2d(c.b, wprj, nprj, c.st());
            h2d(c.c, normImg, nimg, c.st());
            h2d(c.d, img, nimg, c.st());
            os_sqs_ref_layout(g, degs, nBatches, nChannels, nviewPerSubset, nSubsets, c.d.as<float>(),
                              c.a.as<float>(), c.b.as<float>(), c.c.as<float>(), lambda, lb, ub, c.st());
            d2h(res, c.d, nimg, c.st());
            return;
        }
        const size_t nimg = (size_t)nBatches * nx * ny, nprj = (size_t)nBatches * nu * nview;
        OsSqs eng;
        eng.setup(g, degs, nBatches, nviewPerSubset, nSubsets, c.st());
        h2d(c.a, prj, nprj, c.st());
        h2d(c.b, wprj, nprj, c.st());
        eng.set_data(c.a.as<float>(), c.b.as<float>(), c.st());
        h2d(c.c, normImg, nimg, c.st());
        eng.set_norm(c.c.as<float>(), c.st());
        h2d(c.d, img, nimg, c.st());
        eng.set_image(c.d.as<float>(), c.st());
        SqsStep step;
        step.lambda = lambda;
        step.lb = lb;
        step.ub = ub;
        eng.pass(step, c.st());
        eng.get_image(c.d.as<float>(), c.st());
        d2h(res, c.d, nimg, c.st());
    });
}

// ----- 3D / arbitrary-geometry exports -----

namespace {
ArbGeo make_arb(int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
                int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    RN_REQUIRE(nx > 0 && ny > 0 && nz > 0, "image dimensions must be positive");
    RN_REQUIRE(nu > 0 && nview > 0 && nv > 0, "projection dimensions must be positive");
    return ArbGeo{nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v};
}

void arb_call(const char* name, bool forward, int kind, float* out, const float* in,
              const float* v0, const float* v1, const float* v2, const float* v3, int nB, int nC, const ArbGeo& g)
{
    (void)name;
    check_bc(nB, nC);
    ThreadCtx& c = ctx();
    const size_t nimg = (size_t)nB * g.nx * g.ny * g.nz * nC, nprj = (size_t)nB * g.nu * g.nview * g.nv * nC;
    const size_t nvec = 3 * (size_t)g.nview;
    h2d(c.a, in, forward ? nimg : nprj, c.st());
    c.b.ensure((forward ? nprj : nimg) * sizeof(float));
    const float* d0 = nullptr; const float* d1 = nullptr; const float* d2 = nullptr; const float* d3 = nullptr;
    h2d(c.c, v0, nvec, c.st()); d0 = c.c.as<float>();
    h2d(c.d, v1, nvec, c.st()); d1 = c.d.as<float>();
    if (v2) { h2d(c.e, v2, nvec, c.st()); d2 = c.e.as<float>(); }
    if (v3) { h2d(c.f, v3, nvec, c.st()); d3 = c.f.as<float>(); }
    switch (kind) {
    case 0: // Siddon parallel: v0 = detCenter, v1 = detU, v2 = detV, v3 = invRayDir
        if (forward) siddon_arb_fp(g, true, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, d2, d3, c.g, c.st());
        else siddon_arb_bp(g, true, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, d2, d3, c.g, c.st());
        break;
    case 1: { // Siddon cone: v3 = src
        ArbGeo gz = g;
        gz.rows_z = arb_rows_along_z(v1, v2, g.nview) ? 1 : 0;
        if (forward) siddon_arb_fp(gz, false, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, d2, d3, c.g, c.st());
        else siddon_arb_bp(gz, false, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, d2, d3, c.g, c.st());
        break;
    }
    case 2: // DD tomo: v0 = detCenter, v1 = src
        if (forward) dd_tomo_fp(g, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, c.g, c.st());
        else dd_tomo_bp(g, nB, nC, c.a.as<float>(), c.b.as<float>(), d0, d1, c.g, c.st());
        break;
    }
    d2h(out, c.b, forward ? nprj : nimg, c.st());
}
} // namespace

extern "C" void cDistanceDrivenTomoProjection(float* prj, const float* img,
        const float* detCenter, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cDistanceDrivenTomoProjection", [&] {
        arb_call("", true, 2, prj, img, detCenter, src, nullptr, nullptr, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cDistanceDrivenTomoBackprojection(float* img, const float* prj,
        const float* detCenter, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cDistanceDrivenTomoBackprojection", [&] {
        arb_call("", false, 2, img, prj, detCenter, src, nullptr, nullptr, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cSiddonParallelProjection(float* prj, const float* img,
        const float* detCenter, const float* detU, const float* detV, const float* invRayDir,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cSiddonParallelProjection", [&] {
        arb_call("", true, 0, prj, img, detCenter, detU, detV, invRayDir, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cSiddonParallelBackprojection(float* img, const float* prj,
        const float* detCenter, const float* detU, const float* detV, const float* invRayDir,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cSiddonParallelBackprojection", [&] {
        arb_call("", false, 0, img, prj, detCenter, detU, detV, invRayDir, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cSiddonConeProjectionAbitrary(float* prj, const float* img,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cSiddonConeProjectionAbitrary", [&] {
        arb_call("", true, 1, prj, img, detCenter, detU, detV, src, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cSiddonConeBackprojectionAbitrary(float* img, const float* prj,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v)
{
    guarded("cSiddonConeBackprojectionAbitrary", [&] {
        arb_call("", false, 1, img, prj, detCenter, detU, detV, src, nBatches, nChannels,
                 make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v));
    });
}

extern "C" void cOSSQSConeAbitrary(float* res, const float* img, const float* prj,
        const float* normImg, const float* mask,
        const float* detCenter, const float* detU, const float* detV, const float* src,
        float lb, float ub,
        int nviewPerSubset, int nSubsets, int nBatches, int nChannels,
        int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy, float cz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        int typeProjector, float lambda)
{
    (void)typeProjector;
    guarded("cOSSQSConeAbitrary", [&] {
        check_bc(nBatches, nChannels);
        RN_REQUIRE(nviewPerSubset > 0 && nSubsets > 0, "subset sizes must be positive");
        ArbGeo g = make_arb(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v);
        g.rows_z = arb_rows_along_z(detU, detV, nview) ? 1 : 0;
        ThreadCtx& c = ctx();
        const size_t nimg = (size_t)nBatches * nx * ny * nz * nChannels;
        const size_t nprj = (size_t)nBatches * nu * nview * nv * nChannels, nvec = 3 * (size_t)nview;
        DevBuf x, b, D, m, C, U, V, S;
        h2d(x, img, nimg, c.st()); h2d(b, prj, nprj, c.st()); h2d(D, normImg, nimg, c.st());
        h2d(m, mask, nimg, c.st()); h2d(C, detCenter, nvec, c.st()); h2d(U, detU, nvec, c.st());
        h2d(V, detV, nvec, c.st()); h2d(S, src, nvec, c.st());
        os_sqs_cone(g, nBatches, nChannels, x.as<float>(), b.as<float>(), D.as<float>(), m.as<float>(),
                    C.as<float>(), U.as<float>(), V.as<float>(), S.as<float>(), lb, ub, nviewPerSubset, nSubsets,
                    lambda, c.g, c.st());
        d2h(res, x, nimg, c.st());
    });
}

extern "C" void cHelicalBPFromParallelConebeamBackprojection(float* pImg, const float* pPrj,
        int nBatches, int nChannels, int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float du, float dv, float off_u, float off_v, float dsd, float dso,
        int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q)
{
    guarded("cHelicalBPFromParallelConebeamBackprojection", [&] {
        check_bc(nBatches, nChannels);
        RN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nview > 0 && nv > 0, "dimensions must be positive");
        RN_REQUIRE(nviewPerPI > 0, "nviewPerPI must be positive");
        RN_REQUIRE(zrot != 0.0f && std::isfinite(zrot) && mPI >= 0, "zrot must be non-zero and mPI >= 0");
        ThreadCtx& c = ctx();
        const size_t nimg = (size_t)nBatches * nx * ny * nz * nChannels, nprj = (size_t)nBatches * nu * nview * nv * nChannels;
        h2d(c.a, pPrj, nprj, c.st());
        c.b.ensure(nimg * sizeof(float));
        HelicalGeo h{nx, ny, nz, dx, dy, dz, nu, nview, nv, du, dv, off_u, off_v, dsd, dso, nviewPerPI, theta0, volZ0,
                     zrot, mPI, Q};
        helical_bp(h, nBatches, nChannels, c.a.as<float>(), c.b.as<float>(), c.g, c.st());
        d2h(pImg, c.b, nimg, c.st());
    });
}

// =====================================================================
// Additive API
// =====================================================================

extern "C" int rn_helical_bp(const float* prj, float* img, int nBatches, int nChannels, int nx, int ny, int nz,
        float dx, float dy, float dz, int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        float dsd, float dso, int nviewPerPI, float theta0, float volZ0, float zrot, int mPI, float Q, void* stream)
{
    return guarded_status("rn_helical_bp", [&] {
        RN_REQUIRE(prj && img, "null pointer");
        check_bc(nBatches, nChannels);
        RN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nview > 0 && nv > 0, "dimensions must be positive");
        RN_REQUIRE(nviewPerPI > 0, "nviewPerPI must be positive");
        RN_REQUIRE(zrot != 0.0f && std::isfinite(zrot) && mPI >= 0, "zrot must be non-zero and mPI >= 0");
        HelicalGeo h{nx, ny, nz, dx, dy, dz, nu, nview, nv, du, dv, off_u, off_v, dsd, dso, nviewPerPI, theta0, volZ0,
                     zrot, mPI, Q};
        helical_bp(h, nBatches, nChannels, prj, img, ctx().g, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_helical_rebin(const float* fan, float* par, int nBatches, int nChannels,
        int nu_f, int nview_f, int nv_f, float du_f, float dv_f, float off_u_f, float off_v_f,
        float beta0, float dbeta,
        int nu_p, int nview_p, int nv_p, float du_p, float dv_p, float off_u_p, float off_v_p,
        float theta0, float dtheta, float dsd, float dso, void* stream)
{
    return guarded_status("rn_helical_rebin", [&] {
        RN_REQUIRE(fan && par, "null pointer");
        RN_REQUIRE(dbeta != 0.0f && dtheta != 0.0f, "angular steps must be non-zero");
        RebinGeo r{nu_f, nview_f, nv_f, du_f, dv_f, off_u_f, off_v_f, nu_p, nview_p, nv_p, du_p, dv_p, off_u_p,
                   off_v_p, dso, dsd, (double)dtheta / dbeta, ((double)theta0 - beta0) / dbeta, 1.0 / dbeta};
        helical_rebin(r, nBatches, nChannels, fan, par, static_cast<hipStream_t>(stream));
    });
}

extern "C" const char* rn_last_error(void) { return last_error(); }

extern "C" int rn_autotune_enable(int on)
{
    return guarded_status("rn_autotune_enable", [&] { tune_set_enabled(on != 0); });
}

extern "C" int rn_autotune_report(char* buf, int len)
{
    int n = 0;
    const int st = guarded_status("rn_autotune_report", [&] {
        const std::string r = tune_report();
        n = (int)r.size();
        if (buf && len > 0) {
            const int m = std::min(n, len - 1);
            std::memcpy(buf, r.data(), m);
            buf[m] = 0;
        }
    });
    return st != 0 ? st : n;
}

extern "C" int rn_autotune_reset(void)
{
    return guarded_status("rn_autotune_reset", [&] { tune_reset(); });
}

extern "C" int rn_autotune_pin(int kernel, int candidate)
{
    return guarded_status("rn_autotune_pin", [&] { tune_pin(kernel, candidate); });
}

extern "C" void* rn_fan_plan_create(int nx, int ny, int nz, float dx, float dy, float dz,
        int nu, int nview, int nv, float da, float dv, float off_a, float off_v,
        float dsd, float dso, const float* deg)
{
    Plan* p = nullptr;
    guarded("rn_fan_plan_create", [&] {
        FanGeo g = make_fan(nx, ny, nz, dx, dy, dz, nu, nview, nv, da, dv, off_a, off_v, dsd, dso);
        std::unique_ptr<Plan> q(new Plan());
        RN_HIP_CHECK(hipGetDevice(&q->device));
        CallStream s;
        q->tables.build(g, deg, s.get());
        p = q.release();
    });
    return p;
}

extern "C" void rn_fan_plan_destroy(void* plan)
{
    guarded("rn_fan_plan_destroy", [&] { delete static_cast<Plan*>(plan); });
}

extern "C" int rn_fan_forward(void* plan, int kind, const float* img, float* prj, int nBatches, int nChannels,
                              void* stream)
{
    return guarded_status("rn_fan_forward", [&] {
        RN_REQUIRE(plan != nullptr && img != nullptr && prj != nullptr, "null pointer");
        RN_REQUIRE(kind == RN_FP_DD || kind == RN_FP_SIDDON, "unknown forward projector kind");
        Plan* p = static_cast<Plan*>(plan);
        fan_forward(p->tables, kind == RN_FP_DD ? FanFP::DD : FanFP::Siddon, nBatches, nChannels, img, prj, p->work,
                    static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_fan_backward(void* plan, int kind, const float* prj, float* img, int nBatches, int nChannels,
                               void* stream)
{
    return guarded_status("rn_fan_backward", [&] {
        RN_REQUIRE(plan != nullptr && img != nullptr && prj != nullptr, "null pointer");
        FanBP k;
        switch (kind) {
        case RN_BP_DD: k = FanBP::DD; break;
        case RN_BP_DD_FBP: k = FanBP::DDFBP; break;
        case RN_BP_DD_ADJOINT: k = FanBP::DDAdjoint; break;
        case RN_BP_SIDDON: k = FanBP::Siddon; break;
        case RN_BP_PIXEL: k = FanBP::Pixel; break;
        default: throw Error("unknown backprojector kind");
        }
        Plan* p = static_cast<Plan*>(plan);
        fan_backward(p->tables, k, nBatches, nChannels, prj, img, p->work, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_fan_filter(void* plan, const float* prj, float* fprj, int nBatches, int nChannels,
                             int typeFilter, int fan, void* stream)
{
    return guarded_status("rn_fan_filter", [&] {
        RN_REQUIRE(plan != nullptr && prj != nullptr && fprj != nullptr, "null pointer");
        Plan* p = static_cast<Plan*>(plan);
        const FanGeo& g = p->tables.geo;
        filter_device(p->filter, prj, fprj, nBatches, nChannels, g.nu, g.nview, g.nv, g.da, g.off_a, g.dso,
                      typeFilter, fan != 0, p->t1, p->t2, static_cast<hipStream_t>(stream));
    });
}

// ----- circular cone-beam plans -----

namespace {
struct ConeHandle {
    ConePlan plan;
    DevBuf work;
};
} // namespace

extern "C" void* rn_cone_plan_create(int nx, int ny, int nz, float dx, float dy, float dz, float cx, float cy,
        float cz, int nu, int nview, int nv, float du, float dv, float off_u, float off_v,
        const float* detCenter, const float* detU, const float* detV, const float* src)
{
    ConeHandle* p = nullptr;
    guarded("rn_cone_plan_create", [&] {
        ArbGeo g{nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nview, nv, du, dv, off_u, off_v};
        std::unique_ptr<ConeHandle> q(new ConeHandle());
        CallStream s;
        q->plan.setup(g, detCenter, detU, detV, src, s.get());
        p = q.release();
    });
    return p;
}

extern "C" void rn_cone_plan_destroy(void* plan)
{
    guarded("rn_cone_plan_destroy", [&] { delete static_cast<ConeHandle*>(plan); });
}

extern "C" int rn_cone_forward(void* plan, int kind, const float* img, float* prj, int nBatches, int nChannels,
                               void* stream)
{
    return guarded_status("rn_cone_forward", [&] {
        RN_REQUIRE(plan != nullptr && img != nullptr && prj != nullptr, "null pointer");
        RN_REQUIRE(nBatches > 0 && nChannels > 0, "nBatches and nChannels must be positive");
        ConeHandle* p = static_cast<ConeHandle*>(plan);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (kind == RN_CONE_DD) p->plan.dd_fp(nBatches, nChannels, img, prj, st);
        else if (kind == RN_CONE_SIDDON) p->plan.siddon_fp(nBatches, nChannels, img, prj, p->work, st);
        else throw Error("unknown cone projector kind");
    });
}

extern "C" int rn_cone_backward(void* plan, int kind, const float* prj, float* img, int nBatches, int nChannels,
                                void* stream)
{
    return guarded_status("rn_cone_backward", [&] {
        RN_REQUIRE(plan != nullptr && img != nullptr && prj != nullptr, "null pointer");
        RN_REQUIRE(nBatches > 0 && nChannels > 0, "nBatches and nChannels must be positive");
        ConeHandle* p = static_cast<ConeHandle*>(plan);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (kind == RN_CONE_DD) p->plan.dd_adj(nBatches, nChannels, prj, img, p->work, st);
        else if (kind == RN_CONE_SIDDON) p->plan.siddon_bp(nBatches, nChannels, prj, img, p->work, st);
        else throw Error("unknown cone backprojector kind");
    });
}

extern "C" void* rn_sqs_create(void* plan, int nslices, int nviewPerSubset, int nSubsets)
{
    Sqs* out = nullptr;
    guarded("rn_sqs_create", [&] {
        RN_REQUIRE(plan != nullptr, "null plan");
        Plan* p = static_cast<Plan*>(plan);
        std::unique_ptr<Sqs> q(new Sqs());
        q->plan = p;
        CallStream s;
        q->engine.setup(p->tables.geo, p->tables.deg.data(), nslices, nviewPerSubset, nSubsets, s.get());
        out = q.release();
    });
    return out;
}

extern "C" void rn_sqs_destroy(void* sqs)
{
    guarded("rn_sqs_destroy", [&] { delete static_cast<Sqs*>(sqs); });
}

#define SQS_CALL(NAME, BODY)                                                                   \
    return guarded_status(NAME, [&] {                                                          \
        RN_REQUIRE(sqs != nullptr, "null engine");                                             \
        OsSqs& e = static_cast<Sqs*>(sqs)->engine;                                             \
        hipStream_t st = static_cast<hipStream_t>(stream);                                    \
        BODY;                                                                                  \
    })

extern "C" int rn_sqs_set_data(void* sqs, const float* prj, const float* wprj, void* stream)
{ SQS_CALL("rn_sqs_set_data", e.set_data(prj, wprj, st)); }
extern "C" int rn_sqs_compute_norm(void* sqs, void* stream)
{ SQS_CALL("rn_sqs_compute_norm", e.compute_norm(st)); }
extern "C" int rn_sqs_set_norm(void* sqs, const float* norm, void* stream)
{ SQS_CALL("rn_sqs_set_norm", e.set_norm(norm, st)); }
extern "C" int rn_sqs_get_norm(void* sqs, float* norm, void* stream)
{
    SQS_CALL("rn_sqs_get_norm", {
        const size_t n = (size_t)e.nslices() * e.tables().geo.nx * e.tables().geo.ny;
        RN_HIP_CHECK(hipMemcpyAsync(norm, e.norm(), n * sizeof(float), hipMemcpyDeviceToDevice, st));
    });
}
extern "C" int rn_sqs_set_image(void* sqs, const float* img, void* stream)
{ SQS_CALL("rn_sqs_set_image", e.set_image(img, st)); }
extern "C" int rn_sqs_get_image(void* sqs, float* img, void* stream)
{ SQS_CALL("rn_sqs_get_image", e.get_image(img, st)); }
extern "C" int rn_sqs_pass(void* sqs, float lambda, float beta, float eps, float lb, float ub, double* track,
                           void* stream)
{
    SQS_CALL("rn_sqs_pass", {
        SqsStep o;
        o.lambda = lambda; o.beta = beta; o.eps = eps; o.lb = lb; o.ub = ub;
        std::vector<double> tr;
        e.pass(o, st, track ? &tr : nullptr);
        if (track) std::copy(tr.begin(), tr.end(), track);
    });
}

extern "C" int rn_sqs_set_profiling(void* sqs, int on)
{
    void* stream = nullptr;
    SQS_CALL("rn_sqs_set_profiling", { (void)st; e.set_profiling(on != 0); });
}

extern "C" int rn_sqs_kernel_times(void* sqs, double* ms_fp, double* ms_bp, double* ms_update, long* steps)
{
    void* stream = nullptr;
    SQS_CALL("rn_sqs_kernel_times", { (void)st; e.kernel_times(ms_fp, ms_bp, ms_update, steps); });
}

extern "C" int rn_tv_iso_sqs(const float* img, float* s1, float* s2, float* var, int nb, int nx, int ny, int nz,
                             int nc, int dims, float eps, void* stream)
{
    return guarded_status("rn_tv_iso_sqs", [&] {
        tv_iso_sqs(img, s1, s2, var, nb, nx, ny, nz, nc, dims, eps, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_tv_sqs(const float* img, float* grad, float* norm, int nx, int ny, int nslices, float eps,
                         double* tv, void* stream)
{
    return guarded_status("rn_tv_sqs", [&] {
        RN_REQUIRE(img && grad && norm, "null pointer");
        DevBuf scratch;
        std::vector<double> t;
        tv_sqs(img, grad, norm, nx, ny, nslices, eps, tv ? &t : nullptr, scratch, static_cast<hipStream_t>(stream));
        if (tv) std::copy(t.begin(), t.end(), tv);
        else RN_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int rn_tv_sqs_noalloc(const float* img, float* grad, float* norm, int nx, int ny, int nslices,
                                 float eps, double* tv_dev, void* stream)
{
    return guarded_status("rn_tv_sqs_noalloc", [&] {
        RN_REQUIRE(img && grad && norm, "null pointer");
        tv_sqs_noalloc(img, grad, norm, nx, ny, nslices, eps, tv_dev, static_cast<hipStream_t>(stream));
    });
}

// ----- learned-prior convolutions (device pointers; bf16 tensors as raw uint16) -----

#include "rn_conv.h"

namespace {
DevBuf& conv_work()
{
    thread_local DevBuf w;
    return w;
}
} // namespace

extern "C" int rn_cnn_conv_in(const float* x, const float* w1, const float* b1, uint16_t* a1, int B, int H, int W,
                              void* stream)
{
    return guarded_status("rn_cnn_conv_in", [&] {
        RN_REQUIRE(x && w1 && b1 && a1 && B > 0 && H > 0 && W > 0, "invalid argument");
        conv_in(x, w1, b1, a1, B, H, W, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_conv_mid(const uint16_t* in, const uint16_t* wt, const float* bias, const uint16_t* mask,
                               uint16_t* out, int B, int H, int W, int relu, void* stream)
{
    return guarded_status("rn_cnn_conv_mid", [&] {
        RN_REQUIRE(in && wt && out && B > 0 && H > 0 && W > 0, "invalid argument");
        conv_mid(in, wt, bias, mask, out, B, H, W, relu != 0, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_conv_out(const uint16_t* a4, const float* w5, float b5, const float* xres, float* y, int B,
                               int H, int W, void* stream)
{
    return guarded_status("rn_cnn_conv_out", [&] {
        RN_REQUIRE(a4 && w5 && y && B > 0 && H > 0 && W > 0, "invalid argument");
        conv_out(a4, w5, b5, xres, y, B, H, W, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_forward(const float* x, const float* w1, const float* b1, const uint16_t* w2, const float* b2,
                              const uint16_t* w3, const float* b3, const uint16_t* w4, const float* b4,
                              const float* w5, float b5, float* y, int B, int H, int W, void* stream)
{
    return guarded_status("rn_cnn_forward", [&] {
        RN_REQUIRE(x && w1 && b1 && w2 && b2 && w3 && b3 && w4 && b4 && w5 && y && B > 0 && H > 0 && W > 0,
                   "invalid argument");
        cnn_forward(x, w1, b1, w2, b2, w3, b3, w4, b4, w5, b5, y, B, H, W, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_out_dgrad(const float* dy, const float* w5, const uint16_t* a4, uint16_t* dz4, int B, int H,
                                int W, void* stream)
{
    return guarded_status("rn_cnn_out_dgrad", [&] {
        RN_REQUIRE(dy && w5 && a4 && dz4 && B > 0 && H > 0 && W > 0, "invalid argument");
        conv_out_dgrad(dy, w5, a4, dz4, B, H, W, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_wgrad_narrow(const float* img, const uint16_t* act, float* grad, int B, int H, int W, int mode,
                                   void* stream)
{
    return guarded_status("rn_cnn_wgrad_narrow", [&] {
        RN_REQUIRE(img && act && grad && B > 0 && H > 0 && W > 0 && (mode == 0 || mode == 1), "invalid argument");
        wgrad_narrow(img, act, grad, B, H, W, mode, conv_work(), static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_bias_grad(const uint16_t* dz, float* grad, long npx, void* stream)
{
    return guarded_status("rn_cnn_bias_grad", [&] {
        RN_REQUIRE(dz && grad && npx > 0, "invalid argument");
        bias_grad(dz, grad, npx, conv_work(), static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_cnn_wgrad_mid(const uint16_t* in, const uint16_t* dz, float* grad, int B, int H, int W,
                                void* stream)
{
    return guarded_status("rn_cnn_wgrad_mid", [&] {
        RN_REQUIRE(in && dz && grad && B > 0 && H > 0 && W > 0, "invalid argument");
        wgrad_mid(in, dz, grad, B, H, W, conv_work(), static_cast<hipStream_t>(stream));
    });
}

// ----- ragged (uneven-size) persistent batches -----

#include "rn_ragged.h"

extern "C" void* rn_ragged_create(int nslices, const int* dims, const float* geo, const float* angles)
{
    RaggedBatch* out = nullptr;
    guarded("rn_ragged_create", [&] {
        RN_REQUIRE(nslices > 0 && dims && geo && angles, "invalid argument");
        std::vector<FanGeo> geos;
        std::vector<const float*> ang;
        const float* a = angles;
        for (int s = 0; s < nslices; s++) {
            const int* d = dims + 4 * s;
            const float* f = geo + 6 * s;
            geos.push_back(make_fan(d[0], d[1], 1, f[0], f[1], 1.0f, d[2], d[3], 1, f[2], 1.0f, f[3], 0.0f, f[4], f[5]));
            ang.push_back(a);
            a += d[3];
        }
        std::unique_ptr<RaggedBatch> r(new RaggedBatch());
        CallStream s;
        r->setup(geos, ang, s.get());
        out = r.release();
    });
    return out;
}

extern "C" void rn_ragged_destroy(void* h)
{
    guarded("rn_ragged_destroy", [&] { delete static_cast<RaggedBatch*>(h); });
}

extern "C" int rn_ragged_forward(void* h, int kind, const float* img, float* prj, void* stream)
{
    return guarded_status("rn_ragged_forward", [&] {
        RN_REQUIRE(h && img && prj && (kind == RN_FP_DD || kind == RN_FP_SIDDON), "invalid argument");
        static_cast<RaggedBatch*>(h)->forward(kind == RN_FP_SIDDON, img, prj, static_cast<hipStream_t>(stream));
    });
}

extern "C" int rn_ragged_backward(void* h, int kind, const float* prj, float* img, void* stream)
{
    return guarded_status("rn_ragged_backward", [&] {
        RN_REQUIRE(h && img && prj && (kind == RN_BP_DD_ADJOINT || kind == RN_BP_SIDDON), "invalid argument");
        static_cast<RaggedBatch*>(h)->backward(kind == RN_BP_SIDDON, prj, img, static_cast<hipStream_t>(stream));
    });
}
