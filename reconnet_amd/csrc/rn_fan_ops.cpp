// Fan-beam operations on reference-layout device buffers: layout changes,
// z/v mixing and the 2D kernels of rn_fan.hip.
//
// The reference runs one kernel pair per (batch, channel) with a 3D texture
// whose z axis interpolates between image planes / detector rows
// (DistanceDriven.cu:86,99,518; PixelDriven.cu:60; Siddon voxel plane
// selection Siddon.cu:45-61).  Here every (batch, plane, channel) becomes one
// 2D slice of a single launch, and the z/v interpolation is a small mixing
// matrix applied while the slices are gathered.

#include "rn_fan_ops.h"

#include <cmath>

namespace rn {

namespace {

void add_lerp(std::vector<float>& M, int row, int ncols, double c)
{
    // linear texture fetch at unnormalised coordinate c with border addressing
    const double p = c - 0.5;
    const double p0 = std::floor(p);
    const double f = p - p0;
    const long i0 = (long)p0;
    if (i0 >= 0 && i0 < ncols) M[(size_t)row * ncols + i0] += (float)(1.0 - f);
    if (i0 + 1 >= 0 && i0 + 1 < ncols) M[(size_t)row * ncols + i0 + 1] += (float)f;
}

std::vector<float> transpose_host(const std::vector<float>& M, int rows, int cols)
{
    std::vector<float> T((size_t)rows * cols);
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) T[(size_t)c * rows + r] = M[(size_t)r * cols + c];
    return T;
}

bool is_identity_1(const std::vector<float>& M, int nB_unused, int nC)
{
    (void)nB_unused;
    return M.size() == 1 && M[0] == 1.0f && nC == 1;
}

const float* upload_matrix(DevBuf& buf, const std::vector<float>& M, hipStream_t st)
{
    buf.ensure(M.size() * sizeof(float));
    RN_HIP_CHECK(hipMemcpyAsync(buf.get(), M.data(), M.size() * sizeof(float), hipMemcpyHostToDevice, st));
    RN_HIP_CHECK(hipStreamSynchronize(st));
    return buf.as<float>();
}

std::vector<float> identity(int n)
{
    std::vector<float> I((size_t)n * n, 0.0f);
    for (int i = 0; i < n; i++) I[(size_t)i * n + i] = 1.0f;
    return I;
}

} // namespace

std::vector<float> fp_mix_matrix(const FanGeo& g, FanFP kind)
{
    std::vector<float> M((size_t)g.nv * g.nz, 0.0f);
    for (int iv = 0; iv < g.nv; iv++) {
        const double z = (iv - (g.nv - 1) / 2.0 + (double)g.off_v) * (double)g.dv;
        const double zc = z / (double)g.dz + g.nz / 2.0;
        if (kind == FanFP::DD) {
            add_lerp(M, iv, g.nz, zc);          // tex3D z coordinate, border
        } else if (zc >= 0.0 && zc < (double)g.nz) {
            M[(size_t)iv * g.nz + (int)std::floor(zc)] = 1.0f;   // Siddon: the plane the ray lies in
        }
    }
    return M;
}

std::vector<float> bp_mix_matrix(const FanGeo& g, FanBP kind)
{
    switch (kind) {
    case FanBP::DDAdjoint: return transpose_host(fp_mix_matrix(g, FanFP::DD), g.nv, g.nz);
    case FanBP::Siddon: return transpose_host(fp_mix_matrix(g, FanFP::Siddon), g.nv, g.nz);
    default: break;
    }
    std::vector<float> M((size_t)g.nz * g.nv, 0.0f);
    for (int iz = 0; iz < g.nz; iz++) {
        const double ivc = (iz - (g.nz - 1) / 2.0) * (double)g.dz / (double)g.dv + (double)g.off_v + (g.nv - 1) / 2.0;
        add_lerp(M, iz, g.nv, ivc + 0.5);
    }
    return M;
}

void fan_forward(const FanTables& t, FanFP kind, int nB, int nC, const float* img_ref, float* prj_ref,
                 FanWork& w, hipStream_t st)
{
    const FanGeo& g = t.geo;
    RN_REQUIRE(nB > 0 && nC > 0, "nBatches and nChannels must be positive");
    const long plane = (long)g.nx * g.ny;
    const long sino = (long)g.nu * g.nview;
    const int ns = nB * g.nv * nC;
    const std::vector<float> M = fp_mix_matrix(g, kind);
    const bool direct_in = (g.nz == 1 && g.nv == 1 && is_identity_1(M, nB, nC));
    const bool direct_out = (g.nv * nC == 1);
    const size_t big = (size_t)ns * std::max(plane, sino) * sizeof(float);
    w.rx.ensure(big);
    w.ry.ensure(big);
    w.sino.ensure(big);

    const float* P = img_ref;
    if (!direct_in) {
        w.planes.ensure((size_t)ns * plane * sizeof(float));
        const float* Md = upload_matrix(w.mix, M, st);
        launch_mix_gather(img_ref, w.planes.as<float>(), Md, nB, (int)plane, g.nz, nC, g.nv, st);
        P = w.planes.as<float>();
    }
    float* out_vu = w.sino.as<float>();
    if (kind == FanFP::DD && ns >= 4) {
        // groups of 4 slices interleaved, padded lines (zero pads):
        // QY[q][x][pad + y] = (P[x-1][y] + P[x][y]) / 2,  QX[q][y][pad + x] = (P[x][y-1] + P[x][y]) / 2
        const size_t qx = (size_t)g.ny * qpitch(g.nx), qy = (size_t)g.nx * qpitch(g.ny);
        const int nq = cdiv(ns, 4);
        w.qdd.ensure((qx + qy) * nq * sizeof(float4));
        float4* QX = w.qdd.as<float4>();
        float4* QY = QX + qx * nq;
        launch_fill(w.qdd.as<float>(), 0.0f, (qx + qy) * nq * 4, st);
        launch_avg_prev(P, w.ry.as<float>(), g.nx, g.ny, ns, 0, st);
        launch_pack_quads(w.ry.as<float>(), QY, nullptr, g.nx, g.ny, ns, true, st);
        launch_avg_prev(P, w.rx.as<float>(), g.nx, g.ny, ns, 1, st);
        launch_pack_quads(w.rx.as<float>(), nullptr, QX, g.nx, g.ny, ns, true, st);
        launch_dd_fan_fp_q(t.k, QX, QY, out_vu, ns, 0, g.nview, st);
    } else if (kind == FanFP::DD) {
        // RY[x][y] = (P[x-1][y] + P[x][y]) / 2 ;  RX = transpose((P[x][y-1] + P[x][y]) / 2)
        launch_avg_prev(P, w.ry.as<float>(), g.nx, g.ny, ns, 0, st);
        launch_avg_prev(P, out_vu, g.nx, g.ny, ns, 1, st);
        launch_transpose(out_vu, w.rx.as<float>(), g.nx, g.ny, ns, st);
        launch_dd_fan_fp(t.k, w.rx.as<float>(), w.ry.as<float>(), out_vu, ns, 0, g.nview, st);
    } else {
        launch_transpose(P, w.rx.as<float>(), g.nx, g.ny, ns, st);
        launch_siddon_fan_fp(t.k, w.rx.as<float>(), P, out_vu, ns, 0, g.nview, st);
    }
    if (direct_out) {
        launch_transpose(out_vu, prj_ref, g.nview, g.nu, ns, st);
    } else {
        launch_transpose(out_vu, w.ry.as<float>(), g.nview, g.nu, ns, st);
        const float* I = upload_matrix(w.mix, identity(g.nv), st);
        launch_mix_scatter(w.ry.as<float>(), prj_ref, I, nB, (int)sino, g.nv, nC, g.nv, false, st);
    }
}

void fan_backward(const FanTables& t, FanBP kind, int nB, int nC, const float* prj_ref, float* img_ref,
                  FanWork& w, hipStream_t st)
{
    const FanGeo& g = t.geo;
    RN_REQUIRE(nB > 0 && nC > 0, "nBatches and nChannels must be positive");
    const long plane = (long)g.nx * g.ny;
    const long sino = (long)g.nu * g.nview;
    const int ns = nB * g.nz * nC;
    const std::vector<float> M = bp_mix_matrix(g, kind);
    const bool direct_in = (g.nz == 1 && g.nv == 1 && is_identity_1(M, nB, nC));
    const bool direct_out = (g.nz * nC == 1);
    const size_t big = (size_t)ns * std::max(plane, sino) * sizeof(float);
    w.rx.ensure(big);
    w.ry.ensure(big);
    w.sino.ensure(big);

    const float* Q = prj_ref;   // [slice][nu][nview]
    if (!direct_in) {
        const float* Md = upload_matrix(w.mix, M, st);
        launch_mix_gather(prj_ref, w.ry.as<float>(), Md, nB, (int)sino, g.nv, nC, g.nz, st);
        Q = w.ry.as<float>();
    }
    float* W = w.sino.as<float>();   // [slice][nview][nu]
    launch_transpose(Q, W, g.nu, g.nview, ns, st);
    float* out = direct_out ? img_ref : w.rx.as<float>();
    switch (kind) {
    case FanBP::DD:
    case FanBP::DDFBP: {
        // the FBP variant backprojects the sinogram as given; the iterative
        // one pre-weights it by the line length factor first
        const bool fbp = kind == FanBP::DDFBP;
        const float* P = W;
        if (!fbp) {
            launch_dd_preweight(t.k, W, w.ry.as<float>(), ns, st);
            P = w.ry.as<float>();
        }
        if (ns >= 4) {
            // groups of 4 slices interleaved: one float4 load per bin and group
            w.quad.ensure((size_t)cdiv(ns, 4) * sino * sizeof(float4));
            launch_pack_quads(P, w.quad.as<float4>(), nullptr, g.nview, g.nu, ns, false, st);
            launch_dd_fan_bp_q(t.k, w.quad.as<float4>(), out, ns, fbp, st);
        } else {
            launch_dd_fan_bp(t.k, P, out, ns, fbp, st);
        }
        break;
    }
    case FanBP::DDAdjoint:
        launch_dd_fan_adj(t.k, W, out, ns, 0, g.nview, false, st);
        break;
    case FanBP::Siddon:
        launch_siddon_fan_bp(t.k, W, out, ns, 0, g.nview, false, st);
        break;
    case FanBP::Pixel:
        launch_pixel_fan_bp(t.k, W, out, ns, st);
        break;
    }
    if (!direct_out) {
        const float* I = upload_matrix(w.mix, identity(g.nz), st);
        launch_mix_scatter(out, img_ref, I, nB, (int)plane, g.nz, nC, g.nz, false, st);
    }
}

} // namespace rn
