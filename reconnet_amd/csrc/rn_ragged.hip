// Persistent projector launches over a batch of 2D fan slices of uneven size
// (no reference counterpart: the reference ABI takes one geometry per call).
//
// Each slice has its own geometry, angle set and trig tables.  One launch of a
// fixed grid serves the whole batch: every wave repeatedly takes the next work
// item from a global counter (a 64-bin strip of one view for the FP, an 8 x 8
// pixel tile for the BP), finds its descriptor and runs the same per-ray /
// per-pixel code as the single-geometry kernels, so the results are identical
// to independent per-slice launches.  Consecutive slices of one geometry and
// angle set share a descriptor, up to kRagS of them: a lane then carries the
// ray or pixel through all of them, its geometry work done once (each slice
// with the arithmetic of a one-slice call; the forward projection takes a
// group in chunks of kRagFp).  Descriptors are ordered by
// decreasing cost so the heaviest start first.  Every wave leaves the loop once
// the counter passes the item total.

#include "rn_ragged.h"

#include "rn_fan_device.h"

#include <algorithm>
#include <numeric>

namespace rn {

namespace {

using namespace fan;

constexpr int kRagS = 8;    // slices of one geometry per backprojection descriptor
constexpr int kRagFp = 4;   // per forward descriptor (a group's chunks): the ray
                            // loops carry fewer registers and the launch more items
constexpr int kRagSidBp = 4;   // per Siddon backprojection descriptor (the same)

struct RagSlice {
    FanK g;
    const float* pa;     // FP: RX (DD) or Pyx (Siddon); BP: prj [nview][nu]
    const float* pb;     // FP: RY (DD) or Pxy (Siddon)
    float* out;          // FP: prj [nu][nview]; BP: img [nx][ny]
    int item0, tiles, nitems, nsl;   // nsl consecutive slices, nx * ny (image) / nu * nview (sinogram) apart
};

__device__ __forceinline__ int find_slice(const RagSlice* __restrict__ sl, int nsl, int item)
{
    int s = 0;
    while (s + 1 < nsl && sl[s + 1].item0 <= item) ++s;
    return s;
}

__device__ __forceinline__ int next_item(int* counter, int lane)
{
    int it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    return __shfl(it, 0);
}

__global__ __launch_bounds__(256) void k_ragged_fp(const RagSlice* __restrict__ sl, int nsl, int total,
                                                   int* __restrict__ counter, int siddon)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int item = next_item(counter, lane);
        if (item >= total) break;
        const RagSlice& d = sl[find_slice(sl, nsl, item)];
        const int local = item - d.item0;
        const int view = local / d.tiles, iu = (local % d.tiles) * 64 + lane;
        const bool active = iu < d.g.nu;
        const long plane = (long)d.g.nx * d.g.ny, sino = (long)d.g.nu * d.g.nview;
        // every lane takes part (wave-uniform line loops)
        float acc[kRagFp];
        if (siddon)
            siddon_fp_ray_slices<float, kRagFp>(d.g, d.pa, d.pb, plane, d.nsl, view, active ? iu : d.g.nu - 1, active,
                                                acc);
        else
            dd_fp_ray_slices<float, kRagFp>(d.g, d.pa, d.pb, plane, d.nsl, view, active ? iu : d.g.nu - 1, active,
                                            acc);
        if (active) {
            #pragma unroll
            for (int q = 0; q < kRagFp; ++q)
                if (q < d.nsl) d.out[q * sino + (long)iu * d.g.nview + view] = acc[q];
        }
    }
}

// S: the descriptors' slice bound (kRagSidBp for Siddon, kRagS for the DD adjoint)
template <bool SIDDON, int S>
__global__ __launch_bounds__(256) void k_ragged_bp(const RagSlice* __restrict__ sl, int nsl, int total,
                                                   int* __restrict__ counter)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int item = next_item(counter, lane);
        if (item >= total) break;
        const RagSlice& d = sl[find_slice(sl, nsl, item)];
        const int local = item - d.item0;
        const int ix = (local / d.tiles) * 8 + (lane >> 3), iy = (local % d.tiles) * 8 + (lane & 7);
        if (ix < d.g.nx && iy < d.g.ny) {
            const long plane = (long)d.g.nx * d.g.ny, sino = (long)d.g.nu * d.g.nview;
            float acc[S];
            if (SIDDON)
                siddon_bp_pixel_slices<float, 4, S>(d.g, d.pa, sino, d.nsl, 0, d.g.nview, ix, iy, acc);
            else if (d.g.unit_angle)
                dd_adj_pixel_slices<true, S>(d.g, d.pa, sino, d.nsl, 0, d.g.nview, ix, iy, acc);
            else
                dd_adj_pixel_slices<false, S>(d.g, d.pa, sino, d.nsl, 0, d.g.nview, ix, iy, acc);
            #pragma unroll
            for (int q = 0; q < S; ++q)
                if (q < d.nsl) d.out[q * plane + (long)ix * d.g.ny + iy] = acc[q];
        }
    }
}

__global__ void k_zero_int(int* p) { *p = 0; }

int persistent_blocks()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
        hipSuccess || ncu <= 0)
        ncu = 256;
    return 4 * ncu;
}

} // namespace

struct RaggedBatch::Impl {
    std::vector<FanTables> tables;
    std::vector<long> img_off, sino_off;
    std::vector<int> first, count;   // descriptor k: slices first[k] .. first[k] + count[k] - 1
    std::vector<int> order;          // descriptors by decreasing cost
    long img_total = 0, sino_total = 0;
    DevBuf desc, counter, wa, wb, wc;
    std::vector<RagSlice> host;
};

RaggedBatch::RaggedBatch() : p_(new Impl) {}
RaggedBatch::~RaggedBatch() { delete p_; }

void RaggedBatch::setup(const std::vector<FanGeo>& geos, const std::vector<const float*>& angles, hipStream_t st)
{
    RN_REQUIRE(!geos.empty() && geos.size() == angles.size(), "ragged batch needs one angle set per slice");
    const int n = (int)geos.size();
    p_->tables.resize(n);
    p_->img_off.resize(n);
    p_->sino_off.resize(n);
    long io = 0, so = 0;
    for (int s = 0; s < n; s++) {
        RN_REQUIRE(geos[s].nz == 1 && geos[s].nv == 1, "ragged batches hold 2D slices");
        // the per-slice gathers address a plane, a ray table and a sinogram with 32-bit byte offsets
        RN_REQUIRE((long)geos[s].nx * geos[s].ny * 4L < 0x7fffffffL &&
                       (long)geos[s].nu * geos[s].nview * 16L < 0x7fffffffL,
                   "a ragged slice's image, ray table or sinogram exceeds 2 GiB");
        p_->tables[s].build(geos[s], angles[s], st);
        p_->img_off[s] = io;
        p_->sino_off[s] = so;
        io += (long)geos[s].nx * geos[s].ny;
        so += (long)geos[s].nu * geos[s].nview;
    }
    p_->img_total = io;
    p_->sino_total = so;
    // runs of consecutive slices with one geometry and angle set, in chunks of kRagS
    auto same = [&](int a, int b) {
        const FanGeo& x = geos[a];
        const FanGeo& y = geos[b];
        return x.nx == y.nx && x.ny == y.ny && x.nz == y.nz && x.dx == y.dx && x.dy == y.dy && x.dz == y.dz &&
               x.nu == y.nu && x.nview == y.nview && x.nv == y.nv && x.da == y.da && x.dv == y.dv &&
               x.off_a == y.off_a && x.off_v == y.off_v && x.dsd == y.dsd && x.dso == y.dso &&
               std::equal(angles[a], angles[a] + x.nview, angles[b]);
    };
    p_->first.clear();
    p_->count.clear();
    for (int s = 0; s < n; s++) {
        if (!p_->count.empty() && p_->count.back() < kRagS && same(p_->first.back(), s)) {
            ++p_->count.back();
        } else {
            p_->first.push_back(s);
            p_->count.push_back(1);
        }
    }
    const int nd = (int)p_->first.size();
    p_->order.resize(nd);
    std::iota(p_->order.begin(), p_->order.end(), 0);
    std::stable_sort(p_->order.begin(), p_->order.end(), [&](int a, int b) {
        const FanGeo& ga = geos[p_->first[a]];
        const FanGeo& gb = geos[p_->first[b]];
        return (double)ga.nx * ga.ny * ga.nview > (double)gb.nx * gb.ny * gb.nview;
    });
    p_->desc.alloc(nd * ((kRagS + kRagFp - 1) / kRagFp) * sizeof(RagSlice));
    p_->counter.alloc(sizeof(int));
    p_->wa.alloc(io * sizeof(float));
    p_->wb.alloc(std::max(io, so) * sizeof(float));
    p_->wc.alloc(io * sizeof(float));
}

int RaggedBatch::size() const { return (int)p_->tables.size(); }
long RaggedBatch::image_elems() const { return p_->img_total; }
long RaggedBatch::sino_elems() const { return p_->sino_total; }

void RaggedBatch::forward(bool siddon, const float* img, float* prj, hipStream_t st)
{
    const int nd = (int)p_->first.size();
    float* A = p_->wa.as<float>();
    float* Bb = p_->wb.as<float>();
    p_->host.clear();
    int item = 0;
    for (int k = 0; k < nd; k++) {
        const int s = p_->first[p_->order[k]], cnt = p_->count[p_->order[k]];
        const FanGeo& g = p_->tables[s].geo;
        const float* P = img + p_->img_off[s];
        float* a = A + p_->img_off[s];
        float* b = Bb + p_->img_off[s];
        const float* pa;
        const float* pb;
        if (siddon) {
            launch_transpose(P, a, g.nx, g.ny, cnt, st);    // rows contiguous
            pa = a;
            pb = P;
        } else {
            // RY[x][y] = (P[x-1][y] + P[x][y]) / 2 ;  RX = transpose((P[x][y-1] + P[x][y]) / 2)
            float* c = p_->wc.as<float>() + p_->img_off[s];
            launch_avg_prev(P, b, g.nx, g.ny, cnt, 0, st);
            launch_avg_prev(P, c, g.nx, g.ny, cnt, 1, st);
            launch_transpose(c, a, g.nx, g.ny, cnt, st);
            pa = a;
            pb = b;
        }
        // the group's slices in chunks of kRagFp, consecutive in both layouts
        const long plane = (long)g.nx * g.ny, sino = (long)g.nu * g.nview;
        for (int c0 = 0; c0 < cnt; c0 += kRagFp) {
            RagSlice d{};
            d.g = p_->tables[s].k;
            d.nsl = std::min(kRagFp, cnt - c0);
            d.pa = pa + c0 * plane;
            d.pb = pb + c0 * plane;
            d.out = prj + p_->sino_off[s] + c0 * sino;
            d.tiles = cdiv(g.nu, 64);
            d.nitems = d.tiles * g.nview;
            d.item0 = item;
            item += d.nitems;
            p_->host.push_back(d);
        }
    }
    const int nfd = (int)p_->host.size();
    RN_HIP_CHECK(hipMemcpyAsync(p_->desc.get(), p_->host.data(), nfd * sizeof(RagSlice), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p_->counter.as<int>());
    hipLaunchKernelGGL(k_ragged_fp, dim3(persistent_blocks()), dim3(256), 0, st, p_->desc.as<RagSlice>(), nfd, item,
                       p_->counter.as<int>(), siddon ? 1 : 0);
    check_launch("ragged_fp");
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

void RaggedBatch::backward(bool siddon, const float* prj, float* img, hipStream_t st)
{
    const int nd = (int)p_->first.size();
    float* V = p_->wb.as<float>();   // view-major copies of the sinograms
    p_->host.clear();
    const int chunk = siddon ? kRagSidBp : kRagS;
    int item = 0;
    for (int k = 0; k < nd; k++) {
        const int s = p_->first[p_->order[k]], cnt = p_->count[p_->order[k]];
        const FanGeo& g = p_->tables[s].geo;
        float* v = V + p_->sino_off[s];
        launch_transpose(prj + p_->sino_off[s], v, g.nu, g.nview, cnt, st);
        const long plane = (long)g.nx * g.ny, sino = (long)g.nu * g.nview;
        for (int c0 = 0; c0 < cnt; c0 += chunk) {
            RagSlice d{};
            d.g = p_->tables[s].k;
            d.nsl = std::min(chunk, cnt - c0);
            d.pa = v + c0 * sino;
            d.pb = nullptr;
            d.out = img + p_->img_off[s] + c0 * plane;
            d.tiles = cdiv(g.ny, 8);
            d.nitems = d.tiles * cdiv(g.nx, 8);
            d.item0 = item;
            item += d.nitems;
            p_->host.push_back(d);
        }
    }
    const int nbd = (int)p_->host.size();
    RN_HIP_CHECK(hipMemcpyAsync(p_->desc.get(), p_->host.data(), nbd * sizeof(RagSlice), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p_->counter.as<int>());
    if (siddon)
        hipLaunchKernelGGL((k_ragged_bp<true, kRagSidBp>), dim3(persistent_blocks()), dim3(256), 0, st,
                           p_->desc.as<RagSlice>(), nbd, item, p_->counter.as<int>());
    else
        hipLaunchKernelGGL((k_ragged_bp<false, kRagS>), dim3(persistent_blocks()), dim3(256), 0, st,
                           p_->desc.as<RagSlice>(), nbd, item, p_->counter.as<int>());
    check_launch("ragged_bp");
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

} // namespace rn
