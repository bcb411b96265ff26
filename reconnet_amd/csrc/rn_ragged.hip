// Persistent projector launches over a batch of 2D fan slices of uneven size
// (no reference counterpart: the reference ABI takes one geometry per call).
//
// Each slice has its own geometry, angle set and trig tables.  One launch of a
// fixed grid serves the whole batch: every wave repeatedly takes the next work
// item from a global counter (a 64-bin strip of one view for the FP, an 8 x 8
// pixel tile for the BP), finds its slice in the descriptor table and runs the
// same per-ray / per-pixel code as the single-geometry kernels, so the results
// are identical to independent per-slice launches.  Descriptors are ordered
// by decreasing cost so the heaviest slices start first.  Every wave leaves the
// loop once the counter passes the item total.

#include "rn_ragged.h"

#include "rn_fan_device.h"

#include <algorithm>
#include <numeric>

namespace rn {

namespace {

using namespace fan;

struct RagSlice {
    FanK g;
    const float* pa;     // FP: RX (DD) or Pyx (Siddon); BP: prj [nview][nu]
    const float* pb;     // FP: RY (DD) or Pxy (Siddon)
    float* out;          // FP: prj [nu][nview]; BP: img [nx][ny]
    int item0, tiles, nitems, pad;
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
        // every lane takes part (wave-uniform line loops)
        float v;
        if (siddon) {
            v = siddon_fp_ray<float>(d.g, d.pa, d.pb, view, active ? iu : d.g.nu - 1, active);
        } else {
            float acc[1];
            dd_fp_ray_slices<float, 1>(d.g, d.pa, d.pb, 1, 1, view, active ? iu : d.g.nu - 1, active, acc);
            v = acc[0];
        }
        if (active) d.out[(long)iu * d.g.nview + view] = v;
    }
}

__global__ __launch_bounds__(256) void k_ragged_bp(const RagSlice* __restrict__ sl, int nsl, int total,
                                                   int* __restrict__ counter, int siddon)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int item = next_item(counter, lane);
        if (item >= total) break;
        const RagSlice& d = sl[find_slice(sl, nsl, item)];
        const int local = item - d.item0;
        const int ix = (local / d.tiles) * 8 + (lane >> 3), iy = (local % d.tiles) * 8 + (lane & 7);
        if (ix < d.g.nx && iy < d.g.ny) {
            float v;
            if (siddon) {
                v = siddon_bp_pixel<float>(d.g, d.pa, 0, d.g.nview, ix, iy);
            } else {
                float acc[1];
                if (d.g.unit_angle) dd_adj_pixel_slices<true, 1>(d.g, d.pa, 0, 1, 0, d.g.nview, ix, iy, acc);
                else dd_adj_pixel_slices<false, 1>(d.g, d.pa, 0, 1, 0, d.g.nview, ix, iy, acc);
                v = acc[0];
            }
            d.out[(long)ix * d.g.ny + iy] = v;
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
    std::vector<int> order;       // slices by decreasing cost
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
    p_->order.resize(n);
    std::iota(p_->order.begin(), p_->order.end(), 0);
    std::stable_sort(p_->order.begin(), p_->order.end(), [&](int a, int b) {
        const double ca = (double)geos[a].nx * geos[a].ny * geos[a].nview;
        const double cb = (double)geos[b].nx * geos[b].ny * geos[b].nview;
        return ca > cb;
    });
    p_->desc.alloc(n * sizeof(RagSlice));
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
    const int n = size();
    float* A = p_->wa.as<float>();
    float* Bb = p_->wb.as<float>();
    p_->host.assign(n, RagSlice{});
    int item = 0;
    for (int k = 0; k < n; k++) {
        const int s = p_->order[k];
        const FanGeo& g = p_->tables[s].geo;
        const float* P = img + p_->img_off[s];
        float* a = A + p_->img_off[s];
        float* b = Bb + p_->img_off[s];
        RagSlice& d = p_->host[k];
        d.g = p_->tables[s].k;
        if (siddon) {
            launch_transpose(P, a, g.nx, g.ny, 1, st);      // rows contiguous
            d.pa = a;
            d.pb = P;
        } else {
            // RY[x][y] = (P[x-1][y] + P[x][y]) / 2 ;  RX = transpose((P[x][y-1] + P[x][y]) / 2)
            float* c = p_->wc.as<float>() + p_->img_off[s];
            launch_avg_prev(P, b, g.nx, g.ny, 1, 0, st);
            launch_avg_prev(P, c, g.nx, g.ny, 1, 1, st);
            launch_transpose(c, a, g.nx, g.ny, 1, st);
            d.pa = a;
            d.pb = b;
        }
        d.out = prj + p_->sino_off[s];
        d.tiles = cdiv(g.nu, 64);
        d.nitems = d.tiles * g.nview;
        d.item0 = item;
        item += d.nitems;
    }
    RN_HIP_CHECK(hipMemcpyAsync(p_->desc.get(), p_->host.data(), n * sizeof(RagSlice), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p_->counter.as<int>());
    hipLaunchKernelGGL(k_ragged_fp, dim3(persistent_blocks()), dim3(256), 0, st, p_->desc.as<RagSlice>(), n, item,
                       p_->counter.as<int>(), siddon ? 1 : 0);
    check_launch("ragged_fp");
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

void RaggedBatch::backward(bool siddon, const float* prj, float* img, hipStream_t st)
{
    const int n = size();
    float* V = p_->wb.as<float>();   // view-major copies of the sinograms
    p_->host.assign(n, RagSlice{});
    int item = 0;
    for (int k = 0; k < n; k++) {
        const int s = p_->order[k];
        const FanGeo& g = p_->tables[s].geo;
        float* v = V + p_->sino_off[s];
        launch_transpose(prj + p_->sino_off[s], v, g.nu, g.nview, 1, st);
        RagSlice& d = p_->host[k];
        d.g = p_->tables[s].k;
        d.pa = v;
        d.pb = nullptr;
        d.out = img + p_->img_off[s];
        d.tiles = cdiv(g.ny, 8);
        d.nitems = d.tiles * cdiv(g.nx, 8);
        d.item0 = item;
        item += d.nitems;
    }
    RN_HIP_CHECK(hipMemcpyAsync(p_->desc.get(), p_->host.data(), n * sizeof(RagSlice), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p_->counter.as<int>());
    hipLaunchKernelGGL(k_ragged_bp, dim3(persistent_blocks()), dim3(256), 0, st, p_->desc.as<RagSlice>(), n, item,
                       p_->counter.as<int>(), siddon ? 1 : 0);
    check_launch("ragged_bp");
    RN_HIP_CHECK(hipStreamSynchronize(st));
}

} // namespace rn
