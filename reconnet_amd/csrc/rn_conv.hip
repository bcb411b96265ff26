// 3x3 "same" convolutions of the learned prior (no reference counterpart).
//
// Network: 1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, 3x3 kernels, ReLU after
// the first four layers, residual output y = x + conv5(...).
// Activations are BF16, NHWC ([B][H][W][32], 64 bytes per pixel); every
// product accumulates in FP32.
//
// The 32 -> 32 layers (forward and data gradient) run on gfx950 BF16 matrix
// cores: v_mfma_f32_16x16x32_bf16 with A = weights (16 output channels x 32
// input channels of one tap) and B = 16 consecutive pixels x 32 input
// channels; 9 taps x 2 channel tiles = 18 MFMAs per 16-pixel segment.  The
// weight gradient of those layers is an MFMA GEMM over pixels with LDS
// transposes.  The 1 -> 32 and 32 -> 1 layers (K = 9, N = 1) are VALU kernels.

#include "rn_conv.h"
#include "rn_tune.h"

#include <type_traits>

namespace rn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> two bf16 (round to nearest even, NaN stays NaN) packed low | high << 16,
// one hardware conversion (v_cvt_pk_bf16_f32) for the pair
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    const f32x2 f = {lo, hi};
    const bf16x2 h = __builtin_convertvector(f, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &h, sizeof(u));
    return u;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v)
{
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

// ---------------------------------------------------------------------------
// layer 1: x (fp32 [B][H][W]) -> relu(conv(x) + b) (bf16 NHWC), w [9][32]
//
// One wave per 62 pixels of an image row: lane l holds column x0 + l - 1 of
// the three input rows (one load per row, zero outside the image), the left
// and right neighbours come from the adjacent lanes, and lanes 1 .. 62
// compute their pixel's 32 channels (lanes 0 and 63 only supply neighbours).
// The wave's 62 x 64 output bytes are gathered in LDS and written with
// contiguous 16-byte stores, 1 KB per instruction.
// ---------------------------------------------------------------------------
constexpr int kConvInTile = 62;

__global__ __launch_bounds__(256) void k_conv_in(const float* __restrict__ x, const float* __restrict__ w,
                                                 const float* __restrict__ b, uint16_t* __restrict__ out,
                                                 int H, int W, int tiles, long nwaves)
{
    // weights and biases are read at wave-uniform addresses (scalar loads)
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= nwaves) return;                  // whole waves
    const int tile = (int)(wave % tiles);
    const long row = wave / tiles;               // image * H + yh
    const int yh = (int)(row % H);
    const long img = (row - yh) * W;             // first pixel of the image
    const int xc = tile * kConvInTile + lane - 1;
    const bool col = xc >= 0 && xc < W;
    float c[3];
    #pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
        const int yy = yh + dy - 1;
        c[dy] = (col && yy >= 0 && yy < H) ? x[img + (long)yy * W + xc] : 0.0f;
    }
    float v[9];                                  // tap t = 3 dy + dx
    #pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
        v[3 * dy] = __shfl_up(c[dy], 1);
        v[3 * dy + 1] = c[dy];
        v[3 * dy + 2] = __shfl_down(c[dy], 1);
    }
    __shared__ uint4 sb[4][kConvInTile * 4];   // per wave: 62 pixels x 4 x 16 bytes
    uint4* ws = sb[(threadIdx.x >> 6) & 3];
    if (lane >= 1 && lane <= kConvInTile) {
        uint32_t packed[16];
        #pragma unroll
        for (int ch = 0; ch < 32; ch += 2) {
            float z0 = b[ch], z1 = b[ch + 1];
            #pragma unroll
            for (int t = 0; t < 9; ++t) {
                z0 = fmaf(w[t * 32 + ch], v[t], z0);
                z1 = fmaf(w[t * 32 + ch + 1], v[t], z1);
            }
            packed[ch / 2] = pack_bf16(fmaxf(z0, 0.0f), fmaxf(z1, 0.0f));
        }
        #pragma unroll
        for (int q = 0; q < 4; ++q)
            ws[(lane - 1) * 4 + q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint16_t* o = out + (img + (long)yh * W + tile * kConvInTile) * 32;
    const int npix = min(kConvInTile, W - tile * kConvInTile);
    #pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = q * 64 + lane;               // 16-byte unit: pixel e / 4, channels 8 (e % 4) ..
        if (e / 4 < npix) reinterpret_cast<uint4*>(o)[e] = ws[e];
    }
}

// ---------------------------------------------------------------------------
// 32 -> 32 layer on MFMA: v_mfma_f32_32x32x16_bf16, which on gfx950 moves
// twice the flops per cycle of the 16x16x32 form (measured 2.17 against 1.22
// PFLOP/s on MI355X).  wt: [9][32 co][32 ci] bf16 (tap t = 3*dy + dx).
// mode 0: out = relu?(acc + bias); mode 1: out = acc * (mask > 0)
//
// Operands for tap t and input-channel half kc (ci 16 kc .. 16 kc + 15):
// A[m = co][k] on lane l: co = l % 32, ci = 16 kc + 8 (l / 32) .. + 7 (one
// 16-byte load of wt); B[k][n] on lane l: pixel n = l % 32 of a 32-pixel tile,
// the same 8 channels; D on lane l, register i: co = 8 (i / 4) + 4 (l / 32) +
// i % 4 of pixel l % 32.  Every output accumulates taps 0 .. 8, each with
// kc = 0 then 1, in that order, in the layer-by-layer and the fused kernels
// alike, so they give the same bits.
//
// Activation rows in LDS: [channel group g][pixel] 16-byte entries (g = 8
// channels) with a pitch of 4 (mod 16) entries, so the 32 lanes of a fragment
// read (same g, consecutive pixels) and a staging write (4 pixels x 4 groups
// per 16 lanes) hit distinct banks.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16()
{
    f32x16 z;
    #pragma unroll
    for (int i = 0; i < 16; ++i) z[i] = 0.0f;
    return z;
}

// the 18 weight fragments of a 32 -> 32 layer
__device__ __forceinline__ void load_w32(const uint16_t* __restrict__ wt, int lane, bf16x8 (&wa)[9][2])
{
    #pragma unroll
    for (int tp = 0; tp < 9; ++tp)
        #pragma unroll
        for (int kc = 0; kc < 2; ++kc)
            wa[tp][kc] = as_bf16x8(*reinterpret_cast<const uint4*>(
                wt + ((tp * 32 + (lane & 31)) * 32 + 16 * kc + 8 * (lane >> 5))));
}

// One input row of a 32-pixel tile: f[dx][kc] are its fragments at column
// offset dx - 1; it adds taps 6 .. 8 to A (output row above), 3 .. 5 to Bm
// (same row) and 0 .. 2 to Cm (row below).
__device__ __forceinline__ void mid32_row(const bf16x8 (&wa)[9][2], const bf16x8 (&f)[3][2], f32x16& A, f32x16& Bm,
                                          f32x16& Cm)
{
    #pragma unroll
    for (int dx = 0; dx < 3; ++dx)
        #pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            A = mfma32(wa[6 + dx][kc], f[dx][kc], A);
            Bm = mfma32(wa[3 + dx][kc], f[dx][kc], Bm);
            Cm = mfma32(wa[dx][kc], f[dx][kc], Cm);
        }
}

// fragments of the tile whose dx = 0 pixel is entry p of the row buffer
template <int NPP>
__device__ __forceinline__ void frags32(const uint4* __restrict__ rb, int p, int lane, bf16x8 (&f)[3][2])
{
    #pragma unroll
    for (int dx = 0; dx < 3; ++dx)
        #pragma unroll
        for (int kc = 0; kc < 2; ++kc) f[dx][kc] = as_bf16x8(rb[(2 * kc + (lane >> 5)) * NPP + p + dx]);
}

template <int NW>
struct ConvRowStage {
    static constexpr int NP = 32 * NW + 2;                        // pixels of a staged row (1-pixel halo)
    static constexpr int NPP = NP + (((4 - NP) % 16) + 16) % 16;  // pitch = 4 (mod 16)
    static constexpr int NCH = NP * 4;                            // 16-byte chunks of a row
    static constexpr int NT = 64 * NW;                            // threads
    static constexpr int K = (NCH + NT - 1) / NT;                 // chunks per thread
};

template <int NW>
__device__ __forceinline__ void stage_load(const uint16_t* __restrict__ in, long img, int yy, int H, int W, int x0,
                                           int t, uint4 (&v)[ConvRowStage<NW>::K])
{
    using S = ConvRowStage<NW>;
    const bool rowok = yy >= 0 && yy < H;
    #pragma unroll
    for (int k = 0; k < S::K; ++k) {
        const int e = t + k * S::NT;
        const int x = x0 - 1 + (e >> 2);
        v[k] = make_uint4(0, 0, 0, 0);
        if (rowok && e < S::NCH && x >= 0 && x < W)
            v[k] = *reinterpret_cast<const uint4*>(in + (img + (long)yy * W + x) * 32 + 8 * (e & 3));
    }
}

template <int NW>
__device__ __forceinline__ void stage_store(uint4* __restrict__ buf, int t, const uint4 (&v)[ConvRowStage<NW>::K])
{
    using S = ConvRowStage<NW>;
    #pragma unroll
    for (int k = 0; k < S::K; ++k) {
        const int e = t + k * S::NT;
        if (e < S::NCH) buf[(e & 3) * S::NPP + (e >> 2)] = v[k];
    }
}

// One completed output row of a wave's 32 pixels: relu?(acc + bias) (mode 0)
// or acc * (mask > 0) (mode 1) as bf16, gathered in the wave's LDS tile
// (80-byte pixel pitch: conflict-free 8-byte writes) and written as two
// contiguous 16-byte stores per lane.
template <int MODE>
__device__ __forceinline__ void store32(const f32x16& acc, const float (&bv)[16], const uint16_t* __restrict__ mask,
                                        uint16_t* __restrict__ out, long rowbase, int xw, int W, int lane, int relu,
                                        uint4* __restrict__ tile)
{
    const int n = lane & 31, h = lane >> 5;
    uint2* tw = reinterpret_cast<uint2*>(tile);
    #pragma unroll
    for (int q = 0; q < 4; ++q) {
        float z[4];
        if (MODE == 0) {
            #pragma unroll
            for (int e = 0; e < 4; ++e) {
                z[e] = acc[4 * q + e] + bv[4 * q + e];
                if (relu) z[e] = fmaxf(z[e], 0.0f);
            }
        } else {
            uint2 m = make_uint2(0, 0);
            if (xw + n < W) m = *reinterpret_cast<const uint2*>(mask + (rowbase + xw + n) * 32 + 8 * q + 4 * h);
            const uint16_t mk[4] = {(uint16_t)(m.x & 0xffff), (uint16_t)(m.x >> 16), (uint16_t)(m.y & 0xffff),
                                    (uint16_t)(m.y >> 16)};
            #pragma unroll
            for (int e = 0; e < 4; ++e) z[e] = bf2f(mk[e]) > 0.0f ? acc[4 * q + e] : 0.0f;
        }
        tw[n * 10 + 2 * q + h] = make_uint2(pack_bf16(z[0], z[1]), pack_bf16(z[2], z[3]));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    #pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int e = lane + 64 * k;                 // 16-byte chunk: pixel e / 4, channels 8 (e % 4) ..
        const uint4 v = tile[(e >> 2) * 5 + (e & 3)];
        if (xw + (e >> 2) < W) *reinterpret_cast<uint4*>(out + (rowbase + xw) * 32 + (long)e * 8) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// A workgroup of NW waves takes 32 NW adjacent columns and walks down L output
// rows of one image.  Each input row (the columns and a 1-pixel halo) is
// loaded once per workgroup, two rows ahead of its use, into one of two LDS
// row buffers, and feeds the three output rows that use it through three
// rolling accumulators; the oldest is complete after its third row and is
// stored.  The 18 weight fragments stay in registers.
template <int L, int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void k_conv_mid(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
                                                    const float* __restrict__ bias,
                                                    const uint16_t* __restrict__ mask, uint16_t* __restrict__ out,
                                                    int H, int W, int relu)
{
    using S = ConvRowStage<NW>;
    __shared__ uint4 rows[2][4 * S::NPP];
    __shared__ uint4 tiles[NW][32 * 5];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int x0 = blockIdx.x * 32 * NW;                 // first column of the workgroup
    const int xw = x0 + 32 * wv;                         // first column of this wave
    const int y0 = blockIdx.y * L;
    const long img = (long)blockIdx.z * H * W;
    bf16x8 wa[9][2];
    load_w32(wt, lane, wa);
    float bv[16];
    #pragma unroll
    for (int i = 0; i < 16; ++i) bv[i] = (MODE == 0) ? bias[8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)] : 0.0f;
    f32x16 a0 = zero16(), a1 = zero16(), a2 = zero16();
    // step ir uses input row y0 - 1 + ir from buffer ir & 1
    uint4 nx[S::K], nn[S::K];
    stage_load<NW>(in, img, y0 - 1, H, W, x0, t, nx);
    stage_store<NW>(rows[0], t, nx);
    stage_load<NW>(in, img, y0, H, W, x0, t, nx);
    const int p0 = 32 * wv + (lane & 31);                // entry of the pixel left of this lane's pixel
    auto step = [&](int ir, f32x16& A, f32x16& Bm, f32x16& Cm) {
        __syncthreads();                                 // buffer ir & 1 complete, buffer (ir + 1) & 1 free
        bf16x8 f[3][2];
        frags32<S::NPP>(rows[ir & 1], p0, lane, f);
        if (ir + 2 < L + 2) stage_load<NW>(in, img, y0 + ir + 1, H, W, x0, t, nn);
        Cm = zero16();
        mid32_row(wa, f, A, Bm, Cm);
        const int yo = y0 + ir - 2;                      // the row A completes
        if (ir >= 2 && yo < H) store32<MODE>(A, bv, mask, out, img + (long)yo * W, xw, W, lane, relu, tiles[wv]);
        if (ir + 1 < L + 2) stage_store<NW>(rows[(ir + 1) & 1], t, nx);
        #pragma unroll
        for (int k = 0; k < S::K; ++k) nx[k] = nn[k];
    };
    for (int ir = 0; ir < L + 2; ir += 3) {
        step(ir, a0, a1, a2);
        if (ir + 1 < L + 2) step(ir + 1, a1, a2, a0);
        if (ir + 2 < L + 2) step(ir + 2, a2, a0, a1);
    }
}

// ---------------------------------------------------------------------------
// layer 5: y = xres + b + sum_t sum_ci a4[p + off_t][ci] w5[t][ci]   (fp32 out)
//
// As a 1x1 convolution on MFMA followed by a 3x3 sum on the VALU.  The FP32
// weights are split into w5 = hi + lo (hi = bf16(w5), lo = bf16(w5 - hi),
// about 16 significant bits together) and held as the 18 rows m = 2 t + h of
// a 32-row A operand (rows 18 .. 31 zero), so one v_mfma_f32_32x32x16_bf16
// pair per 32 pixels gives U[p][2 t + h] = sum_ci a4[p][ci] w5_h[t][ci] for
// all 9 taps; then S_h = sum over t = 0 .. 8 in order of U[p + off_t][2 t + h]
// and y = xres + (b + (S_0 + S_1)).  The layer-by-layer and the fused kernels
// both compute exactly this.
// ---------------------------------------------------------------------------
// A fragments (kc = 0, 1) of the split layer-5 weights
__device__ __forceinline__ void load_w5(const float* __restrict__ w, int lane, bf16x8 (&wa)[2])
{
    const int m = lane & 31, t = m >> 1;
    #pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        uint32_t u[4];
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            float v[2];
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int ci = 16 * kc + 8 * (lane >> 5) + 2 * e + q;
                const float wf = m < 18 ? w[t * 32 + ci] : 0.0f;
                const float hi = (float)(__bf16)wf;
                v[q] = (m & 1) ? wf - hi : hi;
            }
            u[e] = pack_bf16(v[0], v[1]);
        }
        wa[kc] = as_bf16x8(make_uint4(u[0], u[1], u[2], u[3]));
    }
}

// the U values (rows m < 18) of D on this lane into a U row [pixel][18]
__device__ __forceinline__ void put_u(const f32x16& U, float* __restrict__ urow, int p, int lane)
{
    #pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (m < 18) urow[p * 18 + m] = U[i];
    }
}

// S_0 + S_1 of the pixel whose left neighbour is entry p of the U rows
// r0 (row above), r1, r2 (row below)
__device__ __forceinline__ float sum_u(const float* __restrict__ r0, const float* __restrict__ r1,
                                       const float* __restrict__ r2, int p)
{
    const float* rr[3] = {r0, r1, r2};
    float s0 = 0.0f, s1 = 0.0f;
    #pragma unroll
    for (int t = 0; t < 9; ++t) {
        const float* u = rr[t / 3] + (p + t % 3) * 18 + 2 * t;
        s0 += u[0];
        s1 += u[1];
    }
    return s0 + s1;
}

// One wave: U for the 32 columns xw - 1 .. xw + 30 of each input row, outputs
// for the 30 columns xw .. xw + 29 of each of L rows (3 U rows in LDS).
template <int L>
__global__ __launch_bounds__(256) void k_conv_out(const uint16_t* __restrict__ a, const float* __restrict__ w,
                                                  float b5, const float* __restrict__ xres, float* __restrict__ y,
                                                  int H, int W)
{
    __shared__ float ur[4][3][32 * 18];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int n = lane & 31, h = lane >> 5;
    const int xw = (blockIdx.x * 4 + wv) * 30;
    const int col = xw - 1 + n;                         // this lane's U column
    const int y0 = blockIdx.y * L;
    const long img = (long)blockIdx.z * H * W;
    bf16x8 wa[2];
    load_w5(w, lane, wa);
    for (int ir = 0; ir < L + 2; ++ir) {
        const int yy = y0 + ir - 1;                     // input row
        bf16x8 f[2];
        #pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < H && col >= 0 && col < W)
                v = *reinterpret_cast<const uint4*>(a + (img + (long)yy * W + col) * 32 + 16 * kc + 8 * h);
            f[kc] = as_bf16x8(v);
        }
        f32x16 U = mfma32(wa[0], f[0], zero16());
        U = mfma32(wa[1], f[1], U);
        put_u(U, ur[wv][ir % 3], n, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int yo = yy - 1;
        if (ir >= 2 && yo < H && lane < 30 && xw + lane < W) {
            const float sv = sum_u(ur[wv][(ir + 1) % 3], ur[wv][(ir + 2) % 3], ur[wv][ir % 3], lane);
            const long q = img + (long)yo * W + xw + lane;
            y[q] = (xres ? xres[q] : 0.0f) + (b5 + sv);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// The whole forward network in one launch (inference; no activation is kept).
//
// A workgroup computes a strip of 16 T computed columns (T tiles of 16
// pixels) over L output rows of one image, walking down the rows.  Each of
// its 4 waves owns layers: wave 0 the x rows and the 32 -> 1 output layer
// (MFMA), wave k = 1, 2, 3 the 32 -> 32 layer k + 1 (MFMA, its 72 VGPRs of
// weight fragments resident); wave 0 also runs the 1 -> 32 layer, on FP32
// MFMA.  They run as a row pipeline: at step s wave 0 writes a1 row
// y0 - 4 + s and wave k writes row y0 - 4 - 2k + s of its layer, each consuming the row its
// producer wrote at step s - 1 from a double-buffered LDS row (bf16,
// [channel group][pixel], 16 bytes per entry; pitch = 4 (mod 16) entries
// keeps the fragment reads and the row writes free of bank conflicts); one
// barrier per step; the output layer completes y row y0 - 12 + s.  Every
// stored value comes from the same operations in the same order as the
// layer-by-layer kernels (k_conv_in, k_conv_mid, k_conv_out: rolling
// accumulators over the three input rows, zero rows and columns outside the
// image), so the result is the same bits.  The strip's two edge columns see
// zero neighbours; that error moves one column inwards per 3x3 layer, so of
// the 16 T computed columns the middle 16 T - 8 are output (strips overlap by
// 8 columns), and the first 12 steps fill the pipeline.
// ---------------------------------------------------------------------------
template <int T>
struct FusedCnn {
    static constexpr int NP = 16 * T + 2;                        // pixels of a row incl. the zero halo
    static constexpr int NPP = NP + (((4 - NP) % 16) + 16) % 16;  // entry pitch of a channel group
    static constexpr int WO = 16 * T - 8;                         // output columns of a strip
};

template <int T>
__global__ __launch_bounds__(256, 2) void k_cnn_fwd(const float* __restrict__ x, const float* __restrict__ w1,
                                                   const float* __restrict__ b1, const uint16_t* __restrict__ w2,
                                                   const float* __restrict__ b2, const uint16_t* __restrict__ w3,
                                                   const float* __restrict__ b3, const uint16_t* __restrict__ w4,
                                                   const float* __restrict__ b4, const float* __restrict__ w5, float b5,
                                                   float* __restrict__ y, int H, int W, int nstrips, long total,
                                                   long per)
{
    using F = FusedCnn<T>;
    static_assert(T % 2 == 0, "32-pixel tiles of the 32 -> 32 layers");
    constexpr int NPP = F::NPP, NP = F::NP;
    __shared__ uint4 act[4][2][4 * NPP];     // rows of a1 .. a4
    __shared__ float xr[4][NP];              // x ring: rows y - 1, y, y + 1 of the a1 row and the next
    __shared__ float ur[3][NP * 18];         // U rows of the output layer (wave 0)
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    // The workgroup's rows: [blockIdx.x per, + per) of the (image, strip, row)
    // sequence, walked as segments of consecutive rows of one strip (a new
    // segment refills the pipeline).  Every wave runs the same segments.
    long i0 = (long)blockIdx.x * per;
    const long i1 = i0 + per < total ? i0 + per : total;
    int c0 = 0, y0 = 0, L = 0;               // first computed column, first row, rows of the segment
    long img = 0;
    auto next_segment = [&]() {
        if (i0 >= i1) return false;
        const long si = i0 / H;              // image * nstrips + strip
        y0 = (int)(i0 - si * H);
        L = (int)(i1 - i0 < (long)(H - y0) ? i1 - i0 : (long)(H - y0));
        c0 = (int)(si % nstrips) * F::WO - 4;
        img = (si / nstrips) * (long)H * W;
        i0 += L;
        __syncthreads();                     // the previous segment's last reads are done
        for (int i = t; i < 4 * 2 * 4 * NPP; i += 256) (&act[0][0][0])[i] = make_uint4(0, 0, 0, 0);
        return true;
    };

    // The 1 -> 32 layer on wave 0's matrix core, whose SIMD the 32 -> 32 layers
    // leave idle: per 32-pixel tile five v_mfma_f32_32x32x2f32 add taps 2 i
    // and 2 i + 1 to an accumulator that starts at the bias (tap 9 has weight
    // 0).  This is k_conv_in's arithmetic (z = b; z = fma(w[t], v[t], z) for
    // t = 0..8): the FP32 MFMA adds its two products in order with one
    // rounding each, the same bits as the FMA chain.  A lane holds w1[tap
    // 2 i + l / 32][co = l % 32] and the x value of that tap at pixel l % 32;
    // D is laid out as the 32 -> 32 layers' (channel 8 (r / 4) + 4 (l / 32) +
    // r % 4 of pixel l % 32 in register r).  At step s the a1 row y0 - 4 + s
    // uses x rows y0 - 5 + s .. y0 - 3 + s, in ring slots s % 4 .. (s + 2) % 4;
    // x row y0 - 2 + s goes into slot (s + 3) % 4 for the next step.
    if (wv == 0) {
        // ---- x rows, the 1 -> 32 layer and 32 -> 1 (1x1 MFMA + 3x3 sum, as k_conv_out) ----
        float wa1[5], ba1[16];
        {
            const int co = lane & 31, hk = lane >> 5;
            #pragma unroll
            for (int i = 0; i < 5; ++i) wa1[i] = 2 * i + hk < 9 ? w1[(2 * i + hk) * 32 + co] : 0.0f;
            #pragma unroll
            for (int r = 0; r < 16; ++r) ba1[r] = b1[8 * (r >> 2) + 4 * hk + (r & 3)];
        }
        auto conv_in_row = [&](int s) {
            const int r1 = y0 - 4 + s;
            uint2* a1w = reinterpret_cast<uint2*>(act[0][s & 1]);
            const int n = lane & 31, hk = lane >> 5;
            #pragma unroll
            for (int j = 0; j < T / 2; ++j) {
                const int pp = 32 * j + n;
                f32x16 z;
                #pragma unroll
                for (int r = 0; r < 16; ++r) z[r] = ba1[r];
                #pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int tp = 2 * i + hk;           // tap 3 dy + dx
                    const float v = tp < 9 ? xr[(s + tp / 3) & 3][pp + tp % 3] : 0.0f;
                    z = __builtin_amdgcn_mfma_f32_32x32x2f32(wa1[i], v, z, 0, 0, 0);
                }
                const int c = c0 + pp;
                const bool ok = r1 >= 0 && r1 < H && c >= 0 && c < W;
                #pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint2 ov = make_uint2(0, 0);
                    if (ok) {
                        ov.x = pack_bf16(fmaxf(z[4 * q], 0.0f), fmaxf(z[4 * q + 1], 0.0f));
                        ov.y = pack_bf16(fmaxf(z[4 * q + 2], 0.0f), fmaxf(z[4 * q + 3], 0.0f));
                    }
                    a1w[(q * NPP + pp + 1) * 2 + hk] = ov;
                }
            }
        };
        bf16x8 wa5[2];
        load_w5(w5, lane, wa5);
        // x rows: lane l holds columns c0 - 1 + l + 64 k of a row (zero outside the
        // image); each row is loaded one step before it enters the LDS ring, and
        // the residual values of the y row (columns c0 + l + 64 k) one step before
        // they are added
        constexpr int KX = (NP + 63) / 64;
        constexpr int KY = (16 * T + 63) / 64;
        auto fetch_x = [&](int row, float (&v)[KX]) {
            const bool rok = row >= 0 && row < H;
            #pragma unroll
            for (int k = 0; k < KX; ++k) {
                const int i = lane + 64 * k;
                const int c = c0 - 1 + i;
                v[k] = (rok && i < NP && c >= 0 && c < W) ? x[img + (long)row * W + c] : 0.0f;
            }
        };
        auto put_x = [&](int slot, const float (&v)[KX]) {
            #pragma unroll
            for (int k = 0; k < KX; ++k)
                if (lane + 64 * k < NP) xr[slot][lane + 64 * k] = v[k];
        };
        auto fetch_res = [&](int row, float (&v)[KY]) {
            #pragma unroll
            for (int k = 0; k < KY; ++k) {
                const int c = c0 + lane + 64 * k;
                v[k] = (row >= 0 && row < H && c >= 0 && c < W) ? x[img + (long)row * W + c] : 0.0f;
            }
        };
        float xn[KX], xres[KY];
        while (next_segment()) {
        for (int i = lane; i < 3 * NP * 18; i += 64) (&ur[0][0])[i] = 0.0f;
        fetch_x(y0 - 5, xn);
        put_x(0, xn);
        fetch_x(y0 - 4, xn);
        put_x(1, xn);
        fetch_x(y0 - 3, xn);
        put_x(2, xn);
        fetch_x(y0 - 2, xn);
        fetch_res(y0 - 12, xres);
        __syncthreads();
        // one pipeline step (U row y0 - 13 + s + dy of the output layer in ring
        // slot (s + 1 + dy) % 3; PH = s % 3)
        auto step = [&](auto ph, int s) {
            constexpr int PH = decltype(ph)::value;
            put_x((s + 3) & 3, xn);                     // x row y0 - 2 + s, for step s + 1
            fetch_x(y0 - 1 + s, xn);                    // for the step after
            conv_in_row(s);
            // output layer: U row of the a4 row y0 - 11 + s written at step s - 1,
            // then the 3x3 sum of y row y0 - 12 + s
            const uint4* a4r = act[3][(s + 1) & 1];
            const int n = lane & 31, h = lane >> 5;
            #pragma unroll
            for (int j = 0; j < T / 2; ++j) {
                const bf16x8 f0 = as_bf16x8(a4r[h * NPP + 32 * j + n + 1]);
                const bf16x8 f1 = as_bf16x8(a4r[(2 + h) * NPP + 32 * j + n + 1]);
                f32x16 U = mfma32(wa5[0], f0, zero16());
                U = mfma32(wa5[1], f1, U);
                put_u(U, ur[PH], 32 * j + n + 1, lane);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const int ry = y0 - 12 + s;
            #pragma unroll
            for (int k = 0; k < KY; ++k) {
                const int pp = lane + 64 * k;           // computed column c0 + pp
                if (s >= 12 && ry < H && pp >= 4 && pp < 4 + F::WO && c0 + pp < W) {
                    const float sv = sum_u(ur[(PH + 1) % 3], ur[(PH + 2) % 3], ur[PH], pp);
                    y[img + (long)ry * W + c0 + pp] = xres[k] + (b5 + sv);
                }
            }
            fetch_res(ry + 1, xres);                    // for the next step
            __syncthreads();
        };
        for (int s = 0; s < L + 12; s += 3) {
            step(std::integral_constant<int, 0>{}, s);
            if (s + 1 < L + 12) step(std::integral_constant<int, 1>{}, s + 1);
            if (s + 2 < L + 12) step(std::integral_constant<int, 2>{}, s + 2);
        }
        }
    } else {
        // ---- the 32 -> 32 layer wv + 1: relu(conv(a_wv) + b), T / 2 tiles of 32 pixels ----
        constexpr int T2 = T / 2;
        bf16x8 wa[9][2];
        load_w32(wv == 1 ? w2 : (wv == 2 ? w3 : w4), lane, wa);
        const float* bias = wv == 1 ? b2 : (wv == 2 ? b3 : b4);
        float bv[16];
        #pragma unroll
        for (int i = 0; i < 16; ++i) bv[i] = bias[8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)];
        const int n = lane & 31, h = lane >> 5;
        f32x16 m0[T2], m1[T2], m2[T2];
        while (next_segment()) {
        #pragma unroll
        for (int j = 0; j < T2; ++j) m0[j] = m1[j] = m2[j] = zero16();
        __syncthreads();
        // one pipeline step (accumulator roles rotate as in wave 0's steps)
        auto step = [&](int s, f32x16 (&A)[T2], f32x16 (&Bm)[T2], f32x16 (&Cm)[T2]) {
            const uint4* in = act[wv - 1][(s + 1) & 1];   // written at step s - 1
            uint2* ow = reinterpret_cast<uint2*>(act[wv][s & 1]);
            const int ro = y0 - 4 - 2 * wv + s;            // the row this step completes
            const bool rok = ro >= 0 && ro < H;
            // every tile's MFMAs first, then the completed rows
            #pragma unroll
            for (int j = 0; j < T2; ++j) {
                bf16x8 f[3][2];
                frags32<NPP>(in, 32 * j + n, lane, f);
                Cm[j] = zero16();
                mid32_row(wa, f, A[j], Bm[j], Cm[j]);
            }
            #pragma unroll
            for (int j = 0; j < T2; ++j) {
                // this lane holds channels 8 q + 4 h .. + 3 of pixel n of the tile
                const int c = c0 + 32 * j + n;
                const bool ok = rok && c >= 0 && c < W;
                #pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float z[4];
                    #pragma unroll
                    for (int e = 0; e < 4; ++e) z[e] = fmaxf(A[j][4 * q + e] + bv[4 * q + e], 0.0f);
                    uint2 ov = make_uint2(0, 0);
                    if (ok) { ov.x = pack_bf16(z[0], z[1]); ov.y = pack_bf16(z[2], z[3]); }
                    ow[(q * NPP + 32 * j + n + 1) * 2 + h] = ov;
                }
            }
            __syncthreads();
        };
        for (int s = 0; s < L + 12; s += 3) {
            step(s, m0, m1, m2);
            if (s + 1 < L + 12) step(s + 1, m1, m2, m0);
            if (s + 2 < L + 12) step(s + 2, m2, m0, m1);
        }
        }
    }
}

// layer-5 data gradient: dz4[p][ci] = (a4[p][ci] > 0) * sum_t dy[p - off_t] w5[t][ci]
__global__ __launch_bounds__(256) void k_out_dgrad(const float* __restrict__ dy, const float* __restrict__ w,
                                                   const uint16_t* __restrict__ a4, uint16_t* __restrict__ dz,
                                                   int B, int H, int W)
{
    __shared__ float sw[288];
    for (int i = threadIdx.x; i < 288; i += blockDim.x) sw[i] = w[i];
    __syncthreads();
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long npx = (long)B * H * W;
    if (p >= npx) return;
    const int xw = (int)(p % W), yh = (int)((p / W) % H);
    const long base = p - (long)yh * W - xw;
    float g[9];
    #pragma unroll
    for (int t = 0; t < 9; ++t) {
        // output pixel q = p - off_t receives a4[p] through tap t
        const int yy = yh - (t / 3 - 1), xx = xw - (t % 3 - 1);
        g[t] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? dy[base + (long)yy * W + xx] : 0.0f;
    }
    const uint4* am = reinterpret_cast<const uint4*>(a4 + p * 32);
    uint4* o = reinterpret_cast<uint4*>(dz + p * 32);
    #pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 m = am[q];
        const uint32_t mu[4] = {m.x, m.y, m.z, m.w};
        uint32_t r[4];
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            float z[2];
            #pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = 8 * q + 2 * e + h;
                float s = 0.0f;
                #pragma unroll
                for (int t = 0; t < 9; ++t) s = fmaf(sw[t * 32 + c], g[t], s);
                const uint16_t mk = h ? (uint16_t)(mu[e] >> 16) : (uint16_t)(mu[e] & 0xffff);
                z[h] = bf2f(mk) > 0.0f ? s : 0.0f;
            }
            r[e] = pack_bf16(z[0], z[1]);
        }
        o[q] = make_uint4(r[0], r[1], r[2], r[3]);
    }
}

// ---------------------------------------------------------------------------
// weight gradients of the narrow layers (reduction over pixels, VALU):
// mode 0 (layer 1): part[t][c] = sum_p x[p + off_t] dz1[p][c], part[9][c] = sum_p dz1[p][c]
// mode 1 (layer 5): part[t][c] = sum_p dy[p] a4[p + off_t][c], part[9][0] = sum_p dy[p]
// one workgroup: 8 groups of 32 lanes (lane = channel c), partials per workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad_narrow(const float* __restrict__ img, const uint16_t* __restrict__ act,
                                                      float* __restrict__ part, int B, int H, int W, int mode)
{
    __shared__ float red[8][10 * 32];
    const int c = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const long npx = (long)B * H * W;
    float acc[10];
    #pragma unroll
    for (int t = 0; t < 10; ++t) acc[t] = 0.0f;
    for (long p = (long)blockIdx.x * 8 + grp; p < npx; p += (long)gridDim.x * 8) {
        const int xw = (int)(p % W), yh = (int)((p / W) % H);
        const long base = p - (long)yh * W - xw;
        if (mode == 0) {
            const float d = bf2f(act[p * 32 + c]);
            #pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = yh + t / 3 - 1, xx = xw + t % 3 - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) acc[t] = fmaf(img[base + (long)yy * W + xx], d, acc[t]);
            }
            acc[9] += d;
        } else {
            const float d = img[p];
            #pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = yh + t / 3 - 1, xx = xw + t % 3 - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    acc[t] = fmaf(d, bf2f(act[(base + (long)yy * W + xx) * 32 + c]), acc[t]);
            }
            acc[9] += d;
        }
    }
    #pragma unroll
    for (int t = 0; t < 10; ++t) red[grp][t * 32 + c] = acc[t];
    __syncthreads();
    for (int i = threadIdx.x; i < 320; i += blockDim.x) {
        float s = 0.0f;
        #pragma unroll
        for (int gidx = 0; gidx < 8; ++gidx) s += red[gidx][i];
        part[(long)blockIdx.x * 320 + i] = s;
    }
}

// bias gradient of a 32-channel NHWC tensor: part[wg][c] = sum_p dz[p][c]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dz, float* __restrict__ part, long npx)
{
    __shared__ float red[8][32];
    const int c = threadIdx.x & 31, grp = threadIdx.x >> 5;
    float s = 0.0f;
    for (long p = (long)blockIdx.x * 8 + grp; p < npx; p += (long)gridDim.x * 8) s += bf2f(dz[p * 32 + c]);
    red[grp][c] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.0f;
        for (int gidx = 0; gidx < 8; ++gidx) t += red[gidx][threadIdx.x];
        part[(long)blockIdx.x * 32 + threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// 32 -> 32 weight gradient on MFMA:
//   dW[t][co][ci] = sum_p dz[p][co] in[p + off_t][ci]
// GEMM with M = co, N = ci, K = pixels.  Each wave takes 32-pixel row chunks;
// per chunk it transposes dz into LDS as [co][px] and the 3 x 34 input pixels
// into 9 shifted copies [t][ci][px], so every fragment is one aligned 16-byte
// LDS read.  All waves of a workgroup run the same number of chunk rounds
// (out-of-range chunks contribute zeros) so block barriers are uniform.
// ---------------------------------------------------------------------------
constexpr int WG_WAVES = 4;
constexpr int LDS_PER_WAVE = 32 * 32 + 9 * 32 * 32;   // bf16 elements

__global__ __launch_bounds__(256) void k_wgrad_mid(const uint16_t* __restrict__ in, const uint16_t* __restrict__ dz,
                                                   float* __restrict__ part, int B, int H, int W)
{
    extern __shared__ uint16_t lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    uint16_t* dzT = lds + wv * LDS_PER_WAVE;        // [32 co][32 px]
    uint16_t* inS = dzT + 32 * 32;                  // [9 t][32 ci][32 px]
    const int chunks_per_row = (W + 31) / 32;
    const long nchunk = (long)B * H * chunks_per_row;
    const long stride = (long)gridDim.x * WG_WAVES;
    f32x4 acc[9][2][2];
    #pragma unroll
    for (int t = 0; t < 9; ++t)
        #pragma unroll
        for (int a = 0; a < 2; ++a)
            #pragma unroll
            for (int b = 0; b < 2; ++b) acc[t][a][b] = f32x4{0, 0, 0, 0};
    const long rounds = (nchunk + stride - 1) / stride;
    for (long rd = 0; rd < rounds; ++rd) {
        const long ch = rd * stride + (long)blockIdx.x * WG_WAVES + wv;
        const bool live = ch < nchunk;
        long bimg = 0; int yh = 0, x0 = 0;
        if (live) {
            x0 = (int)(ch % chunks_per_row) * 32;
            const long by = ch / chunks_per_row;
            yh = (int)(by % H);
            bimg = by / H;
        }
        // dz chunk -> dzT[co][px]
        for (int j = lane; j < 128; j += 64) {
            const int px = j >> 2, c8 = (j & 3) * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (live && x0 + px < W) v = *reinterpret_cast<const uint4*>(dz + ((bimg * H + yh) * W + x0 + px) * 32 + c8);
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
            #pragma unroll
            for (int e = 0; e < 8; ++e)
                dzT[(c8 + e) * 32 + px] = (uint16_t)((e & 1) ? (u[e >> 1] >> 16) : (u[e >> 1] & 0xffff));
        }
        // input rows yh-1..yh+1, pixels x0-1..x0+32 -> inS[t][ci][px_out], px_out = px_in - dx
        for (int j = lane; j < 3 * 34 * 4; j += 64) {
            const int r = j / 136, rem = j % 136, pin = rem >> 2, c8 = (rem & 3) * 8;
            const int yy = yh + r - 1, xx = x0 + pin - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (live && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *reinterpret_cast<const uint4*>(in + ((bimg * H + yy) * W + xx) * 32 + c8);
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
            #pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int po = pin - dx;
                if (po < 0 || po >= 32) continue;
                uint16_t* dst = inS + ((r * 3 + dx) * 32 + c8) * 32 + po;
                #pragma unroll
                for (int e = 0; e < 8; ++e)
                    dst[e * 32] = (uint16_t)((e & 1) ? (u[e >> 1] >> 16) : (u[e >> 1] & 0xffff));
            }
        }
        __syncthreads();
        bf16x8 af[2];
        #pragma unroll
        for (int mt = 0; mt < 2; ++mt)
            af[mt] = as_bf16x8(*reinterpret_cast<const uint4*>(dzT + (16 * mt + r16) * 32 + 8 * g4));
        #pragma unroll
        for (int t = 0; t < 9; ++t) {
            #pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const bf16x8 bfrag = as_bf16x8(*reinterpret_cast<const uint4*>(inS + (t * 32 + 16 * nt + r16) * 32 + 8 * g4));
                acc[t][0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bfrag, acc[t][0][nt], 0, 0, 0);
                acc[t][1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bfrag, acc[t][1][nt], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // partial of this wave: part[wave][t][co][ci], D lane layout n = ci = 16 nt + r16, m = co = 16 mt + 4 g4 + i
    float* dst = part + ((long)blockIdx.x * WG_WAVES + wv) * 9216;
    #pragma unroll
    for (int t = 0; t < 9; ++t)
        #pragma unroll
        for (int mt = 0; mt < 2; ++mt)
            #pragma unroll
            for (int nt = 0; nt < 2; ++nt)
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    dst[(t * 32 + 16 * mt + 4 * g4 + i) * 32 + 16 * nt + r16] = acc[t][mt][nt][i];
}

// out[i] (+)= sum_r part[r][i]
__global__ void k_reduce_rows(const float* __restrict__ part, float* __restrict__ out, int nrows, int n, int accumulate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int r = 0; r < nrows; ++r) s += part[(long)r * n + i];
    out[i] = accumulate ? out[i] + s : s;
}

constexpr int kNConvMid = 5;

int grid_for(long work, int per_block, int cap)
{
    long g = (work + per_block - 1) / per_block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

} // namespace

void conv_in(const float* x, const float* w, const float* b, uint16_t* out, int B, int H, int W, hipStream_t st)
{
    RN_REQUIRE(B > 0 && H > 0 && W > 0, "conv_in: bad batch or image size");
    const int tiles = cdiv(W, kConvInTile);
    const long nwaves = (long)B * H * tiles;
    RN_REQUIRE(nwaves / 4 < 0x7fffffffL, "conv_in: batch too large for one launch");
    hipLaunchKernelGGL(k_conv_in, dim3((unsigned)cdiv(nwaves * 64, 256)), dim3(256), 0, st, x, w, b, out, H, W, tiles,
                       nwaves);
    check_launch("conv_in");
}

void conv_mid(const uint16_t* in, const uint16_t* wt, const float* bias, const uint16_t* mask, uint16_t* out,
              int B, int H, int W, bool relu, hipStream_t st)
{
    RN_REQUIRE(mask != nullptr || bias != nullptr, "conv_mid needs a bias (forward) or a mask (data gradient)");
    RN_REQUIRE(B > 0 && H > 0 && W > 0 && B <= 65535, "conv_mid: bad batch or image size");
    const int rl = relu ? 1 : 0;
    // candidates: workgroups of 2, 4 or 8 waves of 32 columns, 16 or 32 rows
    auto launch = [&](int c) {
        auto go = [&](auto kern, int nw, int rows) {
            const dim3 grd((unsigned)cdiv(W, 32 * nw), (unsigned)cdiv(H, rows), (unsigned)B);
            hipLaunchKernelGGL(kern, grd, dim3(64 * nw), 0, st, in, wt, bias, mask, out, H, W, rl);
        };
        if (c == 0) mask ? go(k_conv_mid<16, 1, 4>, 4, 16) : go(k_conv_mid<16, 0, 4>, 4, 16);
        else if (c == 1) mask ? go(k_conv_mid<16, 1, 2>, 2, 16) : go(k_conv_mid<16, 0, 2>, 2, 16);
        else if (c == 2) mask ? go(k_conv_mid<32, 1, 4>, 4, 32) : go(k_conv_mid<32, 0, 4>, 4, 32);
        else if (c == 3) mask ? go(k_conv_mid<32, 1, 2>, 2, 32) : go(k_conv_mid<32, 0, 2>, 2, 32);
        else mask ? go(k_conv_mid<16, 1, 8>, 8, 16) : go(k_conv_mid<16, 0, 8>, 8, 16);
        check_launch("conv_mid");
    };
    const int c = tune_choose({TK_CONV_MID, B, H, W, mask ? 1 : 0, rl}, kNConvMid, launch, st, true);
    launch(c);
}

void conv_out(const uint16_t* a4, const float* w5, float b5, const float* xres, float* y, int B, int H, int W,
              hipStream_t st)
{
    RN_REQUIRE(B > 0 && H > 0 && W > 0 && B <= 65535, "conv_out: bad batch or image size");
    constexpr int kRows = 16;
    const dim3 grd((unsigned)cdiv(W, 4 * 30), (unsigned)cdiv(H, kRows), (unsigned)B);
    hipLaunchKernelGGL(k_conv_out<kRows>, grd, dim3(256), 0, st, a4, w5, b5, xres, y, H, W);
    check_launch("conv_out");
}

void cnn_forward(const float* x, const float* w1, const float* b1, const uint16_t* w2, const float* b2,
                 const uint16_t* w3, const float* b3, const uint16_t* w4, const float* b4, const float* w5, float b5,
                 float* y, int B, int H, int W, hipStream_t st)
{
    RN_REQUIRE(B > 0 && H > 0 && W > 0, "cnn_forward: bad batch or image size");
    int dev = 0, ncu = 0;
    RN_HIP_CHECK(hipGetDevice(&dev));
    RN_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // candidates: strips of T = 4 or 6 tiles of 16 columns; the (image, strip,
    // row) sequence cut into equal shares for 2 workgroups per CU (one round),
    // or into shares of 128 or 64 rows
    auto launch = [&](int c) {
        auto go = [&](auto kern, int wo, long per_rows) {
            const int nstrips = cdiv(W, wo);
            const long total = (long)B * nstrips * H;
            const long per = per_rows > 0 ? per_rows : (total + 2L * ncu - 1) / (2L * ncu);
            const long nwg = (total + per - 1) / per;
            RN_REQUIRE(nwg <= 0x7fffffffL, "cnn_forward: grid too large");
            hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(256), 0, st, x, w1, b1, w2, b2, w3, b3, w4, b4, w5, b5,
                               y, H, W, nstrips, total, per);
        };
        if (c == 0) go(k_cnn_fwd<4>, FusedCnn<4>::WO, 0);
        else if (c == 1) go(k_cnn_fwd<6>, FusedCnn<6>::WO, 0);
        else if (c == 2) go(k_cnn_fwd<4>, FusedCnn<4>::WO, 128);
        else go(k_cnn_fwd<4>, FusedCnn<4>::WO, 64);
        check_launch("cnn_forward");
    };
    const int c = tune_choose({TK_CNN_FWD, B, H, W, 0, 0}, 4, launch, st, true);
    launch(c);
}

void conv_out_dgrad(const float* dy, const float* w5, const uint16_t* a4, uint16_t* dz4, int B, int H, int W,
                    hipStream_t st)
{
    const long npx = (long)B * H * W;
    hipLaunchKernelGGL(k_out_dgrad, dim3(cdiv(npx, 256)), dim3(256), 0, st, dy, w5, a4, dz4, B, H, W);
    check_launch("conv_out_dgrad");
}

void wgrad_narrow(const float* img, const uint16_t* act, float* grad, int B, int H, int W, int mode, DevBuf& work,
                  hipStream_t st)
{
    const long npx = (long)B * H * W;
    const int blocks = grid_for(npx, 8 * 64, 1024);
    work.ensure((size_t)blocks * 320 * sizeof(float));
    hipLaunchKernelGGL(k_wgrad_narrow, dim3(blocks), dim3(256), 0, st, img, act, work.as<float>(), B, H, W, mode);
    check_launch("wgrad_narrow");
    hipLaunchKernelGGL(k_reduce_rows, dim3(cdiv(320, 256)), dim3(256), 0, st, work.as<float>(), grad, blocks, 320, 0);
    check_launch("reduce_rows");
}

void bias_grad(const uint16_t* dz, float* grad, long npx, DevBuf& work, hipStream_t st)
{
    const int blocks = grid_for(npx, 8 * 64, 1024);
    work.ensure((size_t)blocks * 32 * sizeof(float));
    hipLaunchKernelGGL(k_bias_grad, dim3(blocks), dim3(256), 0, st, dz, work.as<float>(), npx);
    check_launch("bias_grad");
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(32), 0, st, work.as<float>(), grad, blocks, 32, 0);
    check_launch("reduce_rows");
}

void wgrad_mid(const uint16_t* in, const uint16_t* dz, float* grad, int B, int H, int W, DevBuf& work, hipStream_t st)
{
    const long nchunk = (long)B * H * ((W + 31) / 32);
    const int blocks = grid_for(nchunk, WG_WAVES * 8, 1024);
    work.ensure((size_t)blocks * WG_WAVES * 9216 * sizeof(float));
    const size_t lds = (size_t)WG_WAVES * LDS_PER_WAVE * sizeof(uint16_t);
    hipLaunchKernelGGL(k_wgrad_mid, dim3(blocks), dim3(256), lds, st, in, dz, work.as<float>(), B, H, W);
    check_launch("wgrad_mid");
    hipLaunchKernelGGL(k_reduce_rows, dim3(cdiv(9216, 256)), dim3(256), 0, st, work.as<float>(), grad,
                       blocks * WG_WAVES, 9216, 0);
    check_launch("reduce_rows");
}

} // namespace rn
