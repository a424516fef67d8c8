// Refinement-indicator MLP on gfx950: 128 -> 64 -> 64 -> 1 with GELU,
// BF16 GEMM operands on MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation, mean-squared-error training with Adam.  No counterpart in
// the reference; its labels come from the gradient error estimator.
//
// Parameters (FP32 master copy, one flat vector of 12,481):
//   W1[64][128] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
// Rounded to BF16 (nearest even) for the GEMMs: inputs, W1, W2, W3, both
// hidden activations and the two back-propagated deltas.  Biases, bias
// gradients, pre-activations, GELU derivatives and the output stay FP32.
//
// A workgroup (4 wavefronts) takes tiles of R rows; weights are staged in
// LDS once per (persistent) workgroup.  Each 16x16 output block of a GEMM
// is one wavefront's MFMA accumulator.  Training: every workgroup keeps its
// weight-gradient blocks in accumulators across its tiles and writes one
// partial gradient; a second kernel sums the partials in a fixed order
// (deterministic) and applies Adam.
#include "drf_internal.hpp"
#include "mlp.hpp"

namespace drf {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NI = kMlpIn, NH = kMlpHidden;
constexpr int LX = NI + 8; // LDS row pitch (bf16 elements) of X and W1
constexpr int LH = NH + 8; // LDS row pitch of W2, activations and deltas
constexpr int OW1 = 0, OB1 = NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH, NP = OB3 + 1;
static_assert(NP == kMlpParams, "parameter layout");
constexpr float kC = 0.7978845608028654f; // sqrt(2/pi)
constexpr float kA = 0.044715f;

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// GELU, tanh form (Hendrycks & Gimpel): 0.5 z (1 + tanh(u)) with
// u = sqrt(2/pi) (z + 0.044715 z^3), evaluated as z sigma(2u) since
// 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u)): one exp2 and one reciprocal.
// The derivative is sigma + 2 z sigma (1 - sigma) u'(z).
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kG1 = -2.0f * kLog2e * kC, kG3 = kG1 * kA; // -2u log2(e) = z (kG1 + kG3 z^2)
__device__ __forceinline__ float gelu_sigma(float z)
{
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(kG3, z * z, kG1));
    return __builtin_amdgcn_rcpf(1.0f + e); // e = inf gives 0
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }
__device__ __forceinline__ float gelu_d(float z)
{
    const float sg = gelu_sigma(z);
    return sg + 2.0f * z * sg * (1.0f - sg) * kC * (1.0f + 3.0f * kA * z * z);
}

// acc += A[m0:m0+16][0:K] * B[0:K][n0:n0+16]; A(m,k) = a[m*am + k*ak],
// B(k,n) = b[k*bk + n*bn], BF16 in LDS.  Lane l supplies A rows m0 + l%16
// and B columns n0 + l%16 at k = 4(l/16) .. 4(l/16)+3 of each 16-wide step;
// it receives C rows m0 + 4(l/16) + t, column n0 + l%16, in acc[t].  When
// an operand's k index is contiguous (ak == 1 / bk == 1: ACONT / BCONT) its
// four values are one 8-byte LDS read (row pitches are multiples of 8 bytes).
template <int K, bool ACONT, bool BCONT>
__device__ __forceinline__ f32x4 mfma_block(const uint16_t *a, int am, int ak, const uint16_t *b,
                                            int bk, int bn, int m0, int n0, f32x4 acc)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
        if (ACONT) {
            av = *reinterpret_cast<const bf16x4 *>(a + (m0 + r) * am + k0 + kq);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                av[t] = (short)a[(m0 + r) * am + (k0 + kq + t) * ak];
        }
        if (BCONT) {
            bv = *reinterpret_cast<const bf16x4 *>(b + (n0 + r) * bn + k0 + kq);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                bv[t] = (short)b[(k0 + kq + t) * bk + (n0 + r) * bn];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

template <int R, bool TRAIN> struct alignas(16) Smem {
    uint16_t x[R][LX];
    uint16_t w1[NH][LX];
    uint16_t w2[NH][LH];
    uint16_t a1[R][LH], a2[R][LH];
    float w3[NH], b1[NH], b2[NH];
    float pred[R];
    // training only (size 1 otherwise)
    uint16_t d2[TRAIN ? R : 1][LH], d1[TRAIN ? R : 1][LH];
    float g1[TRAIN ? R : 1][NH], g2[TRAIN ? R : 1][NH];
    float d3[TRAIN ? R : 1];
    float red[TRAIN ? 4 * NH : 1];
};

// Weights to LDS: W1, W2 as BF16 (already rounded in wq), W3 as the float
// value of its BF16, biases FP32 from the master copy.
template <int R, bool TRAIN>
__device__ void stage_weights(Smem<R, TRAIN> &s, const float *__restrict__ w,
                              const uint16_t *__restrict__ wq)
{
    for (int i = threadIdx.x; i < NH * NI; i += blockDim.x)
        s.w1[i / NI][i % NI] = wq[i];
    for (int i = threadIdx.x; i < NH * NH; i += blockDim.x)
        s.w2[i / NH][i % NH] = wq[NH * NI + i];
    for (int i = threadIdx.x; i < NH; i += blockDim.x) {
        s.w3[i] = bf2f(wq[NH * NI + NH * NH + i]);
        s.b1[i] = w[OB1 + i];
        s.b2[i] = w[OB2 + i];
    }
}

// Rows [r0, r0+R) of X (BF16, 128 per row, 16-byte aligned) to LDS; rows at
// or beyond n are zero.
template <int R, bool TRAIN>
__device__ void load_tile(Smem<R, TRAIN> &s, const uint16_t *__restrict__ X, int64_t n, int64_t r0)
{
    constexpr int VPR = NI / 8; // 16-byte vectors per row
    for (int v = threadIdx.x; v < R * VPR; v += blockDim.x) {
        const int r = v / VPR, c = (v % VPR) * 8;
        uint4 val = make_uint4(0, 0, 0, 0);
        if (r0 + r < n)
            val = *reinterpret_cast<const uint4 *>(X + (r0 + r) * NI + c);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&s.x[r][c]);
        dst[0] = val.x;
        dst[1] = val.y;
        dst[2] = val.z;
        dst[3] = val.w;
    }
}

// Forward pass of the staged tile: a1, a2 (BF16), pred (FP32); with TRAIN
// also the GELU derivatives g1, g2.
template <int R, bool TRAIN>
__device__ void forward_tile(Smem<R, TRAIN> &s, float b3)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NB = (R / 16) * (NH / 16);
    for (int blk = w; blk < NB; blk += 4) {
        const int m0 = (blk / (NH / 16)) * 16, n0 = (blk % (NH / 16)) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_block<NI, true, true>(&s.x[0][0], LX, 1, &s.w1[0][0], 1, LX, m0, n0, acc);
        const int col = n0 + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = m0 + (lane >> 4) * 4 + t;
            const float z = acc[t] + s.b1[col];
            s.a1[row][col] = f2bf(gelu(z));
            if (TRAIN)
                s.g1[row][col] = gelu_d(z);
        }
    }
    __syncthreads();
    for (int blk = w; blk < NB; blk += 4) {
        const int m0 = (blk / (NH / 16)) * 16, n0 = (blk % (NH / 16)) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_block<NH, true, true>(&s.a1[0][0], LH, 1, &s.w2[0][0], 1, LH, m0, n0, acc);
        const int col = n0 + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = m0 + (lane >> 4) * 4 + t;
            const float z = acc[t] + s.b2[col];
            s.a2[row][col] = f2bf(gelu(z));
            if (TRAIN)
                s.g2[row][col] = gelu_d(z);
        }
    }
    __syncthreads();
    // output layer: TPR lanes per row, each a contiguous chunk of the 64 inputs
    constexpr int TPR = 256 / R, CH = NH / TPR;
    const int row = threadIdx.x / TPR, part = threadIdx.x % TPR;
    float p = 0.f;
#pragma unroll
    for (int i = 0; i < CH; ++i)
        p += s.w3[part * CH + i] * bf2f(s.a2[row][part * CH + i]);
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1)
        p += __shfl_xor(p, o);
    if (part == 0)
        s.pred[row] = p + b3;
    __syncthreads();
}

// Inference, one wavefront per 16-row tile with no workgroup barriers.  The
// layers are computed transposed, H^T = W X^T, so a layer's accumulator
// (lane l: hidden units 16 mb + 4 (l / 16) + t of row l % 16) is exactly the
// next layer's B operand (K index 4 (l / 16) + t of K-step mb, column l % 16):
// activations never leave registers.  W2 lives in registers (its 16 A
// fragments); W1's 32 fragments are read from LDS, laid out per (mb, step,
// lane) so every read is one conflict-free 8-byte access.  Layer 1's K order
// is permuted (step s of lane group g contracts inputs 32 g + 4 s .. + 3) so
// each lane's 32 inputs of a row are contiguous: four 16-byte loads, issued
// for the next tile before the current one is computed.
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2v)); // RNE
}

constexpr int kInferWaves = 4;

__global__ void __launch_bounds__(64 * kInferWaves)
    mlp_infer_kernel(int64_t n, const uint16_t *__restrict__ X, const float *__restrict__ w,
                     const uint16_t *__restrict__ wq, float *__restrict__ pred)
{
    // W1 fragments: f1[(mb * 8 + s) * 64 + lane][t] = W1[16 mb + lane % 16][32 (lane / 16) + 4 s + t]
    __shared__ bf16x4 f1[4 * 8 * 64];
    __shared__ __align__(16) float b1s[NH], b2s[NH], w3s[NH];
    for (int q = threadIdx.x; q < 4 * 8 * 64; q += 64 * kInferWaves) {
        const int ln = q & 63, ms = q >> 6, mb = ms >> 3, st = ms & 7;
        const uint16_t *src = wq + (16 * mb + (ln & 15)) * NI + 32 * (ln >> 4) + 4 * st;
        bf16x4 v;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            v[t] = (short)src[t];
        f1[q] = v;
    }
    for (int q = threadIdx.x; q < NH; q += 64 * kInferWaves) {
        b1s[q] = w[OB1 + q];
        b2s[q] = w[OB2 + q];
        w3s[q] = bf2f(wq[NH * NI + NH * NH + q]);
    }
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    // W2 fragments: A[m = 16 mb + r][k = 16 s + 4 g + t]
    bf16x4 f2[4][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int st = 0; st < 4; ++st)
            f2[mb][st] = *reinterpret_cast<const bf16x4 *>(wq + NH * NI + (16 * mb + r) * NH +
                                                          16 * st + 4 * g);
    const float b3 = w[OB3];
    __syncthreads();
    const int64_t ntiles = (n + 15) / 16;
    const int64_t nwaves = (int64_t)gridDim.x * kInferWaves;
    int64_t tile = (int64_t)blockIdx.x * kInferWaves + (threadIdx.x >> 6);
    uint4 xv[4]; // this lane's 32 inputs of row 16 tile + r
    auto fetch = [&](int64_t tl) {
        const int64_t row = 16 * tl + r;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            xv[q] = make_uint4(0, 0, 0, 0);
        if (row < n) {
            const uint4 *src = reinterpret_cast<const uint4 *>(X + row * NI + 32 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                xv[q] = src[q];
        }
    };
    if (tile < ntiles)
        fetch(tile);
    for (; tile < ntiles; tile += nwaves) {
        uint4 cur[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cur[q] = xv[q];
        if (tile + nwaves < ntiles)
            fetch(tile + nwaves);
        // layer 1: z1^T = W1 X^T, 4 blocks of 16 hidden units x 16 rows
        f32x4 acc[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
            acc[mb] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const uint4 &c = cur[st >> 1];
            bf16x4 bx;
            const uint32_t lo = (st & 1) ? c.z : c.x, hi = (st & 1) ? c.w : c.y;
            bx = __builtin_bit_cast(bf16x4, ((uint64_t)hi << 32) | lo);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
                acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(f1[(mb * 8 + st) * 64 + lane],
                                                                    bx, acc[mb], 0, 0, 0);
        }
        // bias + GELU, rounded to BF16: layer 2's B fragments
        bf16x4 a1[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const float4 bb = *reinterpret_cast<const float4 *>(&b1s[16 * mb + 4 * g]);
            const uint32_t lo = pack_bf16(gelu(acc[mb][0] + bb.x), gelu(acc[mb][1] + bb.y));
            const uint32_t hi = pack_bf16(gelu(acc[mb][2] + bb.z), gelu(acc[mb][3] + bb.w));
            a1[mb] = __builtin_bit_cast(bf16x4, ((uint64_t)hi << 32) | lo);
        }
        // layer 2
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            f32x4 a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int st = 0; st < 4; ++st)
                a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(f2[mb][st], a1[st], a, 0, 0, 0);
            acc[mb] = a;
        }
        // output layer on the BF16-rounded activations, then the 4 lane groups
        float p = 0.f;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const float4 bb = *reinterpret_cast<const float4 *>(&b2s[16 * mb + 4 * g]);
            const float4 ww = *reinterpret_cast<const float4 *>(&w3s[16 * mb + 4 * g]);
            const uint32_t lo = pack_bf16(gelu(acc[mb][0] + bb.x), gelu(acc[mb][1] + bb.y));
            const uint32_t hi = pack_bf16(gelu(acc[mb][2] + bb.z), gelu(acc[mb][3] + bb.w));
            p += ww.x * __uint_as_float(lo << 16) + ww.y * __uint_as_float(lo & 0xffff0000u) +
                 ww.z * __uint_as_float(hi << 16) + ww.w * __uint_as_float(hi & 0xffff0000u);
        }
        p += __shfl_xor(p, 16);
        p += __shfl_xor(p, 32);
        const int64_t row = 16 * tile + r;
        if (g == 0 && row < n)
            pred[row] = p + b3;
    }
}

// Forward + backward over this workgroup's tiles; writes its partial
// gradient (NP floats) and partial loss sum.  inv_b = 1 / n (mean loss).
template <int R>
__global__ void __launch_bounds__(256)
    mlp_grad_kernel(int64_t n, const uint16_t *__restrict__ X, const float *__restrict__ y,
                    const float *__restrict__ w, const uint16_t *__restrict__ wq, float inv_b,
                    float *__restrict__ partial, float *__restrict__ lpart)
{
    __shared__ Smem<R, true> s;
    stage_weights(s, w, wq);
    const float b3 = w[OB3];
    const int w_ = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = threadIdx.x & 63, rg = threadIdx.x >> 6; // bias / W3 ownership
    f32x4 gw1[8], gw2[4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
        gw1[i] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; ++i)
        gw2[i] = {0.f, 0.f, 0.f, 0.f};
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
    const int64_t ntiles = (n + R - 1) / R;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t r0 = tile * R;
        __syncthreads();
        load_tile(s, X, n, r0);
        __syncthreads();
        forward_tile(s, b3);
        // output error
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            float d = 0.f;
            if (r0 + r < n) {
                const float e = s.pred[r] - y[r0 + r];
                loss += e * e;
                d = 2.0f * e * inv_b;
            }
            s.d3[r] = d;
        }
        __syncthreads();
        // delta2 = d3 w3 g2 (FP32 for b2's gradient, BF16 for the GEMMs)
        for (int r = rg; r < R; r += 4) {
            const float d3 = s.d3[r];
            const float d2 = d3 * s.w3[col] * s.g2[r][col];
            gb2 += d2;
            gw3 += d3 * bf2f(s.a2[r][col]);
            s.d2[r][col] = f2bf(d2);
            if (col == 0)
                gb3 += d3;
        }
        __syncthreads();
        // delta1 = (delta2 W2) g1: C[r][i] = sum_o d2[r][o] W2[o][i]
        constexpr int NB = (R / 16) * (NH / 16);
        for (int blk = w_; blk < NB; blk += 4) {
            const int m0 = (blk / (NH / 16)) * 16, n0 = (blk % (NH / 16)) * 16;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_block<NH, true, false>(&s.d2[0][0], LH, 1, &s.w2[0][0], LH, 1, m0, n0, acc);
            const int c = n0 + (lane & 15);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + (lane >> 4) * 4 + t;
                const float d1 = acc[t] * s.g1[row][c];
                s.g1[row][c] = d1; // g1 no longer needed: keep FP32 delta1
                s.d1[row][c] = f2bf(d1);
            }
        }
        __syncthreads();
        for (int r = rg; r < R; r += 4)
            gb1 += s.g1[r][col];
        // dW2[o][i] += sum_r d2[r][o] a1[r][i]: 16 blocks, 4 per wavefront
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int blk = w_ * 4 + q;
            const int m0 = (blk / 4) * 16, n0 = (blk % 4) * 16;
            gw2[q] = mfma_block<R, false, false>(&s.d2[0][0], 1, LH, &s.a1[0][0], LH, 1, m0, n0,
                                                 gw2[q]);
        }
        // dW1[o][i] += sum_r d1[r][o] x[r][i]: 32 blocks, 8 per wavefront
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int blk = w_ * 8 + q;
            const int m0 = (blk / 8) * 16, n0 = (blk % 8) * 16;
            gw1[q] = mfma_block<R, false, false>(&s.d1[0][0], 1, LH, &s.x[0][0], LX, 1, m0, n0,
                                                 gw1[q]);
        }
    }
    // write this workgroup's partial gradient
    float *P = partial + (size_t)blockIdx.x * NP;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int blk = w_ * 8 + q;
        const int m0 = (blk / 8) * 16, n0 = (blk % 8) * 16;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            P[OW1 + (m0 + (lane >> 4) * 4 + t) * NI + n0 + (lane & 15)] = gw1[q][t];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int blk = w_ * 4 + q;
        const int m0 = (blk / 4) * 16, n0 = (blk % 4) * 16;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            P[OW2 + (m0 + (lane >> 4) * 4 + t) * NH + n0 + (lane & 15)] = gw2[q][t];
    }
    __syncthreads();
    // per-column sums over the four row groups, in a fixed order
    float *red = s.red;
    red[rg * NH + col] = gb1;
    __syncthreads();
    if (threadIdx.x < NH)
        P[OB1 + col] = red[col] + red[NH + col] + red[2 * NH + col] + red[3 * NH + col];
    __syncthreads();
    red[rg * NH + col] = gb2;
    __syncthreads();
    if (threadIdx.x < NH)
        P[OB2 + col] = red[col] + red[NH + col] + red[2 * NH + col] + red[3 * NH + col];
    __syncthreads();
    red[rg * NH + col] = gw3;
    __syncthreads();
    if (threadIdx.x < NH)
        P[OW3 + col] = red[col] + red[NH + col] + red[2 * NH + col] + red[3 * NH + col];
    __syncthreads();
    // b3 and loss: block sums
    red[threadIdx.x] = gb3;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.f;
        for (int i = 0; i < 256; ++i)
            a += red[i];
        P[OB3] = a;
    }
    __syncthreads();
    red[threadIdx.x] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.f;
        for (int i = 0; i < 256; ++i)
            a += red[i];
        lpart[blockIdx.x] = a;
    }
}

// Sum of the partial gradients (fixed order), then either write the
// gradient or take one Adam step on the master weights and refresh their
// BF16 copies.
__global__ void __launch_bounds__(256)
    mlp_reduce_kernel(int nparts, const float *__restrict__ partial,
                      const float *__restrict__ lpart, float inv_b, float *__restrict__ grad,
                      float *__restrict__ loss, AdamStep adam, float *__restrict__ w,
                      float *__restrict__ m, float *__restrict__ v, uint16_t *__restrict__ wq)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    // Sums in partial order (deterministic); unrolled so that the loads of
    // a group are in flight together.
    if (p == 0 && loss) {
        float a = 0.f;
#pragma unroll 16
        for (int i = 0; i < nparts; ++i)
            a += lpart[i];
        *loss = a * inv_b;
    }
    if (p >= NP)
        return;
    float g = 0.f;
#pragma unroll 16
    for (int i = 0; i < nparts; ++i)
        g += partial[(size_t)i * NP + p];
    if (grad)
        grad[p] = g;
    if (adam.lr != 0.f) {
        const float mi = adam.beta1 * m[p] + (1.0f - adam.beta1) * g;
        const float vi = adam.beta2 * v[p] + (1.0f - adam.beta2) * g * g;
        m[p] = mi;
        v[p] = vi;
        const float wi = w[p] - adam.lr * (mi * adam.c1) / (sqrtf(vi * adam.c2) + adam.eps);
        w[p] = wi;
        if (p < OB1)
            wq[p] = f2bf(wi);
        else if (p >= OW2 && p < OB2)
            wq[NH * NI + (p - OW2)] = f2bf(wi);
        else if (p >= OW3 && p < OB3)
            wq[NH * NI + NH * NH + (p - OW3)] = f2bf(wi);
    }
}

__global__ void __launch_bounds__(256)
    mlp_classify_kernel(int64_t n, const float *__restrict__ pred, float refine_level,
                        float coarsen_level, uint8_t *__restrict__ action)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float p = pred[i];
        action[i] = (p > refine_level) ? 1 : (p < coarsen_level ? 2 : 0);
    }
}

constexpr int kRowsTrain = 32;

} // namespace

int mlp_grad_parts(int64_t n)
{
    const int64_t tiles = (n + kRowsTrain - 1) / kRowsTrain;
    return (int)(tiles < kMlpMaxParts ? tiles : kMlpMaxParts);
}

hipError_t launch_mlp_forward(int64_t n, const uint16_t *X, const float *w, const uint16_t *wq,
                              float *pred, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    const int64_t tiles = (n + 15) / 16;
    // persistent wavefronts: as many workgroups as are resident at once
    static const int cap = [] {
        int dev = 0, per = 1;
        hipDeviceProp_t p;
        int cus = 256;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            cus = p.multiProcessorCount;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, mlp_infer_kernel, 64 * kInferWaves,
                                                         0) != hipSuccess ||
            per < 1)
            per = 1;
        return cus * per;
    }();
    const int64_t need = (tiles + kInferWaves - 1) / kInferWaves;
    const unsigned grid = (unsigned)(need < cap ? need : cap);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(64 * kInferWaves), 0, s, n, X, w, wq,
                       pred);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(int64_t n, const uint16_t *X, const float *y, float *w, uint16_t *wq,
                           float *m, float *v, float *partial, float *lpart, float *grad,
                           float *loss, const AdamStep &adam, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    const int parts = mlp_grad_parts(n);
    const float inv_b = 1.0f / (float)n;
    hipLaunchKernelGGL(mlp_grad_kernel<kRowsTrain>, dim3(parts), dim3(256), 0, s, n, X, y, w, wq,
                       inv_b, partial, lpart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, parts,
                       partial, lpart, inv_b, grad, loss, adam, w, m, v, wq);
    return hipGetLastError();
}

hipError_t launch_mlp_classify(int64_t n, const float *pred, float refine_level,
                               float coarsen_level, uint8_t *action, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    const int64_t b = (n + 255) / 256;
    hipLaunchKernelGGL(mlp_classify_kernel, dim3((unsigned)(b < 4096 ? b : 4096)), dim3(256), 0, s,
                       n, pred, refine_level, coarsen_level, action);
    return hipGetLastError();
}

} // namespace drf
