// Learned edge prior on MI355X: candidate-edge features, the 8 -> 64 -> 64
// -> 1 MLP (tanh GELU) on bf16 matrix cores with FP32 accumulation, and its
// training (binary cross-entropy on best-tour membership) with hand-written
// forward and backward passes, a deterministic gradient reduction and Adam.
// The reference has no learned prior; the CPU restatement is the oracle's.
//
// Every product runs on v_mfma_f32_16x16x32_bf16 (one wave: a 16x16 tile
// with K = 32; lane l holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and
// D[4(l/16)+r][l%16]). Operands are staged in LDS as bf16 with K contiguous
// for both operands (A row-major, B column-major), so each lane reads each
// operand with one 16-byte LDS load; a wave works on tiles of 32 edges. The
// weights reach a workgroup as one image in that layout (kept current by the
// Adam step), copied with 16-byte loads. bf16 rounding points: edge
// features, hidden activations, output-layer inputs, the two backward
// deltas, and all weights; everything else is FP32.
#include "faco_common.h"

namespace faco {
namespace {

constexpr int kParams = 64 * 8 + 64 + 64 * 64 + 64 + 64 + 1;  // 4801
constexpr int oW1 = 0, oB1 = 512, oW2 = 576, oB2 = 4672, oW3 = 4736, oB3 = 4800;
constexpr int kTile = 32;          // edges per wave tile
constexpr int kWavesPerBlock = 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// C (16x16 fragment) += A[16][32] (row-major, row stride lda) x B[32][16]
// given column-major as Bc[16][32] (column stride ldb): Bc[c][k] = B[k][c].
// Strides and bases are multiples of 8 elements (16-byte aligned loads).
__device__ __forceinline__ f32x4 mma(const __bf16* A, int lda, const __bf16* Bc, int ldb, f32x4 c, int lane) {
    const int i = lane & 15, kb = (lane >> 4) * 8;
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + i * lda + kb);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bc + i * ldb + kb);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float gelu(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return 0.5f * z * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
}
__device__ __forceinline__ float sigmoid(float s) { return 1.0f / (1.0f + expf(-s)); }
__device__ __forceinline__ float softplus(float s) { return s > 0.0f ? s + log1pf(expf(-s)) : log1pf(expf(s)); }

}  // namespace

// The weights in the layouts the products read (B operands column-major):
// the block-shared LDS copy and, byte for byte, the image in device memory.
struct alignas(16) WeightsLds {
    __bf16 w1n[64 * 32];   // [j][k] = W1[j][k], k >= 8 zero
    __bf16 w2n[64 * 64];   // [j][k] = W2[j][k]
    __bf16 w2t[64 * 64];   // [k][j] = W2[j][k]
    __bf16 w3n[16 * 64];   // [0][k] = w3[k], rows 1..15 zero
    float b1[64], b2[64], w3[64], b3;   // w3 rounded through bf16
};

namespace {

// Per-wave tile buffers (B operands column-major).
struct TileLds {
    __bf16 xb[kTile * 32];     // features [e][k], K padded to 32
    __bf16 xt[32 * kTile];     // the same transposed, [k][e]
    float z1[kTile * 64];
    __bf16 h1b[kTile * 64];    // [e][j]
    __bf16 h1t[64 * kTile];    // [j][e]
    float z2[kTile * 64];      // reused for the FP32 first-layer delta in backward
    __bf16 h2b[kTile * 64];
    __bf16 db[kTile * 64];     // second-layer delta, [e][j]
    __bf16 dbt[64 * kTile];    // a delta transposed, [j][e]
    float s[kTile], ds[kTile];
};

__device__ void load_weights(const WeightsLds* image, WeightsLds& w) {
    const uint4* src = reinterpret_cast<const uint4*>(image);
    uint4* dst = reinterpret_cast<uint4*>(&w);
    for (uint32_t i = threadIdx.x; i < sizeof(WeightsLds) / 16; i += blockDim.x) dst[i] = src[i];
}

// Parameter p's entries of the weight image (the zero padding is written
// once, when the image is cleared).
__device__ __forceinline__ void image_entry(WeightsLds* w, uint32_t p, float value) {
    const __bf16 h = static_cast<__bf16>(value);
    if (p < oB1) {
        const uint32_t j = (p - oW1) / 8, k = (p - oW1) % 8;
        w->w1n[j * 32 + k] = h;
    } else if (p < oW2) {
        w->b1[p - oB1] = value;
    } else if (p < oB2) {
        const uint32_t j = (p - oW2) / 64, k = (p - oW2) % 64;
        w->w2n[j * 64 + k] = h;
        w->w2t[k * 64 + j] = h;
    } else if (p < oW3) {
        w->b2[p - oB2] = value;
    } else if (p < oB3) {
        w->w3n[p - oW3] = h;
        w->w3[p - oW3] = static_cast<float>(h);
    } else {
        w->b3 = value;
    }
}

// Forward pass of one tile of `count` edges starting at `first`; with
// Backward, the transposed features and first hidden layer too.
template<bool Backward>
__device__ void forward_tile(const float* features, const uint32_t* indices, uint32_t first, uint32_t count,
                             const WeightsLds& w, TileLds& t, int lane) {
    for (int i = lane; i < kTile * 32; i += 64) {
        const int e = i / 32, k = i % 32;
        float v = 0.0f;
        if (e < static_cast<int>(count) && k < 8) {
            const uint32_t id = indices ? indices[first + e] : first + e;
            v = features[static_cast<size_t>(id) * 8 + k];
        }
        t.xb[i] = static_cast<__bf16>(v);
        if (Backward) t.xt[k * kTile + e] = static_cast<__bf16>(v);
    }
    wsync();
    const int fi = 4 * (lane >> 4), fj = lane & 15;
    for (int rt = 0; rt < 2; ++rt)
        for (int ct = 0; ct < 4; ++ct) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma(t.xb + rt * 16 * 32, 32, w.w1n + ct * 16 * 32, 32, acc, lane);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = rt * 16 + fi + r, j = ct * 16 + fj;
                const float z = acc[r] + w.b1[j];
                const __bf16 h = static_cast<__bf16>(gelu(z));
                t.z1[e * 64 + j] = z;
                t.h1b[e * 64 + j] = h;
                if (Backward) t.h1t[j * kTile + e] = h;
            }
        }
    wsync();
    for (int rt = 0; rt < 2; ++rt)
        for (int ct = 0; ct < 4; ++ct) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < 2; ++ks)
                acc = mma(t.h1b + rt * 16 * 64 + ks * 32, 64, w.w2n + ct * 16 * 64 + ks * 32, 64, acc, lane);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = rt * 16 + fi + r, j = ct * 16 + fj;
                const float z = acc[r] + w.b2[j];
                t.z2[e * 64 + j] = z;
                t.h2b[e * 64 + j] = static_cast<__bf16>(gelu(z));
            }
        }
    wsync();
    for (int rt = 0; rt < 2; ++rt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < 2; ++ks)
            acc = mma(t.h2b + rt * 16 * 64 + ks * 32, 64, w.w3n + ks * 32, 64, acc, lane);
        if (fj == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) t.s[rt * 16 + fi + r] = acc[r] + w.b3;
    }
    wsync();
}

}  // namespace

// Edge features (one thread per candidate edge); the definition is the
// oracle's oracle_prior_features.
__global__ void faco_prior_features_kernel(uint32_t n, uint32_t pw, const uint32_t* coords, const uint32_t* primary,
                                           const double* local_scale, double epsilon, float mean_scale, float* out) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * pw) return;
    const uint32_t a = idx / pw, j = idx % pw;
    const uint32_t* row = primary + static_cast<size_t>(a) * pw;
    const uint32_t b = row[j];
    const uint32_t* brow = primary + static_cast<size_t>(b) * pw;
    const float sa = static_cast<float>(local_scale[a] > epsilon ? local_scale[a] : epsilon);
    const float sb = static_cast<float>(local_scale[b] > epsilon ? local_scale[b] : epsilon);
    const uint32_t d0i = euc2d(coords[a], coords[row[0]]);
    const float d0 = static_cast<float>(d0i > 1u ? d0i : 1u);
    const float d = static_cast<float>(euc2d(coords[a], coords[b]));
    uint32_t rank = pw, shared = 0;
    for (uint32_t k = 0; k < pw; ++k) {
        if (brow[k] == a && rank == pw) rank = k;
        for (uint32_t m = 0; m < pw; ++m) shared += brow[k] == row[m] ? 1u : 0u;
    }
    auto clip = [](float v) { return v < 0.0f ? 0.0f : (v > 4.0f ? 4.0f : v); };
    float* f = out + static_cast<size_t>(idx) * 8;
    f[0] = clip(d / sa);
    f[1] = static_cast<float>(j) / static_cast<float>(pw - 1);
    f[2] = clip(d / d0);
    f[3] = rank < pw ? 1.0f : 0.0f;
    f[4] = static_cast<float>(rank) / static_cast<float>(pw);
    f[5] = clip(d / sb);
    f[6] = clip(d / mean_scale);
    f[7] = static_cast<float>(shared) / static_cast<float>(pw);
}

// Logits of `edges` edges.
__global__ void __launch_bounds__(64 * kWavesPerBlock) faco_prior_infer_kernel(const WeightsLds* image, const float* features,
                                                                               uint32_t edges, float* logits) {
    __shared__ WeightsLds w;
    __shared__ TileLds tiles[kWavesPerBlock];
    load_weights(image, w);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    TileLds& t = tiles[wave];
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    for (uint32_t tile = blockIdx.x * kWavesPerBlock + wave; tile * kTile < edges; tile += waves) {
        const uint32_t first = tile * kTile, count = edges - first < kTile ? edges - first : kTile;
        forward_tile<false>(features, nullptr, first, count, w, t, lane);
        if (lane < static_cast<int>(count)) logits[first + lane] = t.s[lane];
        wsync();
    }
}

// Gradient partials of the weighted BCE loss over `edges` edges: one row
// of kParams + 1 (the loss) per wave.
__global__ void __launch_bounds__(64 * kWavesPerBlock) faco_prior_grad_kernel(const WeightsLds* image, const float* features,
                                                                              const uint8_t* labels,
                                                                              const uint32_t* indices, uint32_t offset,
                                                                              uint32_t edges, float pos_weight,
                                                                              float inv_total, float* partial) {
    __shared__ WeightsLds w;
    __shared__ TileLds tiles[kWavesPerBlock];
    load_weights(image, w);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    TileLds& t = tiles[wave];
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    const int fi = 4 * (lane >> 4), fj = lane & 15;
    f32x4 g2[4][4], g1[4];
    for (int a = 0; a < 4; ++a) {
        g1[a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 4; ++b) g2[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
    for (uint32_t tile = blockIdx.x * kWavesPerBlock + wave; tile * kTile < edges; tile += waves) {
        const uint32_t first = tile * kTile, count = edges - first < kTile ? edges - first : kTile;
        forward_tile<true>(features, indices, offset + first, count, w, t, lane);
        if (lane < kTile) {
            float ds = 0.0f;
            if (lane < static_cast<int>(count)) {
                const uint32_t id = indices ? indices[offset + first + lane] : offset + first + lane;
                const float y = labels[id] ? 1.0f : 0.0f;
                const float we = labels[id] ? pos_weight : 1.0f;
                const float s = t.s[lane];
                ds = we * (sigmoid(s) - y) * inv_total;
                t.ds[lane] = ds;
            } else {
                t.ds[lane] = 0.0f;
            }
        }
        wsync();
        if (lane == 0) {
            for (uint32_t e = 0; e < count; ++e) {
                const uint32_t id = indices ? indices[offset + first + e] : offset + first + e;
                const float y = labels[id] ? 1.0f : 0.0f;
                const float we = labels[id] ? pos_weight : 1.0f;
                loss += we * (softplus(t.s[e]) - y * t.s[e]) * inv_total;
                gb3 += t.ds[e];
            }
        }
        // output layer and second-layer delta: lane j owns column j
        {
            const int j = lane;
            float acc_w3 = 0.0f, acc_b2 = 0.0f;
            for (int e = 0; e < kTile; ++e) {
                const float ds = t.ds[e];
                acc_w3 += ds * static_cast<float>(t.h2b[e * 64 + j]);
                const float dz = ds * w.w3[j] * gelu_grad(t.z2[e * 64 + j]);
                acc_b2 += dz;
                t.db[e * 64 + j] = static_cast<__bf16>(dz);
                t.dbt[j * kTile + e] = static_cast<__bf16>(dz);
            }
            gw3 += acc_w3;
            gb2 += acc_b2;
        }
        wsync();
        // gW2 += dz2^T h1 (K = the 32 edges of the tile)
        for (int jt = 0; jt < 4; ++jt)
            for (int kt = 0; kt < 4; ++kt) g2[jt][kt] = mma(t.dbt + jt * 16 * kTile, kTile, t.h1t + kt * 16 * kTile, kTile, g2[jt][kt], lane);
        // first-layer delta: dz1 = (dz2 W2) * gelu'(z1), FP32 into the z2 buffer
        for (int rt = 0; rt < 2; ++rt)
            for (int ct = 0; ct < 4; ++ct) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int ks = 0; ks < 2; ++ks)
                    acc = mma(t.db + rt * 16 * 64 + ks * 32, 64, w.w2t + ct * 16 * 64 + ks * 32, 64, acc, lane);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = rt * 16 + fi + r, k = ct * 16 + fj;
                    t.z2[e * 64 + k] = acc[r] * gelu_grad(t.z1[e * 64 + k]);
                }
            }
        wsync();
        {
            const int k = lane;
            float acc_b1 = 0.0f;
            for (int e = 0; e < kTile; ++e) {
                const float dz = t.z2[e * 64 + k];
                acc_b1 += dz;
                t.dbt[k * kTile + e] = static_cast<__bf16>(dz);
            }
            gb1 += acc_b1;
        }
        wsync();
        // gW1 += dz1^T x
        for (int jt = 0; jt < 4; ++jt) g1[jt] = mma(t.dbt + jt * 16 * kTile, kTile, t.xt, kTile, g1[jt], lane);
        wsync();
    }
    float* out = partial + static_cast<size_t>(blockIdx.x * kWavesPerBlock + wave) * (kParams + 1);
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = jt * 16 + fi + r, k = fj;
            if (k < 8) out[oW1 + j * 8 + k] = g1[jt][r];
        }
    for (int jt = 0; jt < 4; ++jt)
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[oW2 + (jt * 16 + fi + r) * 64 + kt * 16 + fj] = g2[jt][kt][r];
    out[oB1 + lane] = gb1;
    out[oB2 + lane] = gb2;
    out[oW3 + lane] = gw3;
    if (lane == 0) { out[oB3] = gb3; out[kParams] = loss; }
}

// Sum of the per-wave partials in wave order; with `adam`, one Adam step on
// the FP32 master weights (bias-corrected, step number `step` >= 1).
__global__ void faco_prior_reduce_kernel(const float* partial, uint32_t rows, float* grad, float* params, float* m,
                                         float* v, uint32_t adam, float lr, float beta1, float beta2, float eps,
                                         uint32_t step, float* loss_sum, WeightsLds* image) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > static_cast<uint32_t>(kParams)) return;
    float g = 0.0f;
    for (uint32_t r = 0; r < rows; ++r) g += partial[static_cast<size_t>(r) * (kParams + 1) + p];
    grad[p] = g;
    if (p == static_cast<uint32_t>(kParams)) {
        if (loss_sum) loss_sum[0] += g;
        return;
    }
    if (!adam) return;
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(beta1, static_cast<float>(step)));
    const float vh = vv / (1.0f - powf(beta2, static_cast<float>(step)));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
    image_entry(image, p, params[p]);
}

// Every parameter's entries of the weight image (after the image is cleared).
__global__ void faco_prior_image_kernel(const float* params, WeightsLds* image) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < static_cast<uint32_t>(kParams)) image_entry(image, p, params[p]);
}

// 16x16x32 bf16 product of one wave on global operands (layout check).
__global__ void faco_prior_mma_check_kernel(const float* A, const float* B, float* C) {
    __shared__ __attribute__((aligned(16))) __bf16 a[16 * 32];
    __shared__ __attribute__((aligned(16))) __bf16 bc[16 * 32];   // B column-major: bc[c][k] = B[k][c]
    const int lane = threadIdx.x;
    for (int i = lane; i < 16 * 32; i += 64) {
        a[i] = static_cast<__bf16>(A[i]);
        bc[(i % 16) * 32 + i / 16] = static_cast<__bf16>(B[i]);
    }
    wsync();
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mma(a, 32, bc, 32, acc, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) C[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[r];
}

hipError_t launch_prior_features(uint32_t n, uint32_t pw, const uint32_t* coords, const uint32_t* primary,
                                 const double* local_scale, double epsilon, float mean_scale, float* out,
                                 hipStream_t stream) {
    const uint32_t threads = n * pw;
    hipLaunchKernelGGL(faco_prior_features_kernel, dim3((threads + 255) / 256), dim3(256), 0, stream, n, pw, coords,
                       primary, local_scale, epsilon, mean_scale, out);
    return hipGetLastError();
}

uint32_t prior_param_count() { return kParams; }

uint32_t prior_grad_rows(uint32_t edges) {
    const uint32_t tiles = (edges + kTile - 1) / kTile;
    uint32_t blocks = (tiles + kWavesPerBlock - 1) / kWavesPerBlock;
    blocks = blocks < 512u ? blocks : 512u;
    return (blocks ? blocks : 1u) * kWavesPerBlock;
}

uint32_t prior_image_bytes() { return static_cast<uint32_t>(sizeof(WeightsLds)); }

hipError_t launch_prior_image(const float* params, void* image, hipStream_t stream) {
    hipError_t status = hipMemsetAsync(image, 0, sizeof(WeightsLds), stream);
    if (status != hipSuccess) return status;
    hipLaunchKernelGGL(faco_prior_image_kernel, dim3((kParams + 255) / 256), dim3(256), 0, stream, params,
                       static_cast<WeightsLds*>(image));
    return hipGetLastError();
}

hipError_t launch_prior_infer(const void* image, const float* features, uint32_t edges, float* logits,
                              hipStream_t stream) {
    const uint32_t blocks = prior_grad_rows(edges) / kWavesPerBlock;
    hipLaunchKernelGGL(faco_prior_infer_kernel, dim3(blocks), dim3(64 * kWavesPerBlock), 0, stream,
                       static_cast<const WeightsLds*>(image), features, edges, logits);
    return hipGetLastError();
}

hipError_t launch_prior_grad(const void* image, const float* features, const uint8_t* labels,
                             const uint32_t* indices, uint32_t offset, uint32_t edges, float pos_weight,
                             float inv_total, float* partial, hipStream_t stream) {
    const uint32_t blocks = prior_grad_rows(edges) / kWavesPerBlock;
    hipLaunchKernelGGL(faco_prior_grad_kernel, dim3(blocks), dim3(64 * kWavesPerBlock), 0, stream,
                       static_cast<const WeightsLds*>(image), features, labels, indices, offset, edges, pos_weight,
                       inv_total, partial);
    return hipGetLastError();
}

hipError_t launch_prior_reduce(const float* partial, uint32_t rows, float* grad, float* params, float* m, float* v,
                               bool adam, float lr, float beta1, float beta2, float eps, uint32_t step,
                               float* loss_sum, void* image, hipStream_t stream) {
    hipLaunchKernelGGL(faco_prior_reduce_kernel, dim3((kParams + 1 + 255) / 256), dim3(256), 0, stream, partial, rows,
                       grad, params, m, v, adam ? 1u : 0u, lr, beta1, beta2, eps, step, loss_sum,
                       static_cast<WeightsLds*>(image));
    return hipGetLastError();
}

hipError_t launch_prior_mma_check(const float* A, const float* B, float* C, hipStream_t stream) {
    hipLaunchKernelGGL(faco_prior_mma_check_kernel, dim3(1), dim3(64), 0, stream, A, B, C);
    return hipGetLastError();
}

}  // namespace faco
