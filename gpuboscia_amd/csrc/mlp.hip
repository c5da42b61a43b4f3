// Learned branching scorer: an 8 -> 64 -> 64 -> 64 -> 1 ReLU MLP trained and
// evaluated by the library's own gfx950 kernels.  No reference counterpart
// (the reference branches on the most fractional variables,
// branching.h:24-118); the scorer only reorders branching candidates, so the
// proven optimum is unaffected.
//
// Arithmetic: FP32 master weights; the forward pass feeds BF16-rounded
// weights and activations to v_mfma_f32_16x16x16_bf16 with FP32
// accumulation; the backward pass is FP32 on the same BF16-rounded values
// (so its gradients are those of the network the forward pass evaluates);
// Adam updates the FP32 masters.  Weight gradients are per-block partial sums
// reduced in block order, so training is deterministic.
//
// Parameter layout (floats): W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] w4[64] b4[1]  (W[in][out], row-major).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "gb_device.h"
#include "gb_host.h"
#include "gb_mlp.h"

namespace gb {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kMlpHidden;   // 64
constexpr int F = kMlpFeatures; // 8

__host__ __device__ inline unsigned short f2bf(float f) {
    // round to nearest even (finite inputs)
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short b) {
    const unsigned u = (unsigned)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
__device__ inline float bfround(float f) { return bf2f(f2bf(f)); }

struct MlpLds {
    unsigned short w1[16 * H];   // K padded 8 -> 16
    unsigned short w2[H * H];
    unsigned short w3[H * H];
    float b1[H], b2[H], b3[H], w4[H];
    float b4;
    unsigned short act[4][16 * H];  // per-wave 16-sample activation tile (bf16)
};

__device__ void load_weights(const float* p, MlpLds& s) {
    for (int i = threadIdx.x; i < 16 * H; i += blockDim.x) {
        const int k = i / H, c = i % H;
        s.w1[i] = k < F ? f2bf(p[kOffW1 + k * H + c]) : 0;
    }
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        s.w2[i] = f2bf(p[kOffW2 + i]);
        s.w3[i] = f2bf(p[kOffW3 + i]);
    }
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        s.b1[i] = p[kOffB1 + i];
        s.b2[i] = p[kOffB2 + i];
        s.b3[i] = p[kOffB3 + i];
        s.w4[i] = bfround(p[kOffW4 + i]);
    }
    if (threadIdx.x == 0) s.b4 = p[kOffB4];
    __syncthreads();
}

// One wave: 16 samples through a 64-wide layer.  `a` holds the wave's input
// tile (16 x K bf16, row = sample); out[t][r] = relu(sum + bias) for output
// tile t, row 4*(lane/16)+r, column 16 t + lane%16 (MFMA D layout).
__device__ void layer(const unsigned short* a, int K, const unsigned short* w, const float* bias,
                      float (&out)[4][4]) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < K; kk += 16) {
            bf16x4 av, bv;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = kk + 4 * (l / 16) + i;
                av[i] = (short)a[(l % 16) * H + k];
                bv[i] = (short)w[k * H + 16 * t + (l % 16)];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + bias[16 * t + (l % 16)];
            out[t][r] = z > 0.f ? z : 0.f;
        }
    }
}

// Writes a D-layout activation tile to the wave's LDS tile as bf16 (row =
// sample) and, optionally, to global memory as float (bf16-rounded values).
__device__ void store_tile(float (&h)[4][4], unsigned short* tile, float* gout, int s0, int M) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * (l / 16) + r, col = 16 * t + (l % 16);
            const unsigned short b = f2bf(h[t][r]);
            tile[row * H + col] = b;
            if (gout && s0 + row < M) gout[(size_t)(s0 + row) * H + col] = bf2f(b);
        }
    wave_sync();
}

__global__ void __launch_bounds__(256) mlp_forward_kernel(const float* params, const float* mu,
                                                          const float* sigma, const float* X, int M,
                                                          float* Y, float* Xn, float* H1, float* H2,
                                                          float* H3) {
    __shared__ MlpLds s;
    load_weights(params, s);
    const int l = lane_id();
    const int w = wave_id();
    unsigned short* tile = s.act[w];
    const int tiles = (M + 15) / 16;
    for (int tix = blockIdx.x * wave_count() + w; tix < tiles; tix += gridDim.x * wave_count()) {
        const int s0 = tix * 16;
        // standardized input tile (K padded to 16 with zeros)
        for (int e = l; e < 16 * 16; e += kWave) {
            const int row = e / 16, k = e % 16;
            float v = 0.f;
            if (k < F && s0 + row < M) v = (X[(size_t)(s0 + row) * F + k] - mu[k]) / sigma[k];
            const unsigned short b = f2bf(v);
            tile[row * H + k] = b;
            if (Xn && k < F && s0 + row < M) Xn[(size_t)(s0 + row) * F + k] = bf2f(b);
        }
        wave_sync();
        float h[4][4];
        layer(tile, 16, s.w1, s.b1, h);
        wave_sync();
        store_tile(h, tile, H1, s0, M);
        layer(tile, H, s.w2, s.b2, h);
        wave_sync();
        store_tile(h, tile, H2, s0, M);
        layer(tile, H, s.w3, s.b3, h);
        wave_sync();
        // output: y[row] = sum_c bf16(h3[row][c]) * w4[c] + b4; lanes sharing
        // l/16 hold the same 4 rows and 4 of the 64 columns each.
        float part[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float hv = bfround(h[t][r]);
                acc = fmaf(hv, s.w4[16 * t + (l % 16)], acc);
            }
            part[r] = acc;
        }
        store_tile(h, tile, H3, s0, M);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = part[r];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, kWave);
            const int row = 4 * (l / 16) + r;
            if ((l % 16) == 0 && s0 + row < M) Y[s0 + row] = v + s.b4;
        }
    }
}

// Per-sample backward of the MSE loss mean((y - t)^2): dZ for the three
// hidden layers (ReLU mask from the stored activations) and dy.
__global__ void __launch_bounds__(256) mlp_backward_dz_kernel(const float* params, const float* Y,
                                                              const float* T, int M, float scale,
                                                              const float* H1, const float* H2,
                                                              const float* H3, float* DY, float* DZ1,
                                                              float* DZ2, float* DZ3) {
    __shared__ float w2[H * H], w3[H * H], w4[H];
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        w2[i] = bfround(params[kOffW2 + i]);
        w3[i] = bfround(params[kOffW3 + i]);
    }
    for (int i = threadIdx.x; i < H; i += blockDim.x) w4[i] = bfround(params[kOffW4 + i]);
    __syncthreads();
    for (int sidx = blockIdx.x * blockDim.x + threadIdx.x; sidx < M; sidx += gridDim.x * blockDim.x) {
        const float dy = scale * (Y[sidx] - T[sidx]);
        DY[sidx] = dy;
        float dz3[H], dz2[H];
        const float* h3 = H3 + (size_t)sidx * H;
        const float* h2 = H2 + (size_t)sidx * H;
        const float* h1 = H1 + (size_t)sidx * H;
        for (int k = 0; k < H; ++k) dz3[k] = h3[k] > 0.f ? dy * w4[k] : 0.f;
        for (int k = 0; k < H; ++k) DZ3[(size_t)sidx * H + k] = dz3[k];
        for (int j = 0; j < H; ++j) {
            float acc = 0.f;
            for (int k = 0; k < H; ++k) acc = fmaf(w3[j * H + k], dz3[k], acc);
            dz2[j] = h2[j] > 0.f ? acc : 0.f;
        }
        for (int j = 0; j < H; ++j) DZ2[(size_t)sidx * H + j] = dz2[j];
        for (int j = 0; j < H; ++j) {
            float acc = 0.f;
            for (int k = 0; k < H; ++k) acc = fmaf(w2[j * H + k], dz2[k], acc);
            DZ1[(size_t)sidx * H + j] = h1[j] > 0.f ? acc : 0.f;
        }
    }
}

// Partial weight gradients over a chunk of samples: for each layer,
// dW[j][k] = sum_s A[s][j] dZ[s][k], db[k] = sum_s dZ[s][k]; written to
// partial[blockIdx][param] (deterministic per block).
__global__ void __launch_bounds__(256) mlp_weight_grad_kernel(int M, int chunk, const float* Xn,
                                                              const float* H1, const float* H2,
                                                              const float* H3, const float* DY,
                                                              const float* DZ1, const float* DZ2,
                                                              const float* DZ3, float* partial) {
    const int s_begin = blockIdx.x * chunk;
    const int s_end = min(M, s_begin + chunk);
    float* out = partial + (size_t)blockIdx.x * kMlpParams;
    for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) {
        const float* A = nullptr;
        const float* D = nullptr;
        int ain = 0, j = 0, k = 0;
        bool bias = false;
        if (p < kOffB1) { A = Xn; ain = F; D = DZ1; j = p / H; k = p % H; }
        else if (p < kOffW2) { D = DZ1; k = p - kOffB1; bias = true; }
        else if (p < kOffB2) { A = H1; ain = H; D = DZ2; j = (p - kOffW2) / H; k = (p - kOffW2) % H; }
        else if (p < kOffW3) { D = DZ2; k = p - kOffB2; bias = true; }
        else if (p < kOffB3) { A = H2; ain = H; D = DZ3; j = (p - kOffW3) / H; k = (p - kOffW3) % H; }
        else if (p < kOffW4) { D = DZ3; k = p - kOffB3; bias = true; }
        else if (p < kOffB4) { A = H3; ain = H; D = DY; j = p - kOffW4; k = 0; }
        else { D = DY; k = 0; bias = true; }
        const int dstride = (D == DY) ? 1 : H;
        float acc = 0.f;
        for (int s = s_begin; s < s_end; ++s) {
            const float d = D[(size_t)s * dstride + k];
            acc = bias ? acc + d : fmaf(A[(size_t)s * ain + j], d, acc);
        }
        out[p] = acc;
    }
}

__global__ void mlp_reduce_partials_kernel(const float* partial, int blocks, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float acc = 0.f;
    for (int b = 0; b < blocks; ++b) acc += partial[(size_t)b * kMlpParams + p];
    grad[p] = acc;
}

__global__ void mlp_sq_err_kernel(const float* Y, const float* T, int M, double* partial) {
    __shared__ double red[kMaxWaves];
    double acc = 0.0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < M; s += gridDim.x * blockDim.x) {
        const double d = (double)Y[s] - (double)T[s];
        acc = fma(d, d, acc);
    }
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.f - b1) * gi;
    const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace
}  // namespace gb

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct gb_brancher {
    float* params = nullptr;   // device FP32 masters
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* mu = nullptr;       // feature standardization (device)
    float* sigma = nullptr;
    int adam_t = 0;
    // work buffers (grown on demand)
    int cap = 0;
    float *Y = nullptr, *Xn = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr;
    float *DY = nullptr, *DZ1 = nullptr, *DZ2 = nullptr, *DZ3 = nullptr, *partial = nullptr;
    double* sq = nullptr;
    int partial_blocks = 0;
};

namespace gb {
namespace {

constexpr int kGradChunk = 512;
constexpr int kSqBlocks = 256;

gb_status ensure_cap(gb_brancher* b, int M) {
    if (M <= b->cap) return GB_OK;
    for (float* p : {b->Y, b->Xn, b->H1, b->H2, b->H3, b->DY, b->DZ1, b->DZ2, b->DZ3, b->partial})
        if (p) (void)hipFree(p);
    const int cap = M;
    GB_HIP(hipMalloc(&b->Y, (size_t)cap * 4));
    GB_HIP(hipMalloc(&b->Xn, (size_t)cap * F * 4));
    GB_HIP(hipMalloc(&b->H1, (size_t)cap * H * 4));
    GB_HIP(hipMalloc(&b->H2, (size_t)cap * H * 4));
    GB_HIP(hipMalloc(&b->H3, (size_t)cap * H * 4));
    GB_HIP(hipMalloc(&b->DY, (size_t)cap * 4));
    GB_HIP(hipMalloc(&b->DZ1, (size_t)cap * H * 4));
    GB_HIP(hipMalloc(&b->DZ2, (size_t)cap * H * 4));
    GB_HIP(hipMalloc(&b->DZ3, (size_t)cap * H * 4));
    b->partial_blocks = (cap + kGradChunk - 1) / kGradChunk;
    GB_HIP(hipMalloc(&b->partial, (size_t)b->partial_blocks * kMlpParams * 4));
    b->cap = cap;
    return GB_OK;
}

int forward_blocks(int M) {
    const int tiles = (M + 15) / 16;
    const int b = (tiles + 3) / 4;
    return b < 1024 ? (b < 1 ? 1 : b) : 1024;
}

gb_status forward(gb_brancher* b, const float* X, int M, float* Y, bool keep, hipStream_t st) {
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(forward_blocks(M)), dim3(256), 0, st, b->params, b->mu,
                       b->sigma, X, M, Y, keep ? b->Xn : nullptr, keep ? b->H1 : nullptr,
                       keep ? b->H2 : nullptr, keep ? b->H3 : nullptr);
    GB_HIP(hipGetLastError());
    return GB_OK;
}

// Loss mean((y-t)^2) and its gradient into b->grad (device).
gb_status loss_grad(gb_brancher* b, const float* X, const float* T, int M, double* loss, hipStream_t st) {
    gb_status s = ensure_cap(b, M);
    if (s != GB_OK) return s;
    s = forward(b, X, M, b->Y, true, st);
    if (s != GB_OK) return s;
    const int bb = (M + 255) / 256 < 1024 ? (M + 255) / 256 : 1024;
    hipLaunchKernelGGL(mlp_backward_dz_kernel, dim3(bb), dim3(256), 0, st, b->params, b->Y, T, M,
                       2.0f / (float)M, b->H1, b->H2, b->H3, b->DY, b->DZ1, b->DZ2, b->DZ3);
    GB_HIP(hipGetLastError());
    const int pb = (M + kGradChunk - 1) / kGradChunk;
    hipLaunchKernelGGL(mlp_weight_grad_kernel, dim3(pb), dim3(256), 0, st, M, kGradChunk, b->Xn, b->H1,
                       b->H2, b->H3, b->DY, b->DZ1, b->DZ2, b->DZ3, b->partial);
    GB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_reduce_partials_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st,
                       b->partial, pb, b->grad);
    GB_HIP(hipGetLastError());
    if (loss) {
        hipLaunchKernelGGL(mlp_sq_err_kernel, dim3(kSqBlocks), dim3(256), 0, st, b->Y, T, M, b->sq);
        GB_HIP(hipGetLastError());
        double h[kSqBlocks];
        GB_HIP(hipMemcpyAsync(h, b->sq, sizeof(h), hipMemcpyDeviceToHost, st));
        GB_HIP(hipStreamSynchronize(st));
        double acc = 0.0;
        for (int i = 0; i < kSqBlocks; ++i) acc += h[i];
        *loss = acc / (double)M;
    }
    return GB_OK;
}

// He initialisation from a SplitMix64 stream (Box-Muller normals).
void init_params(std::vector<float>& p, unsigned long long seed) {
    unsigned long long state = seed;
    auto next = [&]() {
        unsigned long long z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    p.assign(kMlpParams, 0.f);
    for (int i = 0; i < F * H; ++i) p[kOffW1 + i] = (float)(normal() * std::sqrt(2.0 / F));
    for (int i = 0; i < H * H; ++i) p[kOffW2 + i] = (float)(normal() * std::sqrt(2.0 / H));
    for (int i = 0; i < H * H; ++i) p[kOffW3 + i] = (float)(normal() * std::sqrt(2.0 / H));
    for (int i = 0; i < H; ++i) p[kOffW4 + i] = (float)(normal() * std::sqrt(1.0 / H));
}

}  // namespace
}  // namespace gb

using namespace gb;

extern "C" gb_status gb_brancher_create(unsigned long long seed, gb_brancher** out) {
    if (!out) return fail(GB_ERR_INVALID_ARGUMENT, "null output");
    auto* b = new gb_brancher();
    std::vector<float> p;
    init_params(p, seed);
    const std::vector<float> zeros(kMlpParams, 0.f), ones(F, 1.f), zf(F, 0.f);
    GB_HIP(hipMalloc(&b->params, kMlpParams * 4));
    GB_HIP(hipMalloc(&b->m, kMlpParams * 4));
    GB_HIP(hipMalloc(&b->v, kMlpParams * 4));
    GB_HIP(hipMalloc(&b->grad, kMlpParams * 4));
    GB_HIP(hipMalloc(&b->mu, F * 4));
    GB_HIP(hipMalloc(&b->sigma, F * 4));
    GB_HIP(hipMalloc(&b->sq, kSqBlocks * sizeof(double)));
    GB_HIP(hipMemcpy(b->params, p.data(), kMlpParams * 4, hipMemcpyHostToDevice));
    GB_HIP(hipMemcpy(b->m, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice));
    GB_HIP(hipMemcpy(b->v, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice));
    GB_HIP(hipMemcpy(b->mu, zf.data(), F * 4, hipMemcpyHostToDevice));
    GB_HIP(hipMemcpy(b->sigma, ones.data(), F * 4, hipMemcpyHostToDevice));
    *out = b;
    return GB_OK;
}

extern "C" gb_status gb_brancher_destroy(gb_brancher* b) {
    if (!b) return GB_OK;
    (void)hipDeviceSynchronize();
    for (float* p : {b->params, b->m, b->v, b->grad, b->mu, b->sigma, b->Y, b->Xn, b->H1, b->H2, b->H3,
                     b->DY, b->DZ1, b->DZ2, b->DZ3, b->partial})
        if (p) (void)hipFree(p);
    if (b->sq) (void)hipFree(b->sq);
    delete b;
    return GB_OK;
}

extern "C" gb_status gb_brancher_set_params(gb_brancher* b, const float* params, int count) {
    if (!b || !params || count != kMlpParams)
        return fail(GB_ERR_INVALID_ARGUMENT, "parameter count must be GB_MLP_PARAMS");
    GB_HIP(hipMemcpy(b->params, params, kMlpParams * 4, hipMemcpyHostToDevice));
    return GB_OK;
}

extern "C" gb_status gb_brancher_get_params(gb_brancher* b, float* params, int count) {
    if (!b || !params || count != kMlpParams)
        return fail(GB_ERR_INVALID_ARGUMENT, "parameter count must be GB_MLP_PARAMS");
    GB_HIP(hipMemcpy(params, b->params, kMlpParams * 4, hipMemcpyDeviceToHost));
    return GB_OK;
}

extern "C" gb_status gb_brancher_set_normalization(gb_brancher* b, const float* mu, const float* sigma) {
    if (!b || !mu || !sigma) return fail(GB_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < F; ++i)
        if (!(sigma[i] > 0.f) || !std::isfinite(mu[i]))
            return fail(GB_ERR_INVALID_ARGUMENT, "normalization requires finite mu and positive sigma");
    GB_HIP(hipMemcpy(b->mu, mu, F * 4, hipMemcpyHostToDevice));
    GB_HIP(hipMemcpy(b->sigma, sigma, F * 4, hipMemcpyHostToDevice));
    return GB_OK;
}

extern "C" gb_status gb_brancher_predict(gb_brancher* b, const float* X, int M, float* Y, void* stream) {
    if (!b || M < 0 || (M > 0 && (!X || !Y))) return fail(GB_ERR_INVALID_ARGUMENT, "invalid predict arguments");
    if (M == 0) return GB_OK;
    return forward(b, X, M, Y, false, reinterpret_cast<hipStream_t>(stream));
}

extern "C" gb_status gb_brancher_loss_grad(gb_brancher* b, const float* X, const float* T, int M,
                                           double* loss, float* grad_host) {
    if (!b || !X || !T || M <= 0) return fail(GB_ERR_INVALID_ARGUMENT, "invalid loss arguments");
    gb_status s = loss_grad(b, X, T, M, loss, nullptr);
    if (s != GB_OK) return s;
    if (grad_host) GB_HIP(hipMemcpy(grad_host, b->grad, kMlpParams * 4, hipMemcpyDeviceToHost));
    GB_HIP(hipDeviceSynchronize());
    return GB_OK;
}

extern "C" gb_status gb_brancher_train(gb_brancher* b, const float* X, const float* T, int M, int epochs,
                                       int batch, float lr, double* final_loss) {
    if (!b || !X || !T || M <= 0 || epochs < 0 || batch <= 0 || !(lr > 0.f))
        return fail(GB_ERR_INVALID_ARGUMENT, "invalid training arguments");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        for (int s0 = 0; s0 < M; s0 += batch) {
            const int mb = (M - s0) < batch ? (M - s0) : batch;
            gb_status s = loss_grad(b, X + (size_t)s0 * F, T + s0, mb, nullptr, nullptr);
            if (s != GB_OK) return s;
            ++b->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)b->adam_t);
            const float c2 = 1.f - std::pow(b2, (float)b->adam_t);
            hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, nullptr,
                               b->params, b->grad, b->m, b->v, lr, b1, b2, eps, c1, c2);
            GB_HIP(hipGetLastError());
        }
    }
    if (final_loss) return loss_grad(b, X, T, M, final_loss, nullptr);
    GB_HIP(hipDeviceSynchronize());
    return GB_OK;
}
