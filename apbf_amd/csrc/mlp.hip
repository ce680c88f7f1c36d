// Learned lambda warm start: a 16-64-64-1 ReLU network on the matrix cores.
//
// BF16 weights and activations feed the matrix cores with FP32 accumulation:
// gfx950's v_mfma_f32_16x16x32_bf16 for the 64-deep products (layer 2 and
// its input gradient), v_mfma_f32_16x16x16_bf16 for the 16-deep ones (layer
// 1, whose input has 16 features, and the weight gradients, summed over the
// 16 samples of a tile). Biases, the last layer and all reductions stay FP32.
// One 64-lane wavefront handles a tile of 16 samples (the tiles of a block's
// waves are independent: LDS is synchronised per wavefront); with the 16x16x16 operand
// layout (A: lane l holds row l%16, k = 4(l/16)..+3; B: lane l holds column
// l%16, k = 4(l/16)..+3; D: lane l holds column l%16, rows 4(l/16)..+3) the
// layer-1 output tile of a lane is exactly the transposed operand the weight
// gradients need, so the backward pass reuses registers and only the
// layer-2 operand and its gradient go through LDS.
//
// Training keeps FP32 master weights and Adam moments; every step writes
// per-block gradient partials, and a second kernel sums them in block order
// (deterministic) and applies Adam.
#include "apbf_internal.h"
#include "mlp.h"

namespace apbf {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d) {
    bf16x4_t v;
    v[0] = (short)f2bf(a);
    v[1] = (short)f2bf(b);
    v[2] = (short)f2bf(c);
    v[3] = (short)f2bf(d);
    return v;
}
__device__ __forceinline__ f32x4_t mfma(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// gfx950's 16x16x32 form for the 64-deep products (layer 2 and its input
// gradient): lane l holds row / column l%16 and k = 8(l/16) .. +7 of A / B;
// D as in the 16x16x16 form.
__device__ __forceinline__ f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8_t load8(const uint16_t* p) {  // 16-byte aligned
    return *reinterpret_cast<const bf16x8_t*>(p);
}
// A wavefront's LDS tile written by some lanes and read by others (and the
// reverse before the next write): a wavefront-level fence, no block barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int MLP_WAVES = 4;               // waves per block
constexpr int MLP_THREADS = 64 * MLP_WAVES;

// Model in LDS: bf16 W1 [64][16], W2 [64][64] and W2 transposed, fp32 rest.
struct MlpShared {
    alignas(16) uint16_t w1[MLP_H * MLP_IN];
    alignas(16) uint16_t w2[MLP_H * MLP_H];
    alignas(16) uint16_t w2t[MLP_H * MLP_H];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H], b3;
    float mean[MLP_IN], scale[MLP_IN], out_scale;
    alignas(16) uint16_t xfer[MLP_WAVES][16 * MLP_H];  // per-wave transposition tile
};

__device__ void load_model(const float* __restrict__ p, const float* __restrict__ norm, MlpShared& sh) {
    for (int t = threadIdx.x; t < MLP_H * MLP_IN; t += blockDim.x) sh.w1[t] = f2bf(p[MLP_OFF_W1 + t]);
    for (int t = threadIdx.x; t < MLP_H * MLP_H; t += blockDim.x) {
        const uint16_t v = f2bf(p[MLP_OFF_W2 + t]);
        sh.w2[t] = v;
        sh.w2t[(t % MLP_H) * MLP_H + t / MLP_H] = v;  // w2t[k][n] = w2[n][k]
    }
    for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
        sh.b1[t] = p[MLP_OFF_B1 + t];
        sh.b2[t] = p[MLP_OFF_B2 + t];
        sh.w3[t] = p[MLP_OFF_W3 + t];
    }
    for (int t = threadIdx.x; t < MLP_IN; t += blockDim.x) {
        sh.mean[t] = norm[t];
        sh.scale[t] = norm[MLP_IN + t];
    }
    if (threadIdx.x == 0) {
        sh.b3 = p[MLP_OFF_B3];
        sh.out_scale = norm[2 * MLP_IN];
    }
    __syncthreads();
}

// Forward state of one 16-sample tile (register tiles in the D layout).
struct Tile {
    f32x4_t z1[4], z2[4];  // pre-activations: tile t holds hidden units 16t + l%16
    float y[4];            // network output for samples 4(l/16) + r (all lanes of the group)
};

__device__ __forceinline__ float feat(const float* __restrict__ x, uint32_t s, uint32_t n, int f,
                                      const MlpShared& sh) {
    return s < n ? (x[(size_t)s * MLP_IN + f] - sh.mean[f]) * sh.scale[f] : 0.0f;
}

// Forward pass of samples base .. base+15 by the calling wave.
__device__ void tile_forward(const float* __restrict__ x, uint32_t n, uint32_t base, MlpShared& sh, int wave,
                             Tile& T) {
    const int l = threadIdx.x & 63;
    const int lo = l & 15, hi = l >> 4;
    // layer 1: A = normalized features (sample lo, features 4hi..4hi+3)
    const uint32_t s = base + lo;
    const bf16x4_t a = pack4(feat(x, s, n, 4 * hi, sh), feat(x, s, n, 4 * hi + 1, sh), feat(x, s, n, 4 * hi + 2, sh),
                             feat(x, s, n, 4 * hi + 3, sh));
    uint16_t* xf = sh.xfer[wave];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int nn = 16 * t + lo;
        const bf16x4_t b = *reinterpret_cast<const bf16x4_t*>(&sh.w1[nn * MLP_IN + 4 * hi]);
        const float bias = sh.b1[nn];
        f32x4_t c = {bias, bias, bias, bias};
        T.z1[t] = mfma(a, b, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) xf[(4 * hi + r) * MLP_H + nn] = f2bf(fmaxf(T.z1[t][r], 0.0f));
    }
    wave_lds_sync();
    // layer 2: A = relu(z1) (sample lo, hidden 32kt + 8hi ..), 16x16x32 products
    bf16x8_t a2[2];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) a2[kt] = load8(&xf[lo * MLP_H + 32 * kt + 8 * hi]);
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int nn = 16 * t + lo;
        const float bias = sh.b2[nn];
        f32x4_t c = {bias, bias, bias, bias};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) c = mfma32(a2[kt], load8(&sh.w2[nn * MLP_H + 32 * kt + 8 * hi]), c);
        T.z2[t] = c;
    }
    // layer 3 (FP32): y = relu(z2) . w3 + b3, reduced over the 16 lanes of a row group
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float p = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) p += fmaxf(T.z2[t][r], 0.0f) * sh.w3[16 * t + lo];
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) p += __shfl_xor(p, m, 64);
        T.y[r] = p + sh.b3;
    }
}

// Inference: out[s] = network(features[s]) * out_scale for s < *d_len.
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_infer(const float* __restrict__ params,
                                                           const float* __restrict__ norm,
                                                           const float* __restrict__ x,
                                                           const uint32_t* __restrict__ d_len,
                                                           float* __restrict__ out) {
    __shared__ MlpShared sh;
    load_model(params, norm, sh);
    const uint32_t n = *d_len;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    // every wave of a block runs the same number of tiles (LDS barriers inside)
    for (uint32_t tile0 = blockIdx.x * MLP_WAVES; tile0 * 16 < n; tile0 += gridDim.x * MLP_WAVES) {
        const uint32_t base = (tile0 + wave) * 16;
        Tile T;
        tile_forward(x, n, base, sh, wave, T);
        if ((l & 15) == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t s = base + 4 * (l >> 4) + r;
                if (s < n) out[s] = T.y[r] * sh.out_scale;
            }
        }
    }
}

// One training step: gradient of 0.5 * mean((y - t/out_scale)^2) over the
// batch perm[start .. start+batch) (indices wrap modulo n), per-block partials
// partial[block][0 .. MLP_PARAMS) and the loss partial at [MLP_PARAMS].
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_grad(const float* __restrict__ params,
                                                          const float* __restrict__ norm,
                                                          const float* __restrict__ target,
                                                          const uint32_t* __restrict__ perm, uint32_t n,
                                                          uint32_t start, uint32_t batch, const float* __restrict__ xs,
                                                          float* __restrict__ partial) {
    __shared__ MlpShared sh;
    __shared__ float acc[MLP_PARAMS + 1];
    load_model(params, norm, sh);
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lo = l & 15, hi = l >> 4;
    const float inv_b = 1.0f / (float)batch;
    f32x4_t gw1[4], gw2[4][4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0.0f, loss = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gw1[t] = f32x4_t{0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < 4; ++u) gw2[t][u] = f32x4_t{0, 0, 0, 0};
    }
    uint16_t* xf = sh.xfer[wave];
    const uint32_t tiles = (batch + 15) / 16;
    for (uint32_t tile0 = blockIdx.x * MLP_WAVES; tile0 < tiles; tile0 += gridDim.x * MLP_WAVES) {
        const uint32_t tb = (tile0 + wave) * 16;  // batch-local offset of this wave's tile
        Tile T;
        tile_forward(xs, batch, tb, sh, wave, T);  // xs: the batch gathered into consecutive rows
        // output gradient per sample 4hi + r
        float dy[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t b = tb + 4 * hi + r;
            float d = 0.0f;
            if (b < batch) {
                const uint32_t s = perm[(start + b) % n];
                d = T.y[r] - target[s] / sh.out_scale;
                if (lo == 0) loss += 0.5f * d * d * inv_b;
            }
            dy[r] = d * inv_b;
            if (lo == 0) gb3 += dy[r];
        }
        // layer 3 and layer-2 pre-activation gradient (D layout)
        f32x4_t dz2[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float w = sh.w3[16 * t + lo];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float h2 = fmaxf(T.z2[t][r], 0.0f);
                gw3[t] += h2 * dy[r];
                const float g = T.z2[t][r] > 0.0f ? dy[r] * w : 0.0f;
                dz2[t][r] = g;
                gb2[t] += g;
            }
        }
        // dW2[k][n] += sum_s relu(z1)[s][k] dz2[s][n]: A = relu(z1) tiles (already in the A^T
        // layout), B = dz2 tiles (already in the B layout)
        bf16x4_t h1b[4], dz2b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            h1b[t] = pack4(fmaxf(T.z1[t][0], 0.f), fmaxf(T.z1[t][1], 0.f), fmaxf(T.z1[t][2], 0.f),
                           fmaxf(T.z1[t][3], 0.f));
            dz2b[t] = pack4(dz2[t][0], dz2[t][1], dz2[t][2], dz2[t][3]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gw2[mt][nt] = mfma(h1b[mt], dz2b[nt], gw2[mt][nt]);
        // dh1 = dz2 . W2 (A = dz2 through LDS: sample lo, n = 16kt + 4hi ..; B = W2 rows)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) xf[(4 * hi + r) * MLP_H + 16 * t + lo] = (uint16_t)dz2b[t][r];
        wave_lds_sync();
        bf16x8_t ad[2];
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) ad[kt] = load8(&xf[lo * MLP_H + 32 * kt + 8 * hi]);
        wave_lds_sync();
        f32x4_t dz1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4_t c = {0, 0, 0, 0};
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)  // B[n][col k]: lane holds W2[n = 32kt + 8hi ..][k = 16t + lo] = w2t[k][n ..]
                c = mfma32(ad[kt], load8(&sh.w2t[(16 * t + lo) * MLP_H + 32 * kt + 8 * hi]), c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float g = T.z1[t][r] > 0.0f ? c[r] : 0.0f;
                dz1[t][r] = g;
                gb1[t] += g;
            }
        }
        // dW1[f][k] += sum_s x[s][f] dz1[s][k]: A = x^T (feature lo, samples 4hi ..)
        float xv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) xv[r] = feat(xs, tb + 4 * hi + r, batch, lo, sh);
        const bf16x4_t xa = pack4(xv[0], xv[1], xv[2], xv[3]);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            gw1[t] = mfma(xa, pack4(dz1[t][0], dz1[t][1], dz1[t][2], dz1[t][3]), gw1[t]);
    }
    // reduce the per-lane bias / layer-3 partials over the four row groups
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gb1[t] += __shfl_xor(gb1[t], 16, 64);
        gb1[t] += __shfl_xor(gb1[t], 32, 64);
        gb2[t] += __shfl_xor(gb2[t], 16, 64);
        gb2[t] += __shfl_xor(gb2[t], 32, 64);
        gw3[t] += __shfl_xor(gw3[t], 16, 64);
        gw3[t] += __shfl_xor(gw3[t], 32, 64);
    }
    gb3 += __shfl_xor(gb3, 16, 64);
    gb3 += __shfl_xor(gb3, 32, 64);
    loss += __shfl_xor(loss, 16, 64);
    loss += __shfl_xor(loss, 32, 64);
    // block sum in wave order (deterministic)
    for (int w = 0; w < MLP_WAVES; ++w) {
        if (wave == w) {
            const bool first = w == 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    // gW1 tile t: lane holds [f = 4hi + r][k = 16t + lo] -> W1[k][f]
                    const int i1 = MLP_OFF_W1 + (16 * t + lo) * MLP_IN + 4 * hi + r;
                    acc[i1] = first ? gw1[t][r] : acc[i1] + gw1[t][r];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        // gW2 tile (t, u): [k = 16t + 4hi + r][n = 16u + lo] -> W2[n][k]
                        const int i2 = MLP_OFF_W2 + (16 * u + lo) * MLP_H + 16 * t + 4 * hi + r;
                        acc[i2] = first ? gw2[t][u][r] : acc[i2] + gw2[t][u][r];
                    }
                }
                if (hi == 0) {
                    const int ib1 = MLP_OFF_B1 + 16 * t + lo, ib2 = MLP_OFF_B2 + 16 * t + lo,
                              iw3 = MLP_OFF_W3 + 16 * t + lo;
                    acc[ib1] = first ? gb1[t] : acc[ib1] + gb1[t];
                    acc[ib2] = first ? gb2[t] : acc[ib2] + gb2[t];
                    acc[iw3] = first ? gw3[t] : acc[iw3] + gw3[t];
                }
            }
            if (l == 0) {
                acc[MLP_OFF_B3] = first ? gb3 : acc[MLP_OFF_B3] + gb3;
                acc[MLP_PARAMS] = first ? loss : acc[MLP_PARAMS] + loss;
            }
        }
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.x * (MLP_PARAMS + 1);
    for (int t = threadIdx.x; t < MLP_PARAMS + 1; t += blockDim.x) out[t] = acc[t];
}

// Gathers the batch rows perm[start ..] of x into consecutive rows of xs.
__global__ __launch_bounds__(BLOCK) void k_mlp_gather(const float* __restrict__ x, const uint32_t* __restrict__ perm,
                                                      uint32_t n, uint32_t start, uint32_t batch,
                                                      float* __restrict__ xs) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < batch * MLP_IN; t += gridDim.x * blockDim.x) {
        const uint32_t b = t / MLP_IN, f = t % MLP_IN;
        xs[t] = x[(size_t)perm[(start + b) % n] * MLP_IN + f];
    }
}

// Sums the block partials (in block order) into grad[0 .. MLP_PARAMS] and,
// when lr > 0, applies one Adam step to the FP32 master parameters. The step
// number for the bias corrections is *step_base + k + 1: the device keeps the
// count, so captured steps replay with the right corrections.
__global__ __launch_bounds__(BLOCK) void k_mlp_adam(const float* __restrict__ partial, uint32_t blocks,
                                                    float* __restrict__ grad, float* __restrict__ params,
                                                    float* __restrict__ m1, float* __restrict__ m2, float lr,
                                                    float beta1, float beta2, float eps,
                                                    const uint32_t* __restrict__ step_base, uint32_t k,
                                                    float* __restrict__ loss_out) {
    const float t = (float)(*step_base + k + 1u);
    const float bc1 = 1.0f - powf(beta1, t), bc2 = 1.0f - powf(beta2, t);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= (uint32_t)MLP_PARAMS; p += gridDim.x * blockDim.x) {
        float g = 0.0f;
#pragma unroll 8
        for (uint32_t b = 0; b < blocks; ++b) g += partial[(size_t)b * (MLP_PARAMS + 1) + p];
        if (p == (uint32_t)MLP_PARAMS) {
            if (loss_out) *loss_out = g;
            continue;
        }
        grad[p] = g;
        if (lr > 0.0f) {
            const float a = beta1 * m1[p] + (1.0f - beta1) * g;
            const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
            m1[p] = a;
            m2[p] = v;
            params[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
        }
    }
}

__global__ void k_mlp_steps_done(uint32_t* __restrict__ step_base, uint32_t steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_base += steps;
}

// ------------------------------------------------------------------- host
static uint32_t mlp_grid_tiles(uint32_t samples) {
    uint32_t tiles = (samples + 15) / 16;
    uint32_t g = (tiles + MLP_WAVES - 1) / MLP_WAVES;
    return g < 1024u ? (g ? g : 1u) : 1024u;
}

void launch_mlp_infer(const float* params, const float* norm, const float* x, const uint32_t* d_len, uint32_t cap,
                      float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_grid_tiles(cap)), dim3(MLP_THREADS), 0, s, params, norm, x, d_len, out);
}

}  // namespace apbf

using namespace apbf;


// The captured training launches hold the buffers' addresses: any
// reallocation drops them.
static void drop_training_graph(apbf_mlp* m) {
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
}

static apbf_status mlp_reserve(apbf_mlp* m, uint32_t batch, uint32_t steps, uint32_t n) {
    const uint32_t blocks = mlp_grid_tiles(batch);
    if (batch > m->max_batch || blocks > m->max_blocks || steps > m->max_steps || n > m->max_n)
        drop_training_graph(m);
    if (batch > m->max_batch) {
        if (m->xs) (void)hipFree(m->xs);
        if (hipMalloc(&m->xs, (size_t)batch * MLP_IN * 4) != hipSuccess) return APBF_ERR_HIP;
        m->max_batch = batch;
    }
    if (blocks > m->max_blocks) {
        if (m->partial) (void)hipFree(m->partial);
        if (hipMalloc(&m->partial, (size_t)blocks * (MLP_PARAMS + 1) * 4) != hipSuccess) return APBF_ERR_HIP;
        m->max_blocks = blocks;
    }
    if (steps > m->max_steps) {
        if (m->loss) (void)hipFree(m->loss);
        if (hipMalloc(&m->loss, (size_t)steps * 4) != hipSuccess) return APBF_ERR_HIP;
        m->max_steps = steps;
    }
    if (n > m->max_n) {
        if (m->perm) (void)hipFree(m->perm);
        if (hipMalloc(&m->perm, (size_t)n * 4) != hipSuccess) return APBF_ERR_HIP;
        m->max_n = n;
    }
    return APBF_OK;
}

// The launches of one training call (steps of gather, gradient and Adam,
// then the step counter), on stream s.
static void enqueue_training(apbf_mlp* m, const float* features, const float* targets, const uint32_t* perm,
                             uint32_t n, uint32_t batch, uint32_t steps, float lr, hipStream_t s) {
    const uint32_t blocks = mlp_grid_tiles(batch);
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (uint32_t k = 0; k < (steps ? steps : 1u); ++k) {
        const uint32_t start = (uint32_t)(((uint64_t)k * batch) % n);
        hipLaunchKernelGGL(k_mlp_gather, dim3(stride_grid((uint64_t)batch * MLP_IN)), dim3(BLOCK), 0, s, features, perm,
                           n, start, batch, m->xs);
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(MLP_THREADS), 0, s, m->params, m->norm, targets, perm, n,
                           start, batch, m->xs, m->partial);
        hipLaunchKernelGGL(k_mlp_adam, dim3(div_up(MLP_PARAMS + 1, BLOCK)), dim3(BLOCK), 0, s, m->partial, blocks,
                           m->grad, m->params, m->m1, m->m2, steps ? lr : 0.0f, beta1, beta2, eps, m->step_base, k,
                           m->loss + k);
    }
    if (steps) hipLaunchKernelGGL(k_mlp_steps_done, dim3(1), dim3(64), 0, s, m->step_base, steps);
}

extern "C" {

apbf_status apbf_mlp_create(apbf_mlp** out) {
    if (!out) return APBF_ERR_INVALID_ARGUMENT;
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev <= 0) return APBF_ERR_NO_DEVICE;
    apbf_mlp* m = new apbf_mlp();
    if (hipMalloc(&m->params, MLP_PARAMS * 4) != hipSuccess || hipMalloc(&m->norm, (2 * MLP_IN + 1) * 4) != hipSuccess ||
        hipMalloc(&m->m1, MLP_PARAMS * 4) != hipSuccess || hipMalloc(&m->m2, MLP_PARAMS * 4) != hipSuccess ||
        hipMalloc(&m->grad, MLP_PARAMS * 4) != hipSuccess || hipMalloc(&m->step_base, 4) != hipSuccess) {
        apbf_mlp_destroy(m);
        return APBF_ERR_HIP;
    }
    (void)hipMemset(m->step_base, 0, 4);
    (void)hipMemset(m->m1, 0, MLP_PARAMS * 4);
    (void)hipMemset(m->m2, 0, MLP_PARAMS * 4);
    (void)hipMemset(m->params, 0, MLP_PARAMS * 4);
    *out = m;
    return hipDeviceSynchronize() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_mlp_destroy(apbf_mlp* m) {
    if (!m) return APBF_ERR_INVALID_ARGUMENT;
    drop_training_graph(m);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    float* ps[] = {m->params, m->norm, m->m1, m->m2, m->grad, m->partial, m->xs, m->loss};
    for (float* p : ps)
        if (p) (void)hipFree(p);
    if (m->perm) (void)hipFree(m->perm);
    if (m->step_base) (void)hipFree(m->step_base);
    delete m;
    return APBF_OK;
}

apbf_status apbf_mlp_set_params(apbf_mlp* m, const float* host_params, const float* host_norm) {
    if (!m || !host_params || !host_norm) return APBF_ERR_INVALID_ARGUMENT;
    if (hipMemcpy(m->params, host_params, MLP_PARAMS * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->norm, host_norm, (2 * MLP_IN + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, MLP_PARAMS * 4) != hipSuccess || hipMemset(m->m2, 0, MLP_PARAMS * 4) != hipSuccess ||
        hipMemset(m->step_base, 0, 4) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_mlp_get_params(apbf_mlp* m, float* host_params, float* host_norm) {
    if (!m) return APBF_ERR_INVALID_ARGUMENT;
    if ((host_params && hipMemcpy(host_params, m->params, MLP_PARAMS * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
        (host_norm && hipMemcpy(host_norm, m->norm, (2 * MLP_IN + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess))
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_mlp_forward(apbf_mlp* m, const float* features, const uint32_t* d_len, uint32_t capacity,
                             float* out, apbf_stream stream) {
    if (!m || !features || !d_len || !out) return APBF_ERR_INVALID_ARGUMENT;
    launch_mlp_infer(m->params, m->norm, features, d_len, capacity, out, reinterpret_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_mlp_train(apbf_mlp* m, const float* features, const float* targets, const uint32_t* perm,
                           uint32_t n, uint32_t batch, uint32_t steps, float lr, float* host_losses,
                           apbf_stream stream) {
    if (!m || !features || !targets || !perm || n == 0 || batch == 0 || batch > n) return APBF_ERR_INVALID_ARGUMENT;
    apbf_status st = mlp_reserve(m, batch, steps ? steps : 1, n);
    if (st != APBF_OK) return st;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // The steps read the sample order from the network's own copy, so the
    // same captured launches serve every call with the same arguments.
    if (hipMemcpyAsync(m->perm, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return APBF_ERR_HIP;
    if (steps < 16) {
        enqueue_training(m, features, targets, m->perm, n, batch, steps, lr, s);
    } else {
        // many steps: one graph of the call's launches on the network's own
        // stream (captured at the first call with these arguments), ordered
        // after the caller's stream and before its later work
        if (!m->stream && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return APBF_ERR_HIP;
        if (!(m->graph && m->g_features == features && m->g_targets == targets && m->g_n == n &&
              m->g_batch == batch && m->g_steps == steps && m->g_lr == lr)) {
            drop_training_graph(m);
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return APBF_ERR_HIP;
            enqueue_training(m, features, targets, m->perm, n, batch, steps, lr, m->stream);
            if (hipStreamEndCapture(m->stream, &g) != hipSuccess || !g) return APBF_ERR_HIP;
            const bool ok = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0) == hipSuccess;
            (void)hipGraphDestroy(g);
            if (!ok) {
                m->graph = nullptr;
                return APBF_ERR_HIP;
            }
            m->g_features = features;
            m->g_targets = targets;
            m->g_n = n;
            m->g_batch = batch;
            m->g_steps = steps;
            m->g_lr = lr;
        }
        hipEvent_t before = nullptr, after = nullptr;
        if (hipEventCreateWithFlags(&before, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&after, hipEventDisableTiming) != hipSuccess)
            return APBF_ERR_HIP;
        const bool ok = hipEventRecord(before, s) == hipSuccess && hipStreamWaitEvent(m->stream, before, 0) == hipSuccess &&
                        hipGraphLaunch(m->graph, m->stream) == hipSuccess && hipEventRecord(after, m->stream) == hipSuccess &&
                        hipStreamWaitEvent(s, after, 0) == hipSuccess;
        (void)hipEventDestroy(before);
        (void)hipEventDestroy(after);
        if (!ok) return APBF_ERR_HIP;
    }
    if (hipGetLastError() != hipSuccess) return APBF_ERR_HIP;
    if (host_losses) {
        if (hipMemcpyAsync(host_losses, m->loss, (size_t)(steps ? steps : 1) * 4, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return APBF_ERR_HIP;
    }
    return APBF_OK;
}

apbf_status apbf_mlp_gradient(apbf_mlp* m, const float* features, const float* targets, const uint32_t* perm,
                              uint32_t n, float* host_grad, float* host_loss, apbf_stream stream) {
    if (!m || !host_grad) return APBF_ERR_INVALID_ARGUMENT;
    apbf_status st = apbf_mlp_train(m, features, targets, perm, n, n, 0, 0.0f, host_loss, stream);
    if (st != APBF_OK) return st;
    if (hipMemcpy(host_grad, m->grad, MLP_PARAMS * 4, hipMemcpyDeviceToHost) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

}  // extern "C"
