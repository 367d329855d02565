// Learned subgrid-stress closure on bfloat16 matrix cores (see sgs_mlp.hpp).
//
// Matrix products use v_mfma_f32_16x16x16_bf16: one wavefront computes a
// 16x16 FP32 tile D += A B over K = 16; lane l holds A(i, k) for
// i = l % 16, k = 4 (l / 16) + j, B(k, n) for n = l % 16 and the same k, and
// D(i, n) for n = l % 16, i = 4 (l / 16) + j (j = 0..3).  Training reads
// both bfloat16 operands from LDS; inference keeps the weight operands in
// registers.  Accumulation, biases, GELU and its derivative, the loss and all
// reductions are FP32 in a fixed order, so a training run repeats bit for
// bit.
#include <stdexcept>

#include "sgs_mlp.hpp"

namespace hit3d {

namespace {

using namespace mlp;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TS = 32;      // samples per training tile
constexpr int NT = 256;     // threads per tile (4 wavefronts)

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = expf(-0.5f * z * z) * 0.39894228040143268f;
    return cdf + z * pdf;
}

// D(i0.., n0..) = sum_k A(i, k) B(k, n) over K (a multiple of 16);
// A(i, k) = TA ? A[k*lda + i] : A[i*lda + k], B(k, n) = TB ? B[n*ldb + k] : B[k*ldb + n].
template <bool TA, bool TB>
__device__ __forceinline__ f32x4 mma(const uint16_t* A, int lda, const uint16_t* B, int ldb, int i0,
                                     int n0, int K) {
    const int l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * gq + j;
            a[j] = (short)(TA ? A[k * lda + i0 + r] : A[(i0 + r) * lda + k]);
            b[j] = (short)(TB ? B[(n0 + r) * ldb + k] : B[k * ldb + n0 + r]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// sum of a per-lane value over the four lane groups sharing l % 16
__device__ __forceinline__ float group_sum(float x) {
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    return x;
}

struct Lds {
    uint16_t xb[TS * KIN];
    uint16_t w1[KIN * NH];
    uint16_t w2[NH * NH];
    uint16_t w3[NH * KOUT];
    float g1[TS * NH];   // gelu'(Z1)
    float g2[TS * NH];   // gelu'(Z2)
    uint16_t h1[TS * NH];  // H1, later dZ1
    uint16_t h2[TS * NH];  // H2, later dZ2
    uint16_t dy[TS * KOUT];
    float b1[NH], b2[NH], b3[KOUT];
    float red[4][NH];
    float lred[4];
};

__device__ void stage_weights(Lds& s, const uint16_t* wb, const float* p) {
    for (int i = threadIdx.x; i < KIN * NH; i += NT) s.w1[i] = wb[OW1 + i];
    for (int i = threadIdx.x; i < NH * NH; i += NT) s.w2[i] = wb[OW2 + i];
    for (int i = threadIdx.x; i < NH * KOUT; i += NT) s.w3[i] = wb[OW3 + i];
    for (int i = threadIdx.x; i < NH; i += NT) {
        s.b1[i] = p[OB1 + i];
        s.b2[i] = p[OB2 + i];
    }
    for (int i = threadIdx.x; i < KOUT; i += NT) s.b3[i] = p[OB3 + i];
}

// forward through the two hidden layers for the TS samples staged in s.xb,
// keeping gelu'(Z) in s.g1, s.g2 for the backward pass
__device__ void hidden_layers(Lds& s) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    const int n0 = 16 * wv;
    for (int mt = 0; mt < TS / 16; ++mt) {
        const f32x4 acc = mma<false, false>(s.xb, KIN, s.w1, NH, 16 * mt, n0, KIN);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 16 * mt + 4 * gq + j, n = n0 + r;
            const float z = acc[j] + s.b1[n];
            s.g1[i * NH + n] = gelu_grad(z);
            s.h1[i * NH + n] = to_bf16(gelu(z));
        }
    }
    __syncthreads();
    for (int mt = 0; mt < TS / 16; ++mt) {
        const f32x4 acc = mma<false, false>(s.h1, NH, s.w2, NH, 16 * mt, n0, NH);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 16 * mt + 4 * gq + j, n = n0 + r;
            const float z = acc[j] + s.b2[n];
            s.g2[i * NH + n] = gelu_grad(z);
            s.h2[i * NH + n] = to_bf16(gelu(z));
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(NT) void k_train(const float* __restrict__ X, const float* __restrict__ T,
                                              int64_t count, int64_t start, int64_t nb_total,
                                              const uint16_t* __restrict__ wb, const float* __restrict__ p,
                                              float* __restrict__ partial, float* __restrict__ lossp) {
    __shared__ Lds s;
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    const int64_t first = start + (int64_t)blockIdx.x * TS;
    for (int i = threadIdx.x; i < TS * KIN; i += NT) {
        const int row = i / KIN, c = i % KIN;
        const int64_t q = (first + row) % count;
        s.xb[i] = c < NIN ? to_bf16(X[q * NIN + c]) : (uint16_t)0;
    }
    stage_weights(s, wb, p);
    __syncthreads();
    hidden_layers(s);
    float* part = partial + (int64_t)blockIdx.x * NPARAM;
    // output layer Y = H2 W3 + b3, loss and dY (waves 0, 1: one row tile each)
    const float inv = 2.0f / (float)(nb_total * NOUT);
    float lsum = 0.f, csum = 0.f;
    if (wv < TS / 16) {
        const f32x4 acc = mma<false, false>(s.h2, NH, s.w3, KOUT, 16 * wv, 0, NH);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 16 * wv + 4 * gq + j, n = r;
            float dy = 0.f;
            if (n < NOUT) {
                const int64_t q = (first + i) % count;
                const float d = (acc[j] + s.b3[n]) - T[q * NOUT + n];
                lsum += d * d;
                dy = inv * d;
            }
            csum += dy;
            s.dy[i * KOUT + n] = to_bf16(dy);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    csum = group_sum(csum);
    if (l == 0) s.lred[wv] = lsum;
    if (gq == 0) s.red[wv][r] = csum;
    __syncthreads();
    if (threadIdx.x < KOUT) {
        float t = 0.f;
        for (int w = 0; w < TS / 16; ++w) t += s.red[w][threadIdx.x];
        part[OB3 + threadIdx.x] = t;
    }
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < 4; ++w) t += (w < TS / 16) ? s.lred[w] : 0.f;
        lossp[blockIdx.x] = t;
    }
    // dW3 = H2^T dY: 64 x 16, one row tile per wave
    {
        const f32x4 acc = mma<true, false>(s.h2, NH, s.dy, KOUT, 16 * wv, 0, TS);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[OW3 + (16 * wv + 4 * gq + j) * KOUT + r] = acc[j];
    }
    __syncthreads();  // h2 is overwritten by dZ2 below
    // dZ2 = (dY W3^T) * gelu'(Z2); db2
    {
        float cs = 0.f;
        for (int mt = 0; mt < TS / 16; ++mt) {
            const f32x4 acc = mma<false, true>(s.dy, KOUT, s.w3, KOUT, 16 * mt, 16 * wv, KOUT);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = 16 * mt + 4 * gq + j, n = 16 * wv + r;
                const float dz = acc[j] * s.g2[i * NH + n];
                cs += dz;
                s.h2[i * NH + n] = to_bf16(dz);
            }
        }
        cs = group_sum(cs);
        if (gq == 0) part[OB2 + 16 * wv + r] = cs;
    }
    __syncthreads();
    // dW2 = H1^T dZ2: 64 x 64, column tile per wave
    for (int mt = 0; mt < NH / 16; ++mt) {
        const f32x4 acc = mma<true, false>(s.h1, NH, s.h2, NH, 16 * mt, 16 * wv, TS);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[OW2 + (16 * mt + 4 * gq + j) * NH + 16 * wv + r] = acc[j];
    }
    // dZ1 = (dZ2 W2^T) * gelu'(Z1); db1  (written into h1 after everyone read it)
    f32x4 dz1[TS / 16];
    for (int mt = 0; mt < TS / 16; ++mt) dz1[mt] = mma<false, true>(s.h2, NH, s.w2, NH, 16 * mt, 16 * wv, NH);
    __syncthreads();
    {
        float cs = 0.f;
        for (int mt = 0; mt < TS / 16; ++mt) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = 16 * mt + 4 * gq + j, n = 16 * wv + r;
                const float dz = dz1[mt][j] * s.g1[i * NH + n];
                cs += dz;
                s.h1[i * NH + n] = to_bf16(dz);
            }
        }
        cs = group_sum(cs);
        if (gq == 0) part[OB1 + 16 * wv + r] = cs;
    }
    __syncthreads();
    // dW1 = X^T dZ1: 16 x 64, column tile per wave
    {
        const f32x4 acc = mma<true, false>(s.xb, KIN, s.h1, NH, 0, 16 * wv, TS);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[OW1 + (4 * gq + j) * NH + 16 * wv + r] = acc[j];
    }
}

// Gradient reduction in two fixed-order levels: the tiles fall into
// kReduceGroups contiguous groups, each summed in tile order (level 1, one
// block row per group), then the group sums in group order (level 2).
__global__ void k_reduce_groups(const float* __restrict__ partial, const float* __restrict__ lossp,
                                int tiles, float* __restrict__ red) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = (tiles + kReduceGroups - 1) / kReduceGroups;
    const int t0 = blockIdx.y * chunk, t1 = min(tiles, t0 + chunk);
    float s = 0.f;
    if (k < NPARAM) {
        for (int t = t0; t < t1; ++t) s += partial[(int64_t)t * NPARAM + k];
    } else if (k == NPARAM) {
        for (int t = t0; t < t1; ++t) s += lossp[t];
    }
    if (k <= NPARAM) red[(int64_t)blockIdx.y * (NPARAM + 1) + k] = s;
}

__global__ void k_reduce_final(const float* __restrict__ red, float* __restrict__ grads,
                               float* __restrict__ loss) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > NPARAM) return;
    float s = 0.f;
    for (int gi = 0; gi < kReduceGroups; ++gi) s += red[(int64_t)gi * (NPARAM + 1) + k];
    if (k < NPARAM) grads[k] = s;
    else *loss = s;
}

__global__ void k_adam(float* p, uint16_t* wb, const float* g, float* m, float* v, float bc1, float bc2,
                       float lr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float w = p[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    p[k] = w;
    wb[k] = to_bf16(w);
}

__global__ void k_to_bf16(const float* p, uint16_t* wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM) wb[k] = to_bf16(p[k]);
}

// ---- inference: weights in registers
// Each wavefront keeps the B operands of its 16 output columns of every
// layer in registers for all the tiles of its block (A operands come from
// LDS as one 8-byte read per lane per product); the products, their order
// and the FP32 epilogues are those of hidden_layers/output_layer, so the
// results are identical.
struct FwdRegs {
    bf16x4 w1, w2[NH / 16], w3[NH / 16];
    float b1, b2, b3;
};

__device__ FwdRegs load_fwd_regs(const uint16_t* __restrict__ wb, const float* __restrict__ p) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    const int n0 = 16 * wv;
    FwdRegs w;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        w.w1[j] = (short)wb[OW1 + (4 * gq + j) * NH + n0 + r];
#pragma unroll
        for (int ks = 0; ks < NH / 16; ++ks) {
            w.w2[ks][j] = (short)wb[OW2 + (16 * ks + 4 * gq + j) * NH + n0 + r];
            w.w3[ks][j] = (short)wb[OW3 + (16 * ks + 4 * gq + j) * KOUT + r];
        }
    }
    w.b1 = p[OB1 + n0 + r];
    w.b2 = p[OB2 + n0 + r];
    w.b3 = p[OB3 + r];
    return w;
}

// D(i0.., n) = sum_k A(i, k) B(k, n), A row-major in LDS (lda a multiple of 4,
// 8-byte aligned rows), B fragments in registers
template <int KSTEPS>
__device__ __forceinline__ f32x4 mma_reg(const uint16_t* A, int lda, int i0, const bf16x4* b) {
    const int l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(A + (i0 + r) * lda + 16 * ks + 4 * gq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[ks], acc, 0, 0, 0);
    }
    return acc;
}

// LDS of an inference tile
struct FwdLds {
    alignas(16) uint16_t xb[TS * KIN];
    alignas(16) uint16_t h1[TS * NH];
    alignas(16) uint16_t h2[TS * NH];
};

__device__ void hidden_layers_reg(FwdLds& s, const FwdRegs& w) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    const int n0 = 16 * wv;
#pragma unroll
    for (int mt = 0; mt < TS / 16; ++mt) {
        const f32x4 acc = mma_reg<1>(s.xb, KIN, 16 * mt, &w.w1);
#pragma unroll
        for (int j = 0; j < 4; ++j) s.h1[(16 * mt + 4 * gq + j) * NH + n0 + r] = to_bf16(gelu(acc[j] + w.b1));
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < TS / 16; ++mt) {
        const f32x4 acc = mma_reg<NH / 16>(s.h1, NH, 16 * mt, w.w2);
#pragma unroll
        for (int j = 0; j < 4; ++j) s.h2[(16 * mt + 4 * gq + j) * NH + n0 + r] = to_bf16(gelu(acc[j] + w.b2));
    }
    __syncthreads();
}

template <typename F>
__device__ void output_layer_reg(FwdLds& s, const FwdRegs& w, F&& f) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, gq = l >> 4;
    if (wv < TS / 16) {
        const f32x4 acc = mma_reg<NH / 16>(s.h2, NH, 16 * wv, w.w3);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (r < NOUT) f(16 * wv + 4 * gq + j, r, acc[j] + w.b3);
    }
}

// inference kernels run this many tiles per block
constexpr int FWD_TILES = 4;

__global__ __launch_bounds__(NT) void k_forward(const float* __restrict__ X, int64_t count,
                                                const uint16_t* __restrict__ wb, const float* __restrict__ p,
                                                float* __restrict__ Y) {
    __shared__ FwdLds s;
    const FwdRegs w = load_fwd_regs(wb, p);
    for (int it = 0; it < FWD_TILES; ++it) {
        const int64_t first = ((int64_t)blockIdx.x * FWD_TILES + it) * TS;
        if (first >= count) break;  // the same for the whole block
        for (int i = threadIdx.x; i < TS * KIN; i += NT) {
            const int row = i / KIN, c = i % KIN;
            const int64_t q = first + row;
            s.xb[i] = (c < NIN && q < count) ? to_bf16(X[q * NIN + c]) : (uint16_t)0;
        }
        __syncthreads();
        hidden_layers_reg(s, w);
        output_layer_reg(s, w, [&](int i, int n, float y) {
            if (first + i < count) Y[(first + i) * NOUT + n] = y;
        });
        __syncthreads();  // s.xb and s.h2 are rewritten by the next tile
    }
}

struct Ptr9 {
    const double* p[9];
};
struct Ptr6 {
    double* p[6];
};

// Inference at every grid point.  The stress is then cleared of local
// backscatter: with S the strain rate of the resolved velocity, a point where
// the subgrid dissipation -tau:S is negative gets tau + ((-tau:S)/(S:S)) S,
// whose dissipation is zero (the part of tau orthogonal to S is kept).
// mix > 0: tau = (1 - mix) tau_network + mix tau_Smagorinsky before the
// backscatter clearing, tau_Smagorinsky = -2 cd2 |S| S with cd2 = (c_s dx)^2
// and |S| = sqrt(2 S:S) (les_model 1's stress, m_les.f90:1142-1215).
__global__ __launch_bounds__(NT) void k_les(Geo g, Ptr9 gr, float delta, MlpNorm nrm,
                                            const uint16_t* __restrict__ wb, const float* __restrict__ p,
                                            Ptr6 tau, unsigned int* bad, float mix, float cd2) {
    __shared__ FwdLds s;
    __shared__ int64_t offs[TS];
    __shared__ float graw[TS][9];
    __shared__ float tl[TS][NOUT];
    const int64_t total = (int64_t)g.n * g.n * g.n;
    // grid point q -> physical slot offset; 32-bit division (n^3 <= 2^30)
    auto offset = [&](int64_t q) {
        const uint32_t qq = (uint32_t)q, nn = (uint32_t)g.n;
        return (int64_t)(qq / nn) * 2 * g.P + (int64_t)(qq % nn);
    };
    // Each thread gathers entries e = threadIdx.x + j*NT of a tile, feature
    // major (feature e / TS of sample e % TS): consecutive lanes read
    // consecutive grid points of one gradient slot.  The next tile's values
    // are loaded while this tile goes through the network.
    static_assert(TS * KIN == 2 * NT, "two gathered entries per thread");
    auto gather = [&](int64_t first, float (&raw)[2]) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = threadIdx.x + j * NT, c = e / TS;
            const int64_t q = first + e % TS;
            raw[j] = (c < 9 && q < total) ? (float)gr.p[c][offset(q)] : 0.f;
        }
    };
    float cur[2];
    gather((int64_t)blockIdx.x * FWD_TILES * TS, cur);
    const FwdRegs w = load_fwd_regs(wb, p);
    for (int it = 0; it < FWD_TILES; ++it) {
        const int64_t first = ((int64_t)blockIdx.x * FWD_TILES + it) * TS;
        if (first >= total) break;  // the same for the whole block
        if (threadIdx.x < TS) {
            const int64_t q = first + threadIdx.x;
            offs[threadIdx.x] = q < total ? offset(q) : -1;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = threadIdx.x + j * NT, c = e / TS, row = e % TS;
            float x = 0.f;
            if (first + row < total && c < NIN) {
                const float raw = c < 9 ? cur[j] : delta;
                if (c < 9) graw[row][c] = raw;
                x = (raw - nrm.xmean[c]) / nrm.xstd[c];
            }
            s.xb[row * KIN + c] = c < NIN ? to_bf16(x) : (uint16_t)0;
        }
        float nxt[2] = {0.f, 0.f};
        if (it + 1 < FWD_TILES && first + TS < total) gather(first + TS, nxt);
        __syncthreads();
        hidden_layers_reg(s, w);
        output_layer_reg(s, w, [&](int i, int n, float y) { tl[i][n] = y * nrm.ystd[n] + nrm.ymean[n]; });
        __syncthreads();
        if (threadIdx.x < TS) {
            const int i = threadIdx.x;
            const int64_t o = offs[i];
            if (o >= 0) {
                const float* gg = graw[i];  // du_i/dx_j at gg[3 i + j]
                const float S[6] = {gg[0], 0.5f * (gg[1] + gg[3]), 0.5f * (gg[2] + gg[6]), gg[4],
                                    0.5f * (gg[5] + gg[7]), gg[8]};
                const float w[6] = {1.f, 2.f, 2.f, 1.f, 2.f, 1.f};  // symmetric double sum
                float t[6], ts = 0.f, ss = 0.f;
                bool fin = true;
#pragma unroll
                for (int c = 0; c < NOUT; ++c) {
                    t[c] = tl[i][c];
                    fin = fin && isfinite(t[c]);
                    ss += w[c] * S[c] * S[c];
                }
                if (mix > 0.f) {
                    const float nut2 = -2.f * cd2 * sqrtf(2.f * ss);
#pragma unroll
                    for (int c = 0; c < NOUT; ++c) t[c] = (1.f - mix) * t[c] + mix * (nut2 * S[c]);
                }
#pragma unroll
                for (int c = 0; c < NOUT; ++c) ts += w[c] * t[c] * S[c];
                if (!fin) atomicAdd(bad, 1u);
                if (ts > 0.f && ss > 0.f) {  // -tau:S < 0: backscatter
                    const float a = -ts / ss;
#pragma unroll
                    for (int c = 0; c < NOUT; ++c) t[c] += a * S[c];
                }
#pragma unroll
                for (int c = 0; c < NOUT; ++c) tau.p[c][o] = (double)t[c];
            }
        }
        __syncthreads();  // offs, graw, tl, s.xb are rewritten by the next tile
        cur[0] = nxt[0];
        cur[1] = nxt[1];
    }
}

__global__ void k_features(Geo g, Ptr9 gr, float delta, float* X) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = (q / g.n) * 2 * g.P + q % g.n;
#pragma unroll
        for (int c = 0; c < 9; ++c) X[q * NIN + c] = (float)gr.p[c][o];
        X[q * NIN + 9] = delta;
    }
}

// one block per coarse (z, y): fine rows z in [r*zc, r*zc + r), y likewise;
// thread x accumulates its column, then r consecutive x are summed.
__global__ void k_box(Geo g, const double* u, const double* v, const double* w, int r, double* ubar,
                      double* tau) {
    extern __shared__ double sh[];  // [9][n]
    const int m = g.n / r;
    const int zc = blockIdx.y, yc = blockIdx.x;
    for (int x = threadIdx.x; x < g.n; x += blockDim.x) {
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int dz = 0; dz < r; ++dz)
            for (int dy = 0; dy < r; ++dy) {
                const int64_t o = ((int64_t)(zc * r + dz) * g.n + (yc * r + dy)) * 2 * g.P + x;
                const double uu = u[o], vv = v[o], ww = w[o];
                a[0] += uu;
                a[1] += vv;
                a[2] += ww;
                a[3] += uu * uu;
                a[4] += uu * vv;
                a[5] += uu * ww;
                a[6] += vv * vv;
                a[7] += vv * ww;
                a[8] += ww * ww;
            }
#pragma unroll
        for (int c = 0; c < 9; ++c) sh[c * g.n + x] = a[c];
    }
    __syncthreads();
    const double inv = 1.0 / ((double)r * r * r);
    const int64_t m3 = (int64_t)m * m * m;
    for (int xc = threadIdx.x; xc < m; xc += blockDim.x) {
        double a[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            double t = 0.0;
            for (int dx = 0; dx < r; ++dx) t += sh[c * g.n + xc * r + dx];
            a[c] = t * inv;
        }
        const int64_t q = ((int64_t)zc * m + yc) * m + xc;
        ubar[q] = a[0];
        ubar[m3 + q] = a[1];
        ubar[2 * m3 + q] = a[2];
        tau[q] = a[3] - a[0] * a[0];
        tau[m3 + q] = a[4] - a[0] * a[1];
        tau[2 * m3 + q] = a[5] - a[0] * a[2];
        tau[3 * m3 + q] = a[6] - a[1] * a[1];
        tau[4 * m3 + q] = a[7] - a[1] * a[2];
        tau[5 * m3 + q] = a[8] - a[2] * a[2];
    }
}

}  // namespace

void launch_mlp_train_tiles(const float* X, const float* T, int64_t count, int64_t start, int tiles,
                            int64_t nb_total, const uint16_t* wb, const float* p, float* partial,
                            float* lossp, hipStream_t s) {
    hipLaunchKernelGGL(k_train, dim3(tiles), dim3(NT), 0, s, X, T, count, start, nb_total, wb, p,
                       partial, lossp);
}
void launch_mlp_reduce(const float* partial, const float* lossp, int tiles, float* red,
                       float* grads, float* loss, hipStream_t s) {
    const unsigned nb = (NPARAM + 1 + 255) / 256;
    hipLaunchKernelGGL(k_reduce_groups, dim3(nb, kReduceGroups), dim3(256), 0, s, partial, lossp,
                       tiles, red);
    hipLaunchKernelGGL(k_reduce_final, dim3(nb), dim3(256), 0, s, red, grads, loss);
}
void launch_mlp_adam(float* p, uint16_t* wb, const float* g, float* m, float* v, int t, float lr,
                     hipStream_t s) {
    const float bc1 = 1.0f - powf(0.9f, (float)t), bc2 = 1.0f - powf(0.999f, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, p, wb, g, m, v, bc1, bc2,
                       lr);
}
void launch_mlp_to_bf16(const float* p, uint16_t* wb, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, s, p, wb);
}
void launch_mlp_forward(const float* X, int64_t count, const uint16_t* wb, const float* p, float* Y,
                        hipStream_t s) {
    const int64_t per = (int64_t)TS * FWD_TILES;
    hipLaunchKernelGGL(k_forward, dim3((unsigned)((count + per - 1) / per)), dim3(NT), 0, s, X,
                       count, wb, p, Y);
}
void launch_mlp_les(const Geo& g, const double* const grad[9], double delta, const MlpNorm& nrm,
                    const uint16_t* wb, const float* p, double* const tau[6], unsigned int* bad,
                    double mix, double cd2, hipStream_t s) {
    Ptr9 gr;
    for (int i = 0; i < 9; ++i) gr.p[i] = grad[i];
    Ptr6 tp;
    for (int i = 0; i < 6; ++i) tp.p[i] = tau[i];
    const int64_t total = (int64_t)g.n * g.n * g.n;
    if (total > ((int64_t)1 << 31)) throw std::runtime_error("closure inference: grid above 1024^3");
    const int64_t per = (int64_t)TS * FWD_TILES;
    hipLaunchKernelGGL(k_les, dim3((unsigned)((total + per - 1) / per)), dim3(NT), 0, s, g, gr,
                       (float)delta, nrm, wb, p, tp, bad, (float)mix, (float)cd2);
}
void launch_sgs_features(const Geo& g, const double* const grad[9], double delta, float* X,
                         hipStream_t s) {
    Ptr9 gr;
    for (int i = 0; i < 9; ++i) gr.p[i] = grad[i];
    const int64_t total = (int64_t)g.n * g.n * g.n;
    int64_t b = (total + 255) / 256;
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_features, dim3((unsigned)b), dim3(256), 0, s, g, gr, (float)delta, X);
}
void launch_box_filter(const Geo& g, const double* u, const double* v, const double* w, int r,
                       double* ubar, double* tau, hipStream_t s) {
    const int m = g.n / r;
    hipLaunchKernelGGL(k_box, dim3(m, m), dim3(256), (size_t)9 * g.n * sizeof(double), s, g, u, v,
                       w, r, ubar, tau);
}

}  // namespace hit3d
