// Axis passes of the 3-D real FFT (forward: x R2C, y C2C, z C2C with the
// (y,z) -> (kz,ky) transpose; inverse in reverse order), replacing the FFTW
// plane/stick loops and the slab transpose of xFFT3d (x_fftw.f90:852-1085).
#include <algorithm>
#include <vector>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fft.hpp"
#include "passes.hpp"

namespace hit3d {

bool supported_n(int n) {
    for (int p = 8; p <= 1024; p *= 2)
        if (n == p) return true;
    return false;
}

// ---------------------------------------------------------------------------
// column pass (y, and z with the transpose)
// ---------------------------------------------------------------------------
template <typename T>
struct ColCfg {
    // columns per block: 128-byte row segments in global memory
    static constexpr int B = 128 / (int)sizeof(cpx<T>);
};

template <int N, typename T>
struct ColE {
    static constexpr int B = ColCfg<T>::B;
    static constexpr int E = (N < 8) ? N : ((B * N / 8 <= 1024) ? 8 : 16);
};

// PH = +1: the input (spectral kx = column, kz = element, ky = group)
// is multiplied by exp(+i theta) as it is loaded; PH = -1: the output (same
// indexing) by exp(-i theta) before it is stored; theta = (kx+ky+kz) dx/2,
// tab[2(kx+ky+kz+N)] = cos, sin (the phase shift of rhs_velocity.f90:248-278).
template <int N, int PH, typename T>
__device__ __forceinline__ cpx<T> phase(cpx<T> a, int kx, int jy, int jz, const double* tab) {
    const int t = kx + wavenum(jy, N) + wavenum(jz, N) + N;
    const T c = (T)tab[2 * t], sn = (T)tab[2 * t + 1];
    if constexpr (PH > 0) return cpx<T>{a.x * c - a.y * sn, a.y * c + a.x * sn};
    else return cpx<T>{a.x * c + a.y * sn, a.y * c - a.x * sn};
}

// up to three slots of one launch, the slot chosen by blockIdx.z
struct Slots3 {
    const void* in[3];
    void* out[3];
};

template <int N, int E, int B, int DIR, typename T, int PH = 0, bool ACC = false>
__global__ __launch_bounds__(B* N / E) void k_cols(Slots3 io, int64_t a_in, int64_t a_out,
                                                   int64_t s_in, int64_t s_out, int ncols,
                                                   const cpx<T>* __restrict__ tw,
                                                   const double* __restrict__ tab, ColAcc acc) {
    constexpr int TPF = N / E;
    constexpr int PS = PadFor<T>::value;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    // select the slot without indexing the argument array at run time
    const void* ip = io.in[0];
    void* op = io.out[0];
    if (blockIdx.z == 1) ip = io.in[1], op = io.out[1];
    if (blockIdx.z == 2) ip = io.in[2], op = io.out[2];
    const cpx<T>* __restrict__ in = static_cast<const cpx<T>*>(ip);
    cpx<T>* out = static_cast<cpx<T>*>(op);
    const int c = threadIdx.x % B;
    const int t = threadIdx.x / B;
    const int col = blockIdx.x * B + c;
    const bool valid = col < ncols;
    const int64_t a = blockIdx.y;
    const cpx<T>* src = in + a * a_in + col;
    cpx<T>* dst = out + a * a_out + col;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s)
        v[s] = valid ? ld_g<kStream<N>>(&src[(int64_t)(t + s * TPF) * s_in]) : cpx<T>{(T)0, (T)0};
    // The phase factors in loops of their own: interleaved with the loads
    // or stores, each table load was issued at its use and waited for on
    // its own, one memory round trip per element.
    if constexpr (PH > 0)
        if (valid)
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = phase<N, 1>(v[s], col, (int)a, t + s * TPF, tab);
    block_fft<N, E, DIR, T, B, PS>(v, t, lds + c, tw + N);
    if (!valid) return;
    if constexpr (PH < 0)
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = phase<N, -1>(v[s], col, (int)a, t + s * TPF, tab);
    if constexpr (ACC) {
        // rhs += -1/2 i k_axis v for each target (rhs_velocity.f90:282-398).
        // Every target load is issued before the first store: the compiler
        // cannot tell the targets apart, and a store between two loads would
        // cost one memory round trip per load.  Blocks over 512 threads (at
        // most 128 registers) do it in two halves.
        constexpr int NCH = (B * N / E > 512) ? 2 : 1, H = E / NCH;
#pragma unroll
        for (int h = 0; h < NCH; ++h) {
            cpx<double> r[2][H];
#pragma unroll
            for (int q = 0; q < H; ++q) {
                const int64_t off = a * a_out + col + (int64_t)(t + (h * H + q) * TPF) * s_out;
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (acc.r[j]) r[j][q] = ld_g<kStream<N>>(&acc.r[j][off]);
            }
#pragma unroll
            for (int q = 0; q < H; ++q) {
                const int s = h * H + q;
                const int64_t off = a * a_out + col + (int64_t)(t + s * TPF) * s_out;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (!acc.r[j]) continue;
                    const int ax = acc.ax[j];
                    const double k = ax == 1 ? (double)col
                                             : (double)wavenum(ax == 2 ? (int)a : t + s * TPF, N);
                    cpx<double> val = {r[j][q].x + 0.5 * k * (double)v[s].y,
                                       r[j][q].y - 0.5 * k * (double)v[s].x};
                    if (acc.mask[j] &&
                        alias_count(col, wavenum((int)a, N), wavenum(t + s * TPF, N), N) >= 2)
                        val = {0.0, 0.0};
                    st_g<kStream<N>>(&acc.r[j][off], val);
                }
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < E; ++s) st_g<kStream<N>>(&dst[(int64_t)(t + s * TPF) * s_out], v[s]);
    }
}

// blocks of kernel k resident per compute unit at the given block size,
// queried once per kernel (the persistent column kernels share one function
// type, so the cache is keyed by the kernel's address)
template <typename K>
static int resident_per_cu(K k, int threads) {
    static std::vector<std::pair<const void*, int>> cache;
    for (const auto& e : cache)
        if (e.first == reinterpret_cast<const void*>(k)) return e.second;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, 0) != hipSuccess)
        throw std::runtime_error("column pass: occupancy query failed");
    per = std::max(1, per);
    cache.emplace_back(reinterpret_cast<const void*>(k), per);
    return per;
}

// compute units of the current device (queried once)
static int device_cus() {
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            throw std::runtime_error("column pass: device query failed");
        n = std::max(1, n);
    }
    return n;
}

// The plain and phase-shifted column passes as a persistent grid at
// 1024^3: a block's 128 KB of LDS leaves one block per compute unit, which
// otherwise holds no loads in flight while it transforms.  Each block loops
// over column groups (column block, group a, slot) with the next group's
// loads issued before the current group's transform.
template <int N, int E, int B, int DIR, typename T, int PH>
__global__ __launch_bounds__(B* N / E) void k_colsp(Slots3 io, int64_t a_in, int64_t a_out,
                                                    int64_t s_in, int64_t s_out, int ncols,
                                                    int ncb, int na, int ngroups,
                                                    const cpx<T>* __restrict__ tw,
                                                    const double* __restrict__ tab) {
    constexpr int TPF = N / E;
    constexpr int PS = 0;  // unpadded, as the fused step's column passes (registers)
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    const int c = threadIdx.x % B;
    const int t = threadIdx.x / B;
    auto at = [&](int q, int& col, int& a, int& z) {
        col = (q % ncb) * B + c;
        const int rest = q / ncb;
        a = rest % na;
        z = rest / na;
    };
    // entry indices inside a slot fit 32 bits (N^2 P < 2^31 up to 1024^3)
    const uint32_t si = (uint32_t)s_in, so = (uint32_t)s_out;
    auto load = [&](int q, cpx<T>(&v)[E]) {
        int col, a, z;
        at(q, col, a, z);
        const void* ip = io.in[0];
        if (z == 1) ip = io.in[1];
        if (z == 2) ip = io.in[2];
        const cpx<T>* src = static_cast<const cpx<T>*>(ip);
        const uint32_t idx = (uint32_t)a * (uint32_t)a_in + (uint32_t)col;
        const bool valid = col < ncols;
#pragma unroll
        for (int s = 0; s < E; ++s)
            v[s] = valid ? ld_g<kStream<N>>(&src[idx + (uint32_t)(t + s * TPF) * si])
                         : cpx<T>{(T)0, (T)0};
    };
    int q = blockIdx.x;
    if (q >= ngroups) return;
    cpx<T> nxt[E];
    load(q, nxt);
    for (; q < ngroups; q += gridDim.x) {
        cpx<T> v[E];
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = nxt[s];
        if (q + (int)gridDim.x < ngroups) load(q + gridDim.x, nxt);
        int col, a, z;
        at(q, col, a, z);
        const bool valid = col < ncols;
        if constexpr (PH > 0)
            if (valid)
#pragma unroll
                for (int s = 0; s < E; ++s) v[s] = phase<N, 1>(v[s], col, a, t + s * TPF, tab);
        // the twiddle pointer hidden from the optimiser: hoisted out of the
        // loop, the twiddles of all stages would stay in registers
        const cpx<T>* twp = tw + N;
        asm volatile("" : "+s"(twp));
        block_fft<N, E, DIR, T, B, PS, sizeof(T) == 8>(v, t, lds + c, twp);
        if (!valid) continue;
        if constexpr (PH < 0)
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = phase<N, -1>(v[s], col, a, t + s * TPF, tab);
        void* op = io.out[0];
        if (z == 1) op = io.out[1];
        if (z == 2) op = io.out[2];
        cpx<T>* dst = static_cast<cpx<T>*>(op);
        const uint32_t idx = (uint32_t)a * (uint32_t)a_out + (uint32_t)col;
#pragma unroll
        for (int s = 0; s < E; ++s) st_g<kStream<N>>(&dst[idx + (uint32_t)(t + s * TPF) * so], v[s]);
    }
}

template <int N, typename T>
static void cols_n(int dir, const cpx<T>* const* in, cpx<T>* const* out, int nb, int64_t a_in,
                   int64_t a_out, int64_t s_in, int64_t s_out, const cpx<T>* tw, hipStream_t s,
                   int ph, const double* tab, ColAcc acc) {
    constexpr int B = ColCfg<T>::B;
    constexpr int E = ColE<N, T>::E;
    const int ncols = N / 2 + 1;
    if (nb < 1 || nb > 3 || (nb > 1 && (ph != 0 || acc.r[0])))
        throw std::runtime_error("column pass: batches of 1..3 plain transforms only");
    Slots3 io{};
    for (int i = 0; i < nb; ++i) {
        io.in[i] = in[i];
        io.out[i] = out[i];
    }
    dim3 grid((ncols + B - 1) / B, N, nb);
    dim3 block(B * N / E);
    if (ph != 0 && !tab) throw std::runtime_error("column pass: phase table missing");
    if constexpr (N >= 1024 && std::is_same_v<T, double>) {
        if (!acc.r[0]) {
            const int ncb = (int)grid.x, ngroups = ncb * N * nb;
            auto go = [&](auto kern) {
                const int nblk = std::min(resident_per_cu(kern, (int)block.x) * device_cus(), ngroups);
                hipLaunchKernelGGL(kern, dim3(nblk), block, 0, s, io, a_in, a_out, s_in, s_out,
                                   ncols, ncb, N, ngroups, tw, tab);
            };
            if constexpr (std::is_same_v<T, double>) {
                if (dir > 0 && ph > 0) return go(k_colsp<N, E, B, 1, T, 1>);
                if (dir < 0 && ph < 0) return go(k_colsp<N, E, B, -1, T, -1>);
            }
            if (ph != 0) throw std::runtime_error("column pass: unsupported option");
            if (dir < 0) return go(k_colsp<N, E, B, -1, T, 0>);
            return go(k_colsp<N, E, B, 1, T, 0>);
        }
    }
    if constexpr (std::is_same_v<T, double>) {
        if (dir < 0 && ph < 0 && acc.r[0]) {
            hipLaunchKernelGGL((k_cols<N, E, B, -1, T, -1, true>), grid, block, 0, s, io, a_in,
                               a_out, s_in, s_out, ncols, tw, tab, acc);
            return;
        }
        if (dir < 0 && ph < 0) {
            hipLaunchKernelGGL((k_cols<N, E, B, -1, T, -1>), grid, block, 0, s, io, a_in,
                               a_out, s_in, s_out, ncols, tw, tab, acc);
            return;
        }
        if (dir > 0 && ph > 0) {
            hipLaunchKernelGGL((k_cols<N, E, B, 1, T, 1>), grid, block, 0, s, io, a_in, a_out,
                               s_in, s_out, ncols, tw, tab, acc);
            return;
        }
    }
    if (ph != 0 || acc.r[0]) throw std::runtime_error("column pass: unsupported option");
    if (dir < 0)
        hipLaunchKernelGGL((k_cols<N, E, B, -1, T>), grid, block, 0, s, io, a_in, a_out, s_in,
                           s_out, ncols, tw, tab, acc);
    else
        hipLaunchKernelGGL((k_cols<N, E, B, 1, T>), grid, block, 0, s, io, a_in, a_out, s_in,
                           s_out, ncols, tw, tab, acc);
}

// ---------------------------------------------------------------------------
// row pass (x): real <-> half-spectrum through an N/2-point complex FFT
// ---------------------------------------------------------------------------
template <int N, typename T>
struct RowCfg {
    static constexpr int M = N / 2;
    static constexpr int E = (M < 8) ? M : 8;
    static constexpr int TPF = M / E;
    static constexpr int ROWS = (256 / TPF < N * N) ? 256 / TPF : N * N;
};

template <int N, typename T>
__global__ __launch_bounds__((RowCfg<N, T>::ROWS * RowCfg<N, T>::TPF)) void k_r2c(
    Slots3 io, int64_t P, const cpx<T>* __restrict__ tw) {
    using C = RowCfg<N, T>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    __shared__ cpx<T> lds[C::ROWS][M + 1];
    void* dp = io.out[0];
    if (blockIdx.z == 1) dp = io.out[1];
    if (blockIdx.z == 2) dp = io.out[2];
    T* data = static_cast<T*>(dp);
    const int t = threadIdx.x % TPF;
    const int r = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + r;
    cpx<T>* rowp = reinterpret_cast<cpx<T>*>(data) + row * P;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = ld_g<kStream<N>>(&rowp[t + s * TPF]);
    block_fft<M, E, -1, T, 1, kSwizzle>(v, t, lds[r], tw + M);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) lds[r][t + s * TPF] = v[s];
    __syncthreads();
    const cpx<T>* twN = tw + N;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int k = t + s * TPF;
        cpx<T> zk = lds[r][k];
        cpx<T> zm = conj(lds[r][(M - k) & (M - 1)]);
        cpx<T> e = scale(zk + zm, (T)0.5);
        cpx<T> o = scale(mul_mi(zk - zm), (T)0.5);
        st_g<kStream<N>>(&rowp[k], e + cmul(twN[k], o));
    }
    if (t == 0) {
        cpx<T> z0 = lds[r][0];
        st_g<kStream<N>>(&rowp[M], cpx<T>{z0.x - z0.y, (T)0});
    }
}

// forward x pass of the pointwise product of two physical slots a*b into out
// (out may be a or b: each row is read before it is written)
template <int N, typename T>
__global__ __launch_bounds__((RowCfg<N, T>::ROWS * RowCfg<N, T>::TPF)) void k_r2c_mul(
    const T* a, const T* b, T* out, int64_t P, const cpx<T>* __restrict__ tw) {
    using C = RowCfg<N, T>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    __shared__ cpx<T> lds[C::ROWS][M + 1];
    const int t = threadIdx.x % TPF;
    const int r = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + r;
    const cpx<T>* ap = reinterpret_cast<const cpx<T>*>(a) + row * P;
    const cpx<T>* bp = reinterpret_cast<const cpx<T>*>(b) + row * P;
    cpx<T>* rowp = reinterpret_cast<cpx<T>*>(out) + row * P;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const cpx<T> x = ld_g<kStream<N>>(&ap[t + s * TPF]), y = ld_g<kStream<N>>(&bp[t + s * TPF]);
        v[s] = cpx<T>{x.x * y.x, x.y * y.y};
    }
    block_fft<M, E, -1, T, 1, kSwizzle>(v, t, lds[r], tw + M);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) lds[r][t + s * TPF] = v[s];
    __syncthreads();
    const cpx<T>* twN = tw + N;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int k = t + s * TPF;
        cpx<T> zk = lds[r][k];
        cpx<T> zm = conj(lds[r][(M - k) & (M - 1)]);
        cpx<T> e = scale(zk + zm, (T)0.5);
        cpx<T> o = scale(mul_mi(zk - zm), (T)0.5);
        st_g<kStream<N>>(&rowp[k], e + cmul(twN[k], o));
    }
    if (t == 0) {
        cpx<T> z0 = lds[r][0];
        st_g<kStream<N>>(&rowp[M], cpx<T>{z0.x - z0.y, (T)0});
    }
}

template <int N, typename T>
__global__ __launch_bounds__((RowCfg<N, T>::ROWS * RowCfg<N, T>::TPF)) void k_c2r(
    Slots3 io, int64_t P, const cpx<T>* __restrict__ tw, T norm) {
    using C = RowCfg<N, T>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    __shared__ cpx<T> lds[C::ROWS][M + 1];
    void* dp = io.out[0];
    if (blockIdx.z == 1) dp = io.out[1];
    if (blockIdx.z == 2) dp = io.out[2];
    T* data = static_cast<T*>(dp);
    const int t = threadIdx.x % TPF;
    const int r = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + r;
    cpx<T>* rowp = reinterpret_cast<cpx<T>*>(data) + row * P;
#pragma unroll
    for (int s = 0; s < E; ++s) lds[r][t + s * TPF] = ld_g<kStream<N>>(&rowp[t + s * TPF]);
    if (t == 0) lds[r][M] = ld_g<kStream<N>>(&rowp[M]);
    __syncthreads();
    const cpx<T>* twN = tw + N;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int k = t + s * TPF;
        cpx<T> xk = lds[r][k];
        cpx<T> xm = lds[r][M - k];
        if (k == 0) {  // c2r ignores Im of the kx=0 and kx=N/2 terms
            xk.y = (T)0;
            xm.y = (T)0;
        }
        xm = conj(xm);
        // Z_k = (X_k + conj X_{M-k}) + i W_N^{-k} (X_k - conj X_{M-k})
        v[s] = (xk + xm) + mul_i(cmulc(xk - xm, twN[k]));
    }
    block_fft<M, E, 1, T, 1, kSwizzle>(v, t, lds[r], tw + M);
#pragma unroll
    for (int s = 0; s < E; ++s) st_g<kStream<N>>(&rowp[t + s * TPF], scale(v[s], norm));
}

static Slots3 batch_of(void* const* slots, int nb) {
    if (nb < 1 || nb > 3) throw std::runtime_error("row pass: batches of 1..3 slots only");
    Slots3 io{};
    for (int i = 0; i < nb; ++i) io.out[i] = slots[i];
    return io;
}

template <int N, typename T>
static void r2c_n(T* const* slots, int nb, int64_t P, const cpx<T>* tw, hipStream_t s) {
    using C = RowCfg<N, T>;
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL((k_r2c<N, T>), dim3(rows / C::ROWS, 1, nb), dim3(C::ROWS * C::TPF), 0, s,
                       batch_of(reinterpret_cast<void* const*>(slots), nb), P, tw);
}

template <int N, typename T>
static void r2c_mul_n(const T* a, const T* b, T* out, int64_t P, const cpx<T>* tw, hipStream_t s) {
    using C = RowCfg<N, T>;
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL((k_r2c_mul<N, T>), dim3(rows / C::ROWS), dim3(C::ROWS * C::TPF), 0, s, a, b,
                       out, P, tw);
}

template <int N, typename T>
static void c2r_n(T* const* slots, int nb, int64_t P, const cpx<T>* tw, T norm, hipStream_t s) {
    using C = RowCfg<N, T>;
    const int64_t rows = (int64_t)N * N;
    hipLaunchKernelGGL((k_c2r<N, T>), dim3(rows / C::ROWS, 1, nb), dim3(C::ROWS * C::TPF), 0, s,
                       batch_of(reinterpret_cast<void* const*>(slots), nb), P, tw, norm);
}

// ---------------------------------------------------------------------------
// dispatch over the supported sizes
// ---------------------------------------------------------------------------
template <typename F>
static void dispatch(int n, F&& f) {
    switch (n) {
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        case 128: f(std::integral_constant<int, 128>{}); break;
        case 256: f(std::integral_constant<int, 256>{}); break;
        case 512: f(std::integral_constant<int, 512>{}); break;
        case 1024: f(std::integral_constant<int, 1024>{}); break;
        default: throw std::runtime_error("unsupported grid size " + std::to_string(n));
    }
}

template <typename T>
void launch_x_r2c(int n, T* slot, int64_t P, const cpx<T>* tw, hipStream_t s) {
    dispatch(n, [&](auto NC) { r2c_n<decltype(NC)::value, T>(&slot, 1, P, tw, s); });
}
template <typename T>
void launch_x_c2r(int n, T* slot, int64_t P, const cpx<T>* tw, T norm, hipStream_t s) {
    dispatch(n, [&](auto NC) { c2r_n<decltype(NC)::value, T>(&slot, 1, P, tw, norm, s); });
}
void launch_x_r2c_batch(int n, double* const* slots, int nb, int64_t P, const cpx<double>* tw,
                        hipStream_t s) {
    dispatch(n, [&](auto NC) { r2c_n<decltype(NC)::value, double>(slots, nb, P, tw, s); });
}
void launch_x_c2r_batch(int n, double* const* slots, int nb, int64_t P, const cpx<double>* tw,
                        double norm, hipStream_t s) {
    dispatch(n, [&](auto NC) { c2r_n<decltype(NC)::value, double>(slots, nb, P, tw, norm, s); });
}
template <typename T>
void launch_x_r2c_mul(int n, const T* a, const T* b, T* out, int64_t P, const cpx<T>* tw,
                      hipStream_t s) {
    dispatch(n, [&](auto NC) { r2c_mul_n<decltype(NC)::value, T>(a, b, out, P, tw, s); });
}
template <typename T>
void launch_cols(int n, int dir, const cpx<T>* in, cpx<T>* out, int64_t a_in, int64_t a_out,
                 int64_t s_in, int64_t s_out, const cpx<T>* tw, hipStream_t s, int ph,
                 const double* tab, ColAcc acc) {
    dispatch(n, [&](auto NC) {
        cols_n<decltype(NC)::value, T>(dir, &in, &out, 1, a_in, a_out, s_in, s_out, tw, s, ph, tab,
                                       acc);
    });
}
void launch_cols_batch(int n, int dir, const cpx<double>* const* in, cpx<double>* const* out,
                       int nb, int64_t a_in, int64_t a_out, int64_t s_in, int64_t s_out,
                       const cpx<double>* tw, hipStream_t s) {
    dispatch(n, [&](auto NC) {
        cols_n<decltype(NC)::value, double>(dir, in, out, nb, a_in, a_out, s_in, s_out, tw, s, 0,
                                            nullptr, ColAcc{});
    });
}

template void launch_x_r2c_mul<double>(int, const double*, const double*, double*, int64_t,
                                       const cpx<double>*, hipStream_t);
template void launch_x_r2c<double>(int, double*, int64_t, const cpx<double>*, hipStream_t);
template void launch_x_r2c<float>(int, float*, int64_t, const cpx<float>*, hipStream_t);
template void launch_x_c2r<double>(int, double*, int64_t, const cpx<double>*, double, hipStream_t);
template void launch_x_c2r<float>(int, float*, int64_t, const cpx<float>*, float, hipStream_t);
template void launch_cols<double>(int, int, const cpx<double>*, cpx<double>*, int64_t, int64_t,
                                  int64_t, int64_t, const cpx<double>*, hipStream_t, int,
                                  const double*, ColAcc);
template void launch_cols<float>(int, int, const cpx<float>*, cpx<float>*, int64_t, int64_t,
                                 int64_t, int64_t, const cpx<float>*, hipStream_t, int,
                                 const double*, ColAcc);

}  // namespace hit3d
