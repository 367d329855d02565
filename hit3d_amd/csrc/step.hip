// Fused time step for the 2/3-rule velocity equation (main.f90:124-167 with
// rhs_scalars.f90:32-39, rhs_velocity.f90:31-144, m_force.f90:101-182,
// main.f90:156-164, pressure.f90:13-47).
//
// The prognostic fields are zero outside the cube |k_i| <= kmax (dealias_all
// at start-up, a masked RHS every step), so every transform is pruned to
// that cube: the inverse transforms read only the kmax+1 kx columns and the
// 2kmax+1 ky/kz rows that can be non-zero, and the forward transforms of the
// six products keep only those outputs.  Five kernels, each one pass over
// HBM:
//   A  inverse z transform of u,v,w, reading fields        (kx, kz, ky) -> (kx, ky, z)
//   B  inverse y transform, in place                        (kx, ky, z)  -> (kx, y, z)
//   C  x: c2r of u,v,w, Courant max, six products u_i u_j, r2c of each product
//   D  forward y transform of the products, in place, keeping kept ky rows
//   E  forward z transform of the products + RHS assembly -i k_j (u_i u_j)^
//      - nu k^2 u + band forcing + Euler/AB2 update + projection, writing
//      fields and rhs_old.
// At 1024^3, C runs as k_xh (each real row one half-length transform inside
// a wavefront) and A, B, D as persistent grids (k_inv_zp, k_yp) that load
// their next column group while transforming the current one.
// The work buffers (wrk slots) use a pruned row pitch PA >= kmax+1 and T =
// double (FP64 form) or float (mixed form: FP32 transforms and nonlinear
// term, FP64 state, update, forcing energy and Courant reduction; rhs_old
// stored in FP32, cpx<float> at the FP64 entry index, which saves a sixth
// of kernel E's traffic and costs nothing the FP32 RHS had not already
// lost).
#include "fft.hpp"
#include "passes.hpp"
#include "state.hpp"

namespace hit3d {

namespace {

struct SG {
    int K;        // kmax
    int NK;       // kept kx columns = K + 1
    int64_t P;    // field row pitch (complex)
    int64_t PA;   // work row pitch (complex)
};

template <typename T>
struct BCols {
    static constexpr int B = 128 / (int)sizeof(cpx<T>);
};

template <int N, typename T>
struct CE {
    static constexpr int B = BCols<T>::B;
    static constexpr int E = (B * N / 8 <= 1024) ? 8 : 16;
    static constexpr int TPF = N / E;
    static constexpr int THREADS = B * TPF;
};

// y passes: FP32 blocks of 16 columns at N >= 512 take radix 16 (half the
// threads, one LDS exchange fewer): 0.377 -> 0.350 ms inverse and 0.735 ->
// 0.704 ms forward at 512^3; the z kernels measured slower that way
template <int N, typename T>
struct CY {
    static constexpr int B = CE<N, T>::B;
    static constexpr int E = (sizeof(T) == 4 && B * N / 8 > 512) ? 16 : CE<N, T>::E;
    static constexpr int TPF = N / E;
    static constexpr int THREADS = B * TPF;
};

// forward z + update at N >= 1024, in 512-thread blocks: the column blocks
// of the other passes are 1024 threads there, with 128 registers each, and
// spilled 412 (FP64) / 452 (FP32) bytes per lane.  FP64: 4 columns, radix 8
// (20.1 -> 17.3-20.5 ms at 1024^3); FP32: 8 columns, radix 16 (13.8 ->
// 9.6 ms)
template <int N, typename T>
struct CZ {
    static constexpr int B = N >= 1024 ? (sizeof(T) == 8 ? 4 : 8) : CE<N, T>::B;
    static constexpr int E = (N >= 1024 && sizeof(T) == 4) ? 16 : CE<N, T>::E;
    static constexpr int TPF = N / E;
    static constexpr int THREADS = B * TPF;
};

// x pass as full-length complex transforms of two real rows at once:
// z = a + i b, TPF threads per row: inside one wavefront up to N = 512, two
// wavefronts at N = 1024 (radix 16 in one wavefront needed 256 VGPRs + 81
// AGPRs, one wavefront per SIMD: 28.7 ms at 1024^3)
template <int N, typename T>
struct XP {
    static constexpr int E = N >= 8 ? 8 : N;
    static constexpr int TPF = N / E;
    static constexpr int ROWS = (256 / TPF < N * N) ? 256 / TPF : N * N;
    static constexpr int THREADS = ROWS * TPF;
};

// The fused step runs the 2/3 rule only, where kmax = N/3 (m_parameters.f90:124).
template <int N>
struct Kc {
    static constexpr int K = N / 3;
    static constexpr int NK = K + 1;
};

__device__ __forceinline__ bool kept(int j, int K, int n) { return j <= K || j >= n - K; }

// Column blocks narrower than a 128-byte row segment split each segment
// between two blocks.  Blocks are dispatched round-robin over the 8 XCDs,
// each with its own L2, so two neighbouring blocks would each fetch the
// whole segment from HBM (the 1024^3 z/update kernel moved 1.6x its
// algorithmic bytes).  With PAIR the two halves go to blocks 8 apart, which
// run on one XCD at about the same time, and the second finds the segment
// in that L2.  The grid must have an even number of column blocks; returns
// this block's logical (column block, row).
// The paired blocks load through the cache (not streaming): a streaming
// load's line is gone from L2 before the partner block reads its other half,
// and the 1024^3 z/update kernel fetched 1.30x its algorithmic bytes that
// way (cached: 0.98x, 12.3-14.5 -> 11.4 ms).
template <typename T, int B>
inline constexpr bool kPairHalves = B * (int)sizeof(cpx<T>) < 128;
template <int N, typename T, int B>
inline constexpr bool kStreamLoad = kStream<N> && !kPairHalves<T, B>;
template <bool PAIR>
__device__ __forceinline__ void block_coords(int& bx, int& by) {
    if constexpr (!PAIR) {
        bx = blockIdx.x;
        by = blockIdx.y;
    } else {
        const int gx = gridDim.x, total = gx * gridDim.y;
        const int l = blockIdx.x + gx * blockIdx.y;
        int logical = l;
        if (l < total - total % 16) {
            const int grp = l / 16, r = l % 16;
            logical = (grp * 8 + (r & 7)) * 2 + (r >> 3);
        }
        bx = logical % gx;
        by = logical / gx;
    }
}
// Whether any element t + s*TPF (0 <= t < TPF) of register slot s is kept;
// constant once the slot loop is unrolled, so dead slots take no registers.
template <int N, int TPF>
__device__ __forceinline__ constexpr bool slot_live(int s) {
    return s * TPF <= Kc<N>::K || s * TPF + TPF - 1 >= N - Kc<N>::K;
}
__device__ __forceinline__ int kept_row(int kp, int K, int n) { return kp <= K ? kp : n - 2 * K - 1 + kp; }

template <typename T>
__device__ __forceinline__ cpx<T> to(cpx<double> a) {
    return cpx<T>{(T)a.x, (T)a.y};
}

struct P3 {
    const double* p[3];
};
template <typename T>
struct W6 {
    T* p[6];
};

// ---------------------------------------------------------------- A
template <int N, typename T>
__global__ __launch_bounds__((CE<N, T>::THREADS)) void k_inv_z(SG g, P3 F, W6<T> A,
                                                             const cpx<T>* __restrict__ tw) {
    constexpr int E = CE<N, T>::E, B = CE<N, T>::B, TPF = CE<N, T>::TPF;
    constexpr int PS = 0;  // padding does not pay in the column passes
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = blockIdx.x * B + c;
    const bool valid = kx < Kc<N>::NK;
    const int jy = kept_row(blockIdx.y, Kc<N>::K, N);
    const int comp = blockIdx.z;
    const cpx<double>* src = reinterpret_cast<const cpx<double>*>(F.p[comp]) + (int64_t)jy * N * g.P + kx;
    cpx<T>* dst = reinterpret_cast<cpx<T>*>(A.p[comp]) + (int64_t)jy * g.PA + kx;
    cpx<T> v[E];
    // Branch-free loads (an in-bounds address for dead lanes, conversion
    // after the load): as a branch, the FP32 form's conversion waited for
    // every load in flight, one round trip per element.  Slots with no kept
    // row in any lane (constant once unrolled) load nothing.
    const cpx<double>* base = reinterpret_cast<const cpx<double>*>(F.p[comp]);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        if (!slot_live<N, TPF>(s)) {
            v[s] = cpx<T>{(T)0, (T)0};
            continue;
        }
        const int e = t + s * TPF;
        const bool ok = valid && kept(e, Kc<N>::K, N);
        const cpx<double> x = ld_g<kStream<N>>(ok ? &src[(int64_t)e * g.P] : base);
        v[s] = ok ? to<T>(x) : cpx<T>{(T)0, (T)0};
    }
    block_fft<N, E, 1, T, B, PS>(v, t, lds + c, tw + N);
    if (valid) {
#pragma unroll
        for (int s = 0; s < E; ++s) st_g<kStream<N>>(&dst[(int64_t)(t + s * TPF) * N * g.PA], v[s]);
    }
}

// A as a persistent grid at 1024^3 in FP64 (see k_yp): groups (column
// block, kept ky row, component), the next group's loads issued before the
// current group's transform: 4.72 -> 4.28 ms.  The FP32 form keeps k_inv_z
// (persistent: 3.21 -> 3.87 ms, its 16-entry columns need the registers).
template <int N, typename T>
__global__ __launch_bounds__((CE<N, T>::THREADS)) void k_inv_zp(SG g, P3 F, W6<T> A,
                                                              const cpx<T>* __restrict__ tw,
                                                              int ncb, int ngroups) {
    constexpr int E = CE<N, T>::E, B = CE<N, T>::B, TPF = CE<N, T>::TPF;
    constexpr int PS = 0;
    constexpr int NKY = 2 * Kc<N>::K + 1;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    // entry indices fit 32 bits: at most N^2 P < 2^31 in a field slot
    auto at = [&](int q, int& kx, int& jy, int& comp) {
        const int bx = q % ncb, rest = q / ncb;
        kx = bx * B + c;
        jy = kept_row(rest % NKY, Kc<N>::K, N);
        comp = rest / NKY;
    };
    const uint32_t P = (uint32_t)g.P;
    auto load = [&](int q, cpx<T>(&v)[E]) {
        int kx, jy, comp;
        at(q, kx, jy, comp);
        const bool valid = kx < Kc<N>::NK;
        const cpx<double>* base = reinterpret_cast<const cpx<double>*>(F.p[0]);
        if (comp == 1) base = reinterpret_cast<const cpx<double>*>(F.p[1]);
        if (comp == 2) base = reinterpret_cast<const cpx<double>*>(F.p[2]);
        const uint32_t idx = (uint32_t)jy * (uint32_t)N * P + (uint32_t)kx;
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (!slot_live<N, TPF>(s)) {
                v[s] = cpx<T>{(T)0, (T)0};
                continue;
            }
            const uint32_t e = t + s * TPF;
            const bool ok = valid && kept(e, Kc<N>::K, N);
            const cpx<double> x = ld_g<kStream<N>>(&base[ok ? idx + e * P : 0u]);
            v[s] = ok ? to<T>(x) : cpx<T>{(T)0, (T)0};
        }
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
        const cpx<T>* twp = tw + N;
        asm volatile("" : "+s"(twp));  // not hoisted out of the loop (registers)
        block_fft<N, E, 1, T, B, PS, sizeof(T) == 8>(v, t, lds + c, twp);
        int kx, jy, comp;
        at(q, kx, jy, comp);
        if (kx < Kc<N>::NK) {
            T* ap = A.p[0];
            if (comp == 1) ap = A.p[1];
            if (comp == 2) ap = A.p[2];
            cpx<T>* dst = reinterpret_cast<cpx<T>*>(ap);
            const uint32_t idx = (uint32_t)jy * (uint32_t)g.PA + (uint32_t)kx;
            const uint32_t zs = (uint32_t)N * (uint32_t)g.PA;
#pragma unroll
            for (int s = 0; s < E; ++s) st_g<kStream<N>>(&dst[idx + (uint32_t)(t + s * TPF) * zs], v[s]);
        }
    }
}

// A for the rotational form: each block loads the three velocity columns
// (kx, ky) once and transforms six: u, v, w and omega = i k x u, formed in
// FP64 from the loaded entries (omega_c = i (k_a u_b - k_b u_a), (a, b) =
// (c+1, c+2) mod 3).  Loading each column once instead of once per output
// that needs it reads a third of the cube traffic of six independent
// blocks.
// Blocks of at most 512 threads (half the columns where the other column
// passes use 1024): the three loaded columns stay in registers across the
// six transforms, which 1024-thread blocks (128 VGPRs) would spill.
template <int N, typename T>
struct CR {
    static constexpr int E = CE<N, T>::E;
    static constexpr int TPF = N / E;
    static constexpr int B = CE<N, T>::B * TPF > 512 ? CE<N, T>::B / 2 : CE<N, T>::B;
    static constexpr int THREADS = B * TPF;
};

template <int N, typename T>
__global__ __launch_bounds__((CR<N, T>::THREADS)) void k_inv_z_rot(SG g, P3 F, W6<T> A,
                                                                 const cpx<T>* __restrict__ tw) {
    constexpr int E = CR<N, T>::E, B = CR<N, T>::B, TPF = CR<N, T>::TPF;
    constexpr int PS = 0;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    int bx, by;
    block_coords<kPairHalves<T, B>>(bx, by);
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = bx * B + c;
    const bool valid = kx < Kc<N>::NK;
    const int jy = kept_row(by, Kc<N>::K, N);
    const int kyv = wavenum(jy, N);
    // the velocity in the work precision (the mixed form's vorticity is
    // formed from the FP32-rounded velocity, in FP64, then rounded)
    cpx<T> uv[3][E];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
        const cpx<double>* base = reinterpret_cast<const cpx<double>*>(F.p[cc]);
        const cpx<double>* src = base + (int64_t)jy * N * g.P + kx;
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (!slot_live<N, TPF>(s)) {
                uv[cc][s] = cpx<T>{(T)0, (T)0};
                continue;
            }
            const int e = t + s * TPF;
            const bool ok = valid && kept(e, Kc<N>::K, N);
            const cpx<double> x = ld_g<kStreamLoad<N, T, B>>(ok ? &src[(int64_t)e * g.P] : base);
            uv[cc][s] = ok ? to<T>(x) : cpx<T>{(T)0, (T)0};
        }
    }
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        cpx<T> v[E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (comp < 3) {
                v[s] = comp == 0 ? uv[0][s] : comp == 1 ? uv[1][s] : uv[2][s];
            } else {
                const int c3 = comp - 3;
                const int kzv = wavenum(t + s * TPF, N);
                const double ka = c3 == 0 ? (double)kyv : c3 == 1 ? (double)kzv : (double)kx;
                const double kb = c3 == 0 ? (double)kzv : c3 == 1 ? (double)kx : (double)kyv;
                const cpx<T> ub = c3 == 0 ? uv[2][s] : c3 == 1 ? uv[0][s] : uv[1][s];
                const cpx<T> ua = c3 == 0 ? uv[1][s] : c3 == 1 ? uv[2][s] : uv[0][s];
                const cpx<double> d{ka * (double)ub.x - kb * (double)ua.x,
                                    ka * (double)ub.y - kb * (double)ua.y};
                v[s] = to<T>(mul_i(d));
            }
        }
        block_fft<N, E, 1, T, B, PS>(v, t, lds + c, tw + N);
        T* ap = A.p[0];
        if (comp == 1) ap = A.p[1];
        if (comp == 2) ap = A.p[2];
        if (comp == 3) ap = A.p[3];
        if (comp == 4) ap = A.p[4];
        if (comp == 5) ap = A.p[5];
        cpx<T>* dst = reinterpret_cast<cpx<T>*>(ap) + (int64_t)jy * g.PA + kx;
        if (valid) {
#pragma unroll
            for (int s = 0; s < E; ++s) st_g<kStream<N>>(&dst[(int64_t)(t + s * TPF) * N * g.PA], v[s]);
        }
    }
}

// ---------------------------------------------------------------- B / D
// DIR=+1: inverse y, loads kept rows only, stores all rows.
// DIR=-1: forward y, loads all rows, stores kept rows only.
template <int N, int DIR, typename T>
__global__ __launch_bounds__((CY<N, T>::THREADS)) void k_y(SG g, W6<T> A,
                                                         const cpx<T>* __restrict__ tw) {
    constexpr int E = CY<N, T>::E, B = CY<N, T>::B, TPF = CY<N, T>::TPF;
    constexpr int PS = 0;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = blockIdx.x * B + c;
    const bool valid = kx < Kc<N>::NK;
    const int z = blockIdx.y;
    cpx<T>* base = reinterpret_cast<cpx<T>*>(A.p[blockIdx.z]) + (int64_t)z * N * g.PA + kx;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int e = t + s * TPF;
        const bool ld = valid && (DIR < 0 || kept(e, Kc<N>::K, N));
        v[s] = ld ? ld_g<kStream<N>>(&base[(int64_t)e * g.PA]) : cpx<T>{(T)0, (T)0};
    }
    block_fft<N, E, DIR, T, B, PS>(v, t, lds + c, tw + N);
    if (valid) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int e = t + s * TPF;
            if (DIR > 0 || kept(e, Kc<N>::K, N)) st_g<kStream<N>>(&base[(int64_t)e * g.PA], v[s]);
        }
    }
}

// The same pass as a persistent grid (1024^3): as many blocks as fit on the
// device at once, each looping over column groups (column block, plane z,
// buffer) with the loads of its next group issued before the transform of
// the current one.  At 1024^3 a block's 128 KB of LDS leaves one block per
// compute unit, which with k_y held no loads in flight while it transformed:
// 6.46 -> 6.06 ms inverse, 10.66 -> 9.83 ms forward (FP64).  At 512^3, two
// blocks per unit already overlap, and the loop's registers cost more than
// the prefetch gains (0.67 -> 0.71 ms, 1.10 -> 1.15 ms).
template <int N, int DIR, typename T>
__global__ __launch_bounds__((CY<N, T>::THREADS)) void k_yp(SG g, W6<T> A,
                                                          const cpx<T>* __restrict__ tw, int ncb,
                                                          int ngroups) {
    constexpr int E = CY<N, T>::E, B = CY<N, T>::B, TPF = CY<N, T>::TPF;
    constexpr int PS = 0;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    // group q: column block q % ncb, plane z, buffer; the buffer's base
    // pointer is uniform, the entry index inside it fits 32 bits (at most
    // N^2 PA < 2^31 entries), which keeps the per-slot offsets small
    auto at = [&](int q, bool& valid, uint32_t& idx) {
        const int bx = q % ncb, rest = q / ncb;
        const int z = rest % N, buf = rest / N;
        const int kx = bx * B + c;
        valid = kx < Kc<N>::NK;
        T* p = A.p[0];
        if (buf == 1) p = A.p[1];
        if (buf == 2) p = A.p[2];
        if (buf == 3) p = A.p[3];
        if (buf == 4) p = A.p[4];
        if (buf == 5) p = A.p[5];
        idx = (uint32_t)z * (uint32_t)(N * g.PA) + (uint32_t)kx;
        return reinterpret_cast<cpx<T>*>(p);
    };
    const uint32_t pa = (uint32_t)g.PA;
    auto load = [&](int q, cpx<T>(&v)[E]) {
        bool valid;
        uint32_t idx;
        const cpx<T>* base = at(q, valid, idx);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const uint32_t e = t + s * TPF;
            const bool ld = valid && (DIR < 0 || kept(e, Kc<N>::K, N));
            v[s] = ld ? ld_g<kStream<N>>(&base[idx + e * pa]) : cpx<T>{(T)0, (T)0};
        }
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
        // the twiddle pointer hidden from the optimiser: hoisted out of the
        // loop, the twiddles of all stages would stay in registers; FP64
        // derives them by recurrence (one table load per butterfly)
        const cpx<T>* twp = tw + N;
        asm volatile("" : "+s"(twp));
        block_fft<N, E, DIR, T, B, PS, sizeof(T) == 8>(v, t, lds + c, twp);
        bool valid;
        uint32_t idx;
        cpx<T>* base = at(q, valid, idx);
        if (valid) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const uint32_t e = t + s * TPF;
                if (DIR > 0 || kept(e, Kc<N>::K, N)) st_g<kStream<N>>(&base[idx + e * pa], v[s]);
            }
        }
    }
}

// blocks of kernel k that fit on the device at once (at least one per unit)
template <typename K>
static int resident_blocks(K k, int threads, int ncu) {
    int per = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, 0),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return std::max(1, per) * std::max(1, ncu);
}

template <int N, int DIR, typename T>
static void launch_y(Lib& L, const SG& g, const W6<T>& Wb, const cpx<T>* tw, int nbuf) {
    constexpr int B = CY<N, T>::B;
    const int ncb = (g.NK + B - 1) / B;
    if constexpr (N >= 1024) {
        const int ngroups = ncb * N * nbuf;
        static int grid = 0;
        if (grid == 0) grid = resident_blocks(k_yp<N, DIR, T>, CY<N, T>::THREADS, L.ncu);
        hipLaunchKernelGGL((k_yp<N, DIR, T>), dim3(std::min(grid, ngroups)),
                           dim3(CY<N, T>::THREADS), 0, L.st, g, Wb, tw, ncb, ngroups);
    } else {
        hipLaunchKernelGGL((k_y<N, DIR, T>), dim3(ncb, N, nbuf), dim3(CY<N, T>::THREADS), 0, L.st,
                           g, Wb, tw);
    }
}

// ---------------------------------------------------------------- C (pairs)
// The x pass with two real rows per complex transform of length N.
// Inverse: Z_k = U_k + i V_k with U_{N-k} = conj U_k (Im of kx=0 ignored,
// as the reference's c2r does, x_fftw.f90:1066-1073); z_n = u_n + i v_n.
// w rides alone as Z_k = W_k.  Forward: z = p + i q for the product pairs
// (uu,uv), (uw,vv), (vw,ww); P_k = (Z_k + conj Z_{N-k})/2,
// Q_k = -i (Z_k - conj Z_{N-k})/2, kept for kx <= kmax.  Output k sits in
// lane t, slot s (k = t + s*TPF); N-k sits in lane (TPF-t) mod TPF, slot
// E-1-s (slot (E-s) mod E for t = 0), so the split is a lane shuffle.  One
// LDS row of N entries per row for the transform exchanges, synchronised per
// wavefront; twiddles by recurrence from one FP64 table load per butterfly
// (computed in FP64 for the FP32 form as well, then rounded once).
template <int N, int TPF, int NK, typename T>
__device__ __forceinline__ cpx<T> herm_load(const cpx<T>* rowp, int t, int s) {
    // entry k = t + s*TPF of the Hermitian-extended half spectrum (zero
    // outside kx < NK).  Cached loads: the row is read twice (k and N-k), and
    // streaming loads made the pass slower (1.60 -> 1.72 ms at 512^3 FP64).
    // Branch-free: every live slot loads from an in-bounds address and the
    // selection and conjugation follow the load; as branches, each
    // conjugation waited for all loads in flight, one round trip per slot.
    // A slot with no live entry in any lane (constant once unrolled) loads
    // nothing.
    if (s * TPF >= NK && s * TPF + TPF - 1 <= N - NK) return cpx<T>{(T)0, (T)0};
    const int k = t + s * TPF;
    const bool mirror = 2 * k > N;
    const int kk = mirror ? N - k : k;
    const bool live = kk < NK;
    const cpx<T> a = rowp[live ? kk : 0];
    const T re = live ? a.x : (T)0;
    const T im = !live ? (T)0 : mirror ? -a.y : (k != 0 ? a.y : (T)0);
    return cpx<T>{re, im};
}

// DEV (the deviatoric form): the products of the trace-free tensor
// D_ij = u_i u_j - delta_ij |u|^2/3, five of them: (D11, D12), (D13, D22) and
// D23 alone into buffers 0..4.  Its divergence differs from that of u_i u_j
// by a gradient, which the projection removes.
template <int N, typename T, bool DEV = false>
__global__ __launch_bounds__((XP<N, T>::THREADS)) void k_xp(SG g, W6<T> Wb,
                                                          const cpx<double>* __restrict__ twd,
                                                          T norm, double* __restrict__ blockmax) {
    using C = XP<N, T>;
    constexpr int E = C::E, TPF = C::TPF;
    constexpr int PS = kSwizzle;
    constexpr int FS = padded_size<PS>(N);
    constexpr int NK = Kc<N>::NK;
    static_assert(64 % TPF == 0 || TPF % 64 == 0, "a row is part of a wavefront or whole ones");
    // a row inside one wavefront synchronises per wavefront and splits the
    // outputs by lane shuffles; a row over several, per block and through LDS
    constexpr bool WAVE = TPF <= 64;
    __shared__ cpx<T> lds[C::ROWS][FS];
    __shared__ double redm[C::THREADS / 64 + 1];
    // With radix 8 throughout (N a power of 8) the five transforms of a row
    // share one set of twiddles, computed once per thread: FP32 x pass 1.02
    // -> 0.85 ms, FP64 1.72 -> 1.68 ms at 512^3; otherwise each transform
    // derives its own (three wavefronts per SIMD forced by the register
    // limit instead spill and double the time).
    // FP64: the shared twiddles live in a per-block LDS table instead of
    // registers (212 -> 158 VGPRs, three wavefronts per SIMD instead of two):
    // 1.51 -> 1.48 ms at 512^3
    constexpr bool CACHE = E == 8 && N >= 64 && N == (1 << (3 * stage_index(N, 8)));
    constexpr bool LC = CACHE && sizeof(T) == 8;
    using TC = std::conditional_t<CACHE, std::conditional_t<LC, TwLds<T, N, E>, TwCache<T, N, E>>,
                                  void>;
    __shared__ cpx<T> twl[LC ? TwLds<T, N, E>::SIZE : 1];
    const int t = threadIdx.x % TPF;
    const int rr = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + rr;
    const int partner = (TPF - t) & (TPF - 1);
    cpx<T>* L = lds[rr];
    TwCache<T, N, E> tcache;
    TwLds<T, N, E> twv{twl, t};
    const TC* tcp = nullptr;
    if constexpr (LC) {
        if (rr == 0) {
            TwCache<T, N, E> tcv;
            fill_twiddles<N, E>(tcv, t, twd + N);
#pragma unroll
            for (int si = 0; si < TwCache<T, N, E>::NS; ++si)
#pragma unroll
                for (int m = 1; m < E; ++m) twl[(si * E + m) * TPF + t] = tcv.w[si][m];
        }
        __syncthreads();
        tcp = &twv;
    } else if constexpr (CACHE) {
        fill_twiddles<N, E>(tcache, t, twd + N);
        tcp = &tcache;
    }
    T u[E], v[E], w[E];
    // w's row is loaded with u's and v's, ahead of the first transform
    // (registers allow it at the same occupancy)
    cpx<T> zw[1][E];
    {
        const cpx<T>* Wr = reinterpret_cast<const cpx<T>*>(Wb.p[2]) + row * g.PA;
#pragma unroll
        for (int s = 0; s < E; ++s) zw[0][s] = herm_load<N, TPF, NK>(Wr, t, s);
    }
    {
        const cpx<T>* U = reinterpret_cast<const cpx<T>*>(Wb.p[0]) + row * g.PA;
        const cpx<T>* V = reinterpret_cast<const cpx<T>*>(Wb.p[1]) + row * g.PA;
        cpx<T> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int k = t + s * TPF;
            const cpx<T> a = herm_load<N, TPF, NK>(U, t, s), b = herm_load<N, TPF, NK>(V, t, s);
            z[0][s] = cpx<T>{a.x - b.y, a.y + b.x};
        }
        multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            u[s] = z[0][s].x * norm;
            v[s] = z[0][s].y * norm;
        }
    }
    {
        cpx<T>(&z)[1][E] = zw;
        lds_sync<WAVE>();  // the previous transform's readers are done
        multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
#pragma unroll
        for (int s = 0; s < E; ++s) w[s] = z[0][s].x * norm;
    }
    // Courant number, rhs_velocity.f90:31-32 (FP64 max)
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const double a = ((double)fabs(u[s]) + (double)fabs(v[s])) + (double)fabs(w[s]);
        m = a > m ? a : m;
    }
    // product pairs (uu,uv), (uw,vv), (vw,ww): rhs_velocity.f90:56-72
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        cpx<T> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if constexpr (DEV) {
                const T t3 = ((u[s] * u[s] + v[s] * v[s]) + w[s] * w[s]) * (T)(1.0 / 3.0);
                if (pr == 0) z[0][s] = cpx<T>{u[s] * u[s] - t3, u[s] * v[s]};
                if (pr == 1) z[0][s] = cpx<T>{u[s] * w[s], v[s] * v[s] - t3};
                if (pr == 2) z[0][s] = cpx<T>{v[s] * w[s], (T)0};
            } else {
                if (pr == 0) z[0][s] = cpx<T>{u[s] * u[s], u[s] * v[s]};
                if (pr == 1) z[0][s] = cpx<T>{u[s] * w[s], v[s] * v[s]};
                if (pr == 2) z[0][s] = cpx<T>{v[s] * w[s], w[s] * w[s]};
            }
        }
        // DEV: the last pair holds one real row, whose second output (zero)
        // is not stored
        const bool store_q = !(DEV && pr == 2);
        lds_sync<WAVE>();
        multi_fft<N, E, -1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
        cpx<T>* outp = reinterpret_cast<cpx<T>*>(Wb.p[2 * pr]) + row * g.PA;
        cpx<T>* outq = reinterpret_cast<cpx<T>*>(Wb.p[2 * pr + 1]) + row * g.PA;
        if constexpr (WAVE) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                if (s * TPF < NK) {  // the same for every lane
                    const cpx<T> mine = (t == 0) ? z[0][(E - s) & (E - 1)] : z[0][E - 1 - s];
                    const cpx<T> zm = conj(cpx<T>{__shfl(mine.x, partner, TPF),
                                                  __shfl(mine.y, partner, TPF)});
                    const int k = t + s * TPF;
                    if (k < NK) {
                        const cpx<T> zk = z[0][s];
                        st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                        if (store_q) st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                    }
                }
            }
        } else {
            // Z_{N-k} comes through LDS (after the transform's last exchange
            // has been read); the next transform synchronises before it
            // writes the row buffer again
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) L[padi<PS>(t + s * TPF)] = z[0][s];
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const int k = t + s * TPF;
                if (s * TPF < NK && k < NK) {
                    const cpx<T> zm = conj(L[padi<PS>((N - k) & (N - 1))]);
                    const cpx<T> zk = z[0][s];
                    st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                    if (store_q) st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                }
            }
        }
    }
    // block max -> one partial per block (folded by the next launch: one
    // same-address atomic per block serialises at the L2)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < (C::THREADS + 63) / 64; ++i) mm = redm[i] > mm ? redm[i] : mm;
        blockmax[blockIdx.x] = mm;
    }
}

// ---------------------------------------------------------------- C (deviatoric, paired rows)
// The deviatoric x pass over two rows A and B per thread group.  Two real
// rows share one complex transform, and k_xp leaves two such slots half
// used per row: w's inverse and the forward transform of D23 = v w, the
// fifth product.  Here w_A + i w_B is one inverse transform and
// D23_A + i D23_B one forward transform, so the two rows take 8 transforms
// instead of 10 (each row: u + i v inverse, (D11, D12) and (D13, D22)
// forward).  w_B waits in registers while row A is processed, D23_A while
// row B is.  Rows B are ROWS after rows A; a block covers 2 ROWS rows.
// The two rows in one kernel cost registers beyond those two arrays (FP64
// at 512^3: 244 VGPRs, two wavefronts per SIMD, against 162 and three for
// k_xp; forced to three it spills), so the pairing pays only where k_xp's
// extra wavefront is worth less than a fifth of the transforms.
template <int N, typename T>
struct XP2 {
    static constexpr int E = XP<N, T>::E;
    static constexpr int TPF = XP<N, T>::TPF;
    static constexpr int ROWS = (256 / TPF < N * N / 2) ? 256 / TPF : N * N / 2;
    static constexpr int THREADS = ROWS * TPF;
};

template <int N, typename T>
__global__ __launch_bounds__((XP2<N, T>::THREADS)) void k_xp_pair(SG g, W6<T> Wb,
                                                                const cpx<double>* __restrict__ twd,
                                                                T norm, double* __restrict__ blockmax) {
    using C = XP2<N, T>;
    constexpr int E = C::E, TPF = C::TPF;
    constexpr int PS = kSwizzle;
    constexpr int FS = padded_size<PS>(N);
    constexpr int NK = Kc<N>::NK;
    static_assert(64 % TPF == 0 || TPF % 64 == 0, "a row is part of a wavefront or whole ones");
    constexpr bool WAVE = TPF <= 64;
    __shared__ cpx<T> lds[C::ROWS][FS];
    __shared__ double redm[C::THREADS / 64 + 1];
    constexpr bool CACHE = E == 8 && N >= 64 && N == (1 << (3 * stage_index(N, 8)));
    constexpr bool LC = CACHE && sizeof(T) == 8;
    using TC = std::conditional_t<CACHE, std::conditional_t<LC, TwLds<T, N, E>, TwCache<T, N, E>>,
                                  void>;
    __shared__ cpx<T> twl[LC ? TwLds<T, N, E>::SIZE : 1];
    const int t = threadIdx.x % TPF;
    const int rr = threadIdx.x / TPF;
    const int64_t rowA = (int64_t)blockIdx.x * (2 * C::ROWS) + rr;
    const int64_t rowB = rowA + C::ROWS;
    const int partner = (TPF - t) & (TPF - 1);
    cpx<T>* L = lds[rr];
    TwCache<T, N, E> tcache;
    TwLds<T, N, E> twv{twl, t};
    const TC* tcp = nullptr;
    if constexpr (LC) {
        if (rr == 0) {
            TwCache<T, N, E> tcv;
            fill_twiddles<N, E>(tcv, t, twd + N);
#pragma unroll
            for (int si = 0; si < TwCache<T, N, E>::NS; ++si)
#pragma unroll
                for (int m = 1; m < E; ++m) twl[(si * E + m) * TPF + t] = tcv.w[si][m];
        }
        __syncthreads();
        tcp = &twv;
    } else if constexpr (CACHE) {
        fill_twiddles<N, E>(tcache, t, twd + N);
        tcp = &tcache;
    }
    // inverse transform of two Hermitian half spectra a, b as one complex
    // row a + i b; the real results come back as (re, im)
    auto inverse_pair = [&](const T* pa, const T* pb, int64_t ra, int64_t rb, T(&ra_out)[E],
                            T(&rb_out)[E]) {
        const cpx<T>* A = reinterpret_cast<const cpx<T>*>(pa) + ra * g.PA;
        const cpx<T>* B = reinterpret_cast<const cpx<T>*>(pb) + rb * g.PA;
        cpx<T> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const cpx<T> a = herm_load<N, TPF, NK>(A, t, s), b = herm_load<N, TPF, NK>(B, t, s);
            z[0][s] = cpx<T>{a.x - b.y, a.y + b.x};
        }
        lds_sync<WAVE>();  // the previous transform's readers are done
        multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            ra_out[s] = z[0][s].x * norm;
            rb_out[s] = z[0][s].y * norm;
        }
    };
    // forward transform of two real rows p + i q, split by the Hermitian
    // symmetry into the half spectra stored at outp and outq
    auto forward_pair = [&](cpx<T>(&z)[1][E], cpx<T>* outp, cpx<T>* outq) {
        lds_sync<WAVE>();
        multi_fft<N, E, -1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
        if constexpr (WAVE) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                if (s * TPF < NK) {  // the same for every lane
                    const cpx<T> mine = (t == 0) ? z[0][(E - s) & (E - 1)] : z[0][E - 1 - s];
                    const cpx<T> zm = conj(cpx<T>{__shfl(mine.x, partner, TPF),
                                                  __shfl(mine.y, partner, TPF)});
                    const int k = t + s * TPF;
                    if (k < NK) {
                        const cpx<T> zk = z[0][s];
                        st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                        st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                    }
                }
            }
        } else {
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) L[padi<PS>(t + s * TPF)] = z[0][s];
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const int k = t + s * TPF;
                if (s * TPF < NK && k < NK) {
                    const cpx<T> zm = conj(L[padi<PS>((N - k) & (N - 1))]);
                    const cpx<T> zk = z[0][s];
                    st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                    st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                }
            }
        }
    };
    double m = 0.0;
    // one row: u + i v, the Courant max, (D11, D12) and (D13, D22); D23 = v w
    // is left in d23
    auto one_row = [&](int64_t row, const T(&w)[E], T(&d23)[E]) {
        T u[E], v[E];
        inverse_pair(Wb.p[0], Wb.p[1], row, row, u, v);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const double a = ((double)fabs(u[s]) + (double)fabs(v[s])) + (double)fabs(w[s]);
            m = a > m ? a : m;
        }
#pragma unroll
        for (int pr = 0; pr < 2; ++pr) {
            cpx<T> z[1][E];
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const T t3 = ((u[s] * u[s] + v[s] * v[s]) + w[s] * w[s]) * (T)(1.0 / 3.0);
                if (pr == 0) z[0][s] = cpx<T>{u[s] * u[s] - t3, u[s] * v[s]};
                if (pr == 1) {
                    z[0][s] = cpx<T>{u[s] * w[s], v[s] * v[s] - t3};
                    d23[s] = v[s] * w[s];  // u, v, w die here
                }
            }
            forward_pair(z, reinterpret_cast<cpx<T>*>(Wb.p[2 * pr]) + row * g.PA,
                         reinterpret_cast<cpx<T>*>(Wb.p[2 * pr + 1]) + row * g.PA);
        }
    };
    T wa[E], wb[E], da[E], db[E];
    inverse_pair(Wb.p[2], Wb.p[2], rowA, rowB, wa, wb);
    one_row(rowA, wa, da);
    one_row(rowB, wb, db);
    {
        cpx<T> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) z[0][s] = cpx<T>{da[s], db[s]};
        cpx<T>* out = reinterpret_cast<cpx<T>*>(Wb.p[4]);
        forward_pair(z, out + rowA * g.PA, out + rowB * g.PA);
    }
    // block max -> one partial per block
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < (C::THREADS + 63) / 64; ++i) mm = redm[i] > mm ? redm[i] : mm;
        blockmax[blockIdx.x] = mm;
    }
}

// ---------------------------------------------------------------- C (half-length rows)
// The x pass with one real row per complex transform of half the length:
// row x_0..x_{N-1} travels as z_m = x_{2m} + i x_{2m+1}, m < M = N/2.  At
// N = 1024 this puts a row's transforms (M = 512, radix 8) inside one
// wavefront, synchronised per wavefront, with the shared twiddle table of
// the 512-point transforms; k_xp's 1024-point row transforms span two
// wavefronts with block barriers at every exchange and ran at 0.33 of the
// HBM bandwidth.  Per row: 3 inverse and NOUT forward M-point transforms.
//   inverse (c2r of X_k, k <= K): Z_k = (X_k + conj X_{M-k})
//            + i e^{+2 pi i k/N} (X_k - conj X_{M-k}), z = M-point inverse of Z;
//   forward (r2c): Z = M-point transform of z,
//            X_k = (Z_k + conj Z_{M-k})/2 - i e^{-2 pi i k/N} (Z_k - conj Z_{M-k})/2.
// Entry k of a transform sits in lane t, slot s (k = t + s*TPF) and M-k in
// lane (TPF-t) mod TPF, slot E-1-s (slot (E-s) mod E for t = 0), so the
// forward split is a lane shuffle; the inverse loads X_{M-k} directly.
// Products in buffer order: divergence form uu, uv, uw, vv, vw, ww; DEV (the
// trace-free tensor) D11, D12, D13, D22, D23.  The e^{-2 pi i k/N} factors
// are one table entry per thread times constants per slot, in FP64.
template <int N, typename T>
struct XH {
    static constexpr int M = N / 2;
    static constexpr int E = 8;
    static constexpr int TPF = M / E;
    static constexpr int ROWS = 4;
    static constexpr int THREADS = ROWS * TPF;
};

template <int N, typename T, bool DEV>
__global__ __launch_bounds__((XH<N, T>::THREADS)) void k_xh(SG g, W6<T> Wb,
                                                          const cpx<double>* __restrict__ twd,
                                                          T norm, double* __restrict__ blockmax) {
    using C = XH<N, T>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    static_assert(N == 1024 && TPF == 64, "one row per wavefront; slot factors e^{-2 pi i s/16}");
    constexpr int PS = kSwizzle;
    constexpr int FS = padded_size<PS>(M);
    constexpr int K = Kc<N>::K;
    static_assert(K < M, "the Nyquist entry is outside the kept range");
    // the shared twiddles in a per-block LDS table, in both precisions
    // (FP32 in registers: 160 VGPRs, three wavefronts per SIMD; in LDS: 128,
    // four)
    using TC = TwLds<T, M, E>;
    __shared__ cpx<T> lds[C::ROWS][FS];
    __shared__ cpx<T> twl[TwLds<T, M, E>::SIZE];
    __shared__ double redm[C::ROWS];
    const int t = threadIdx.x % TPF;
    const int rr = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + rr;
    const int partner = (TPF - t) & (TPF - 1);
    cpx<T>* L = lds[rr];
    TwLds<T, M, E> twv{twl, t};
    const TC* tcp = &twv;
    if (rr == 0) {
        TwCache<T, M, E> tcv;
        fill_twiddles<M, E>(tcv, t, twd + M);
#pragma unroll
        for (int si = 0; si < TwCache<T, M, E>::NS; ++si)
#pragma unroll
            for (int m = 1; m < E; ++m) twl[(si * E + m) * TPF + t] = tcv.w[si][m];
    }
    __syncthreads();
    // e^{-2 pi i k/N} for k = t + s*TPF: a = e^{-2 pi i t/N} times the
    // constant e^{-2 pi i s/16} (TPF/N = 1/16)
    const cpx<double> at = twd[N + t];
    auto wk = [&](int s) -> cpx<double> {
        constexpr double c[8] = {1.0,
                                 0.92387953251128675612818318939678828682,
                                 0.70710678118654752440084436210484903928,
                                 0.38268343236508977172845998403039886676,
                                 0.0,
                                 -0.38268343236508977172845998403039886676,
                                 -0.70710678118654752440084436210484903928,
                                 -0.92387953251128675612818318939678828682};
        // e^{-2 pi i s/16} = cos(pi s/8) - i sin(pi s/8); sin(pi s/8) = cos(pi (4-s)/8)
        constexpr double sn[8] = {0.0,
                                  0.38268343236508977172845998403039886676,
                                  0.70710678118654752440084436210484903928,
                                  0.92387953251128675612818318939678828682,
                                  1.0,
                                  0.92387953251128675612818318939678828682,
                                  0.70710678118654752440084436210484903928,
                                  0.38268343236508977172845998403039886676};
        // recomputed at every use: kept across the transforms, the eight
        // factors would hold 32 registers for the whole kernel
        cpx<double> a = at;
        asm volatile("" : "+v"(a.x), "+v"(a.y));
        const int s4 = s & 3;
        const cpx<double> r = s4 == 0 ? a : cmul(a, cpx<double>{c[s4], -sn[s4]});
        return s >= 4 ? mul_mi(r) : r;  // e^{-2 pi i (s+4)/16} = -i e^{-2 pi i s/16}
    };
    // inverse: half spectrum of buffer b -> x_{2m} + i x_{2m+1} (times norm)
    auto inverse = [&](int b, cpx<T>(&z)[1][E]) {
        const cpx<T>* X = reinterpret_cast<const cpx<T>*>(Wb.p[b]) + row * g.PA;
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int k = t + s * TPF;
            // A = X_k (k <= K; the imaginary part of X_0 is ignored, as by
            // the reference's c2r), B = conj X_{M-k} (M-k <= K); branch-free
            // loads from in-bounds addresses, slots with no live lane load
            // nothing (constant once unrolled)
            cpx<T> A{(T)0, (T)0}, B{(T)0, (T)0};
            if (s * TPF <= K) {
                const bool live = k <= K;
                const cpx<T> a = X[live ? k : 0];
                A = cpx<T>{live ? a.x : (T)0, (live && k != 0) ? a.y : (T)0};
            }
            if ((s + 1) * TPF - 1 >= M - K) {
                const int km = M - k;
                const bool live = km <= K;
                const cpx<T> a = X[live ? km : 0];
                B = cpx<T>{live ? a.x : (T)0, live ? -a.y : (T)0};
            }
            const cpx<double> w = wk(s);
            const cpx<T> d = A - B;
            // i conj(w) d
            const cpx<T> wc{(T)w.x, (T)-w.y};
            const cpx<T> r = mul_i(cmul(wc, d));
            z[0][s] = (A + B) + r;
        }
        multi_fft<M, E, 1, T, 1, 1, FS, PS, true, true, double, TC>(z, t, L, twd + M, tcp);
#pragma unroll
        for (int s = 0; s < E; ++s) z[0][s] = scale(z[0][s], norm);
    };
    // forward: z (x_{2m} + i x_{2m+1}) -> half spectrum k <= K into buffer b
    auto forward = [&](cpx<T>(&z)[1][E], int b) {
        lds_sync<true>();  // the previous transform's readers are done
        multi_fft<M, E, -1, T, 1, 1, FS, PS, true, true, double, TC>(z, t, L, twd + M, tcp);
        cpx<T>* out = reinterpret_cast<cpx<T>*>(Wb.p[b]) + row * g.PA;
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (s * TPF <= K) {  // the same for every lane
                const cpx<T> mine = (t == 0) ? z[0][(E - s) & (E - 1)] : z[0][E - 1 - s];
                const cpx<T> zm = conj(cpx<T>{__shfl(mine.x, partner, TPF),
                                              __shfl(mine.y, partner, TPF)});
                const int k = t + s * TPF;
                if (k <= K) {
                    const cpx<T> zk = z[0][s];
                    const cpx<double> w = wk(s);
                    const cpx<T> wt{(T)w.x, (T)w.y};
                    // (zk + zm)/2 - i w (zk - zm)/2
                    const cpx<T> o = mul_mi(cmul(wt, zk - zm));
                    st_g<kStream<N>>(&out[k], scale((zk + zm) + o, (T)0.5));
                }
            }
        }
    };
    cpx<T> u[1][E], v[1][E], w[1][E];
    inverse(0, u);
    lds_sync<true>();
    inverse(1, v);
    lds_sync<true>();
    inverse(2, w);
    // Courant number, rhs_velocity.f90:31-32 (FP64 max)
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const double a0 = ((double)fabs(u[0][s].x) + (double)fabs(v[0][s].x)) + (double)fabs(w[0][s].x);
        const double a1 = ((double)fabs(u[0][s].y) + (double)fabs(v[0][s].y)) + (double)fabs(w[0][s].y);
        m = a0 > m ? a0 : m;
        m = a1 > m ? a1 : m;
    }
    // products of the two real values each entry holds (rhs_velocity.f90:56-72)
    auto prod = [&](int which, cpx<T>(&z)[1][E]) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const cpx<T> a = u[0][s], b = v[0][s], c = w[0][s];
            cpx<T> r;
            if constexpr (DEV) {
                // D11 = u u - |u|^2/3, D22 = v v - |u|^2/3 (same grouping as k_xp)
                if (which == 0 || which == 3) {
                    // recomputed for D22: kept from D11, the trace held
                    // registers across a transform (FP32: 179 -> 128 VGPRs)
                    T ax = a.x, ay = a.y, bx = b.x, by = b.y, cx = c.x, cy = c.y;
                    asm volatile("" : "+v"(ax), "+v"(ay), "+v"(bx), "+v"(by), "+v"(cx), "+v"(cy));
                    const T t3x = ((ax * ax + bx * bx) + cx * cx) * (T)(1.0 / 3.0);
                    const T t3y = ((ay * ay + by * by) + cy * cy) * (T)(1.0 / 3.0);
                    r = which == 0 ? cpx<T>{a.x * a.x - t3x, a.y * a.y - t3y}
                                   : cpx<T>{b.x * b.x - t3x, b.y * b.y - t3y};
                }
                if (which == 1) r = cpx<T>{a.x * b.x, a.y * b.y};
                if (which == 2) r = cpx<T>{a.x * c.x, a.y * c.y};
                if (which == 4) r = cpx<T>{b.x * c.x, b.y * c.y};
            } else {
                if (which == 0) r = cpx<T>{a.x * a.x, a.y * a.y};
                if (which == 1) r = cpx<T>{a.x * b.x, a.y * b.y};
                if (which == 2) r = cpx<T>{a.x * c.x, a.y * c.y};
                if (which == 3) r = cpx<T>{b.x * b.x, b.y * b.y};
                if (which == 4) r = cpx<T>{b.x * c.x, b.y * c.y};
                if (which == 5) r = cpx<T>{c.x * c.x, c.y * c.y};
            }
            z[0][s] = r;
        }
    };
    // order: the products of u first, so that u dies early
    constexpr int NOUT = DEV ? 5 : 6;
    constexpr int order_dev[5] = {0, 3, 1, 2, 4};
    constexpr int order_div[6] = {0, 1, 2, 3, 4, 5};
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
        const int which = DEV ? order_dev[i] : order_div[i];
        cpx<T> z[1][E];
        prod(which, z);
        forward(z, which);
    }
    // block max -> one partial per block (folded by the next launch)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < C::ROWS; ++i) mm = redm[i] > mm ? redm[i] : mm;
        blockmax[blockIdx.x] = mm;
    }
}

// ---------------------------------------------------------------- C (rotational)
// The x pass of the rotational form: three inverse transforms of row pairs
// (u + i v, w + i omega_1, omega_2 + i omega_3), the Courant max, the cross
// product c = u x omega (c_1 = v w3 - w w2, c_2 = w w1 - u w3,
// c_3 = u w2 - v w1), then two forward transforms (c_1 + i c_2 split by the
// Hermitian symmetry, and c_3 alone) into buffers 0, 1, 2: five transforms
// per row, as in the divergence form.
template <int N, typename T>
__global__ __launch_bounds__((XP<N, T>::THREADS)) void k_xp_rot(SG g, W6<T> Wb,
                                                              const cpx<double>* __restrict__ twd,
                                                              T norm, double* __restrict__ blockmax) {
    using C = XP<N, T>;
    constexpr int E = C::E, TPF = C::TPF;
    constexpr int PS = kSwizzle;
    constexpr int FS = padded_size<PS>(N);
    constexpr int NK = Kc<N>::NK;
    static_assert(64 % TPF == 0 || TPF % 64 == 0, "a row is part of a wavefront or whole ones");
    constexpr bool WAVE = TPF <= 64;
    __shared__ cpx<T> lds[C::ROWS][FS];
    __shared__ double redm[C::THREADS / 64 + 1];
    // twiddles cached in registers, as in k_xp's FP32 form
    constexpr bool CACHE = E == 8 && N >= 64 && N == (1 << (3 * stage_index(N, 8)));
    constexpr bool LC = false;  // the LDS table made this pass slower (1.43 -> 1.66 ms)
    using TC = std::conditional_t<CACHE, std::conditional_t<LC, TwLds<T, N, E>, TwCache<T, N, E>>,
                                  void>;
    __shared__ cpx<T> twl[LC ? TwLds<T, N, E>::SIZE : 1];
    const int t = threadIdx.x % TPF;
    const int rr = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + rr;
    const int partner = (TPF - t) & (TPF - 1);
    cpx<T>* L = lds[rr];
    TwCache<T, N, E> tcache;
    TwLds<T, N, E> twv{twl, t};
    const TC* tcp = nullptr;
    if constexpr (LC) {
        if (rr == 0) {
            TwCache<T, N, E> tcv;
            fill_twiddles<N, E>(tcv, t, twd + N);
#pragma unroll
            for (int si = 0; si < TwCache<T, N, E>::NS; ++si)
#pragma unroll
                for (int m = 1; m < E; ++m) twl[(si * E + m) * TPF + t] = tcv.w[si][m];
        }
        __syncthreads();
        tcp = &twv;
    } else if constexpr (CACHE) {
        fill_twiddles<N, E>(tcache, t, twd + N);
        tcp = &tcache;
    }
    auto rowp = [&](int b) { return reinterpret_cast<const cpx<T>*>(Wb.p[b]) + row * g.PA; };
    // z = a + i b of two Hermitian-extended half spectra
    auto pair_load = [&](int a, int b, cpx<T>(&z)[1][E]) {
        const cpx<T>* A = rowp(a);
        const cpx<T>* Bp = rowp(b);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const cpx<T> x = herm_load<N, TPF, NK>(A, t, s), y = herm_load<N, TPF, NK>(Bp, t, s);
            z[0][s] = cpx<T>{x.x - y.y, x.y + y.x};
        }
    };
    T u[E], v[E], w[E], o1[E];
    cpx<T> z1[1][E];
    pair_load(2, 3, z1);  // w + i omega_1, loaded ahead of the first transform
    {
        cpx<T> z[1][E];
        pair_load(0, 1, z);
        multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z, t, L, twd + N, tcp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            u[s] = z[0][s].x * norm;
            v[s] = z[0][s].y * norm;
        }
    }
    cpx<T> z2[1][E];
    pair_load(4, 5, z2);  // omega_2 + i omega_3, in flight during the next transform
    lds_sync<WAVE>();
    multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z1, t, L, twd + N, tcp);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        w[s] = z1[0][s].x * norm;
        o1[s] = z1[0][s].y * norm;
    }
    lds_sync<WAVE>();
    multi_fft<N, E, 1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z2, t, L, twd + N, tcp);
    // Courant number, rhs_velocity.f90:31-32 (FP64 max)
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const double a = ((double)fabs(u[s]) + (double)fabs(v[s])) + (double)fabs(w[s]);
        m = a > m ? a : m;
    }
    // u x omega; c_1 + i c_2 in z1, c_3 in z2 (real)
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const T o2 = z2[0][s].x * norm, o3 = z2[0][s].y * norm;
        z1[0][s] = cpx<T>{v[s] * o3 - w[s] * o2, w[s] * o1[s] - u[s] * o3};
        z2[0][s] = cpx<T>{u[s] * o2 - v[s] * o1[s], (T)0};
    }
    lds_sync<WAVE>();
    multi_fft<N, E, -1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z1, t, L, twd + N, tcp);
    {
        cpx<T>* outp = reinterpret_cast<cpx<T>*>(Wb.p[0]) + row * g.PA;
        cpx<T>* outq = reinterpret_cast<cpx<T>*>(Wb.p[1]) + row * g.PA;
        if constexpr (WAVE) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                if (s * TPF < NK) {
                    const cpx<T> mine = (t == 0) ? z1[0][(E - s) & (E - 1)] : z1[0][E - 1 - s];
                    const cpx<T> zm = conj(cpx<T>{__shfl(mine.x, partner, TPF),
                                                  __shfl(mine.y, partner, TPF)});
                    const int k = t + s * TPF;
                    if (k < NK) {
                        const cpx<T> zk = z1[0][s];
                        st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                        st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                    }
                }
            }
        } else {
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) L[padi<PS>(t + s * TPF)] = z1[0][s];
            lds_sync<false>();
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const int k = t + s * TPF;
                if (s * TPF < NK && k < NK) {
                    const cpx<T> zm = conj(L[padi<PS>((N - k) & (N - 1))]);
                    const cpx<T> zk = z1[0][s];
                    st_g<kStream<N>>(&outp[k], scale(zk + zm, (T)0.5));
                    st_g<kStream<N>>(&outq[k], scale(mul_mi(zk - zm), (T)0.5));
                }
            }
        }
    }
    lds_sync<WAVE>();
    multi_fft<N, E, -1, T, 1, 1, FS, PS, WAVE, true, double, TC>(z2, t, L, twd + N, tcp);
    {
        // a real row's transform is its own half spectrum
        cpx<T>* outr = reinterpret_cast<cpx<T>*>(Wb.p[2]) + row * g.PA;
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int k = t + s * TPF;
            if (s * TPF < NK && k < NK) st_g<kStream<N>>(&outr[k], z2[0][s]);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < (C::THREADS + 63) / 64; ++i) mm = redm[i] > mm ? redm[i] : mm;
        blockmax[blockIdx.x] = mm;
    }
}

// ---------------------------------------------------------------- E
struct Upd {
    double* F[3];
    double* R[3];  // rhs_old: FP64, or FP32 values for the mixed form
    const double* eshell;  // band energy per shell 1..kfmax (device)
    double* energy_out;
    const double* dtp;  // the step size (device value)
    double nu, famp;
    int kfmax, forced, first;
};

// The column kx = ky = 0 of a product holds the product's plane sums, whose
// common part (N^2 times the product's mean over the box) dwarfs every
// other input of the z transforms.  In FP32 the butterflies that cancel it
// leave errors of order eps N^3 |mean| on the high kz of that column, which
// k_j P_ij multiplies by kz: the largest error of the mixed RHS, growing with
// N.  Taking the column's mean out first changes only its kz = 0 output,
// which the RHS multiplies by k = 0.  Called by the whole block (barriers);
// `mine` marks the threads holding the column, t its thread position; lds
// is the block's LDS tile of LDSN entries.
template <int N, int E, int TPF, int LDSN, typename T>
__device__ void remove_column_mean(cpx<T> (&v)[E], int t, bool mine, cpx<T>* lds) {
    static_assert(TPF * sizeof(cpx<double>) <= LDSN * sizeof(cpx<T>), "partials exceed the LDS tile");
    cpx<double>* part = reinterpret_cast<cpx<double>*>(lds);
    cpx<double> s{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < E; ++i) s = {s.x + (double)v[i].x, s.y + (double)v[i].y};
    __syncthreads();  // the previous transform's last LDS reads are done
    if (mine) part[t] = s;
    __syncthreads();
    if (mine) {
        cpx<double> tot{0.0, 0.0};
        for (int i = 0; i < TPF; ++i) tot = {tot.x + part[i].x, tot.y + part[i].y};
        const cpx<T> m{(T)(tot.x / N), (T)(tot.y / N)};
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = {v[i].x - m.x, v[i].y - m.y};
    }
    __syncthreads();  // before the transform writes LDS
}

// FORM 0: six buffers hold the products u_i u_j, r_i = -i k_j (u_i u_j)^ -
// nu k^2 u_i.  FORM 1 (rotational): three hold (u x omega)_i, the nonlinear
// term itself, r_i = (u x omega)_i^ - nu k^2 u_i.  FORM 2 (deviatoric): five
// hold D11, D12, D13, D22, D23 of the trace-free tensor, D33 = -(D11 + D22).
template <int N, typename T, int FORM = 0>
__global__ __launch_bounds__((CZ<N, T>::THREADS)) void k_fwd_z_update(
    SG g, W6<T> Q, Upd u, const cpx<double>* __restrict__ twd) {
    constexpr bool ROT = FORM == 1;
    constexpr int NP = FORM == 1 ? 3 : FORM == 2 ? 5 : 6;
    constexpr int E = CZ<N, T>::E, B = CZ<N, T>::B, TPF = CZ<N, T>::TPF;
    constexpr int PS = 0;
    __shared__ cpx<T> lds[padded_size<PS>(N * B)];
    int bx, by;
    block_coords<kPairHalves<T, B>>(bx, by);
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = bx * B + c;
    const bool valid = kx < Kc<N>::NK;
    const int jy = kept_row(by, Kc<N>::K, N);
    const int ky = wavenum(jy, N);
    const T kxT = (T)kx, kyT = (T)ky;
    // FP64 with radix 8 throughout: the six transforms share the twiddles of
    // each thread position, computed once per block into LDS (registers are
    // full): 1.66 -> 1.63 ms at 512^3 (FP32 gained nothing)
    constexpr bool LCACHE =
        sizeof(T) == 8 && E == 8 && N >= 64 && N == (1 << (3 * stage_index(N, 8)));
    using TL = std::conditional_t<LCACHE, TwLds<T, N, E>, void>;
    __shared__ cpx<T> twl[LCACHE ? TwLds<T, N, E>::SIZE : 1];
    TwLds<T, N, E> twv{twl, t};
    const TL* tlp = nullptr;
    if constexpr (LCACHE) {
        if (c == 0) {
            TwCache<T, N, E> tcv;
            fill_twiddles<N, E>(tcv, t, twd + N);
#pragma unroll
            for (int si = 0; si < TwCache<T, N, E>::NS; ++si)
#pragma unroll
                for (int m = 1; m < E; ++m) twl[(si * E + m) * TPF + t] = tcv.w[si][m];
        }
        __syncthreads();
        tlp = &twv;
    }
    // t_i = k_j P_ij accumulated in the reference's order (rhs_velocity.f90:89-91)
    cpx<T> acc[3][E];
#pragma unroll
    for (int s = 0; s < E; ++s) acc[0][s] = acc[1][s] = acc[2][s] = cpx<T>{(T)0, (T)0};
    // FP64: the column of product p + 1 is loaded while product p is
    // transformed (FP32 blocks have no registers to spare for it)
    constexpr bool PREFETCH = sizeof(T) == 8;
    auto load = [&](int p, cpx<T>(&dst)[E]) {
        // select the buffer without indexing the argument array at run time
        const T* qp = Q.p[0];
        if (p == 1) qp = Q.p[1];
        if (p == 2) qp = Q.p[2];
        if (p == 3) qp = Q.p[3];
        if (p == 4) qp = Q.p[4];
        if (p == 5) qp = Q.p[5];
        const cpx<T>* src = reinterpret_cast<const cpx<T>*>(qp) + (int64_t)jy * g.PA + kx;
#pragma unroll
        for (int s = 0; s < E; ++s)
            dst[s] = valid ? ld_g<kStreamLoad<N, T, B>>(&src[(int64_t)(t + s * TPF) * N * g.PA]) : cpx<T>{(T)0, (T)0};
    };
    cpx<T> nxt[E];
    if (PREFETCH) load(0, nxt);
#pragma unroll 1
    for (int p = 0; p < NP; ++p) {
        cpx<T> v[E];
        if (PREFETCH) {
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = nxt[s];
            if (p < NP - 1) load(p + 1, nxt);
        } else {
            load(p, v);
        }
        // Twiddles by recurrence from one FP64 table load per butterfly (in
        // FP64 for the FP32 form too, then rounded once), from a pointer the
        // compiler cannot see is loop-invariant, so they are recomputed per
        // product instead of holding fourteen complex values in registers
        // across the loop.
        const cpx<double>* twp = twd + N;
        asm volatile("" : "+s"(twp));
        if constexpr (sizeof(T) == 4)
            if (bx == 0 && jy == 0)
                remove_column_mean<N, E, TPF, padded_size<PS>(N * B)>(v, t, c == 0, lds);
        block_fft<N, E, -1, T, B, PS, true, double, TL>(v, t, lds + c, twp, tlp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (!slot_live<N, TPF>(s) || !kept(t + s * TPF, Kc<N>::K, N)) continue;  // never written out
            if constexpr (ROT) {
                if (p == 0) acc[0][s] = v[s];
                if (p == 1) acc[1][s] = v[s];
                if (p == 2) acc[2][s] = v[s];
                continue;
            }
            const T kzT = (T)wavenum(t + s * TPF, N);
            if constexpr (FORM == 2) {
                // t_3 = kx D13 + ky D23 - kz (D11 + D22)
                if (p == 0) {
                    acc[0][s] = scale(v[s], kxT);
                    acc[2][s] = scale(v[s], -kzT);
                }
                if (p == 1) {
                    acc[0][s] = acc[0][s] + scale(v[s], kyT);
                    acc[1][s] = scale(v[s], kxT);
                }
                if (p == 2) {
                    acc[0][s] = acc[0][s] + scale(v[s], kzT);
                    acc[2][s] = acc[2][s] + scale(v[s], kxT);
                }
                if (p == 3) {
                    acc[1][s] = acc[1][s] + scale(v[s], kyT);
                    acc[2][s] = acc[2][s] - scale(v[s], kzT);
                }
                if (p == 4) {
                    acc[1][s] = acc[1][s] + scale(v[s], kzT);
                    acc[2][s] = acc[2][s] + scale(v[s], kyT);
                }
                continue;
            }
            // p: 0 uu, 1 uv, 2 uw, 3 vv, 4 vw, 5 ww
            if (p == 0) acc[0][s] = scale(v[s], kxT);
            if (p == 1) {
                acc[0][s] = acc[0][s] + scale(v[s], kyT);
                acc[1][s] = scale(v[s], kxT);
            }
            if (p == 2) {
                acc[0][s] = acc[0][s] + scale(v[s], kzT);
                acc[2][s] = scale(v[s], kxT);
            }
            if (p == 3) acc[1][s] = acc[1][s] + scale(v[s], kyT);
            if (p == 4) {
                acc[1][s] = acc[1][s] + scale(v[s], kzT);
                acc[2][s] = acc[2][s] + scale(v[s], kyT);
            }
            if (p == 5) acc[2][s] = acc[2][s] + scale(v[s], kzT);
        }
    }
    // band forcing factor, m_force.f90:153-161
    double fac = 0.0;
    if (u.forced) {
        double en = 0.0;
        for (int i = 0; i < u.kfmax; ++i) en += u.eshell[i];
        fac = u.famp / en;
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *u.energy_out = en;
    }
    if (!valid) return;
    const double dt = *u.dtp;
    const double kxd = kx, kyd = ky;
    // The slots in NCH chunks; in each, every fields / rhs_old load is issued
    // before the first store (the stores may alias later loads as far as the
    // compiler knows, which would otherwise serialise one memory round trip
    // per load).  FP32 blocks are twice as wide, so they get half the
    // registers per thread: smaller chunks.
    constexpr int NCH = sizeof(T) == 8 ? 2 : E, H = E / NCH;
#pragma unroll
    for (int h = 0; h < NCH; ++h) {
        cpx<double> f[3][H], o[3][H];
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = h * H + q, e = t + s * TPF;
            if (!slot_live<N, TPF>(s) || !kept(e, Kc<N>::K, N)) continue;
            const int64_t off = ((int64_t)jy * N + e) * g.P + kx;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                f[cc][q] = ld_g<kStreamLoad<N, T, B>>(&reinterpret_cast<const cpx<double>*>(u.F[cc])[off]);
                if (!u.first) {
                    if constexpr (sizeof(T) == 8) {
                        o[cc][q] = ld_g<kStreamLoad<N, T, B>>(&reinterpret_cast<const cpx<double>*>(u.R[cc])[off]);
                    } else {
                        const cpx<float> r32 = ld_g<kStreamLoad<N, T, B>>(&reinterpret_cast<const cpx<float>*>(u.R[cc])[off]);
                        o[cc][q] = cpx<double>{(double)r32.x, (double)r32.y};
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = h * H + q, e = t + s * TPF;
            if (!slot_live<N, TPF>(s) || !kept(e, Kc<N>::K, N)) continue;
            const int kz = wavenum(e, N);
            const double kzd = kz;
            const int64_t off = ((int64_t)jy * N + e) * g.P + kx;
            const double k2 = kxd * kxd + kyd * kyd + kzd * kzd;
            const double lap = -u.nu * k2;
            const bool fnode = u.forced && kx < N / 2 && [&] {
                const int sh = shell_of(kx * kx + ky * ky + kz * kz);
                return sh > 0 && sh <= u.kfmax;
            }();
            cpx<double> fn[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const cpx<double> fc = f[cc][q];
                cpx<double> r;
                if constexpr (ROT) {
                    // r = (u x omega)^ + lap * f
                    r = {(double)acc[cc][s].x + lap * fc.x, (double)acc[cc][s].y + lap * fc.y};
                } else {
                    // r = i * (-t) + lap * f, t = k.P  (rhs_velocity.f90:89-136)
                    const double tre = -(double)acc[cc][s].x, tim = -(double)acc[cc][s].y;
                    r = {-tim + lap * fc.x, tre + lap * fc.y};
                }
                if (fnode) r = {r.x + fac * fc.x, r.y + fac * fc.y};
                if (u.first) {
                    fn[cc] = {fc.x + dt * r.x, fc.y + dt * r.y};
                } else {
                    const cpx<double> oc = o[cc][q];
                    fn[cc] = {fc.x + dt * (1.5 * r.x - 0.5 * oc.x),
                              fc.y + dt * (1.5 * r.y - 0.5 * oc.y)};
                }
                st_g<kStream<N>>(&reinterpret_cast<cpx<T>*>(u.R[cc])[off], to<T>(r));
            }
            // projection, pressure.f90:21-43
            const double dre = kxd * fn[0].x + kyd * fn[1].x + kzd * fn[2].x;
            const double dim = kxd * fn[0].y + kyd * fn[1].y + kzd * fn[2].y;
            const double lapl = k2 == 0.0 ? 9e20 : k2;
            const double pr = dre / lapl, pi = dim / lapl;
            st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[0])[off], cpx<double>{fn[0].x - pr * kxd, fn[0].y - pi * kxd});
            st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[1])[off], cpx<double>{fn[1].x - pr * kyd, fn[1].y - pi * kyd});
            st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[2])[off], cpx<double>{fn[2].x - pr * kzd, fn[2].y - pi * kzd});
        }
    }
}

// ---------------------------------------------------------------- host
template <typename F>
static void dispatch_step(int n, F&& f) {
    switch (n) {
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        case 128: f(std::integral_constant<int, 128>{}); break;
        case 256: f(std::integral_constant<int, 256>{}); break;
        case 512: f(std::integral_constant<int, 512>{}); break;
        case 1024: f(std::integral_constant<int, 1024>{}); break;
        default: throw HipError{"fused step: unsupported grid size"};
    }
}

// column blocks of B columns covering nk, rounded up to an even count where
// the kernel pairs half-segment blocks (block_coords)
template <typename T, int B>
static int col_blocks(int nk) {
    const int c = (nk + B - 1) / B;
    return kPairHalves<T, B> ? c + (c & 1) : c;
}

template <int N, typename T, int FORM>
static void run(Lib& L, const double* dtp, bool first, const cpx<T>* tw) {
    constexpr bool ROT = FORM == 1;
    constexpr int NFWD = FORM == 1 ? 3 : FORM == 2 ? 5 : 6;
    if (L.kmax != Kc<N>::K) throw std::runtime_error("the fused step expects kmax = n/3");
    SG g;
    g.K = L.kmax;
    g.NK = L.kmax + 1;
    g.P = L.g.P;
    g.PA = ((g.NK + 15) / 16) * 16;
    P3 F{{L.F[1], L.F[2], L.F[3]}};
    W6<T> Wb;
    for (int i = 0; i < 6; ++i) Wb.p[i] = reinterpret_cast<T*>(L.W[i + 1]);
    constexpr int B = CE<N, T>::B;
    const int ncb = (g.NK + B - 1) / B;
    const int nky = 2 * g.K + 1;
    const bool forced = L.flow_type == 1;
    if (forced) {
        Timed tm(TC_REDUCE);
        launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kfmax, L.dshell, L.st);
    }
    {
        Timed tm(TC_STEP_INV_Z);
        if constexpr (ROT)
            hipLaunchKernelGGL((k_inv_z_rot<N, T>),
                               dim3(col_blocks<T, CR<N, T>::B>(g.NK), nky),
                               dim3(CR<N, T>::THREADS), 0, L.st, g, F, Wb, tw);
        else if constexpr (N >= 1024 && sizeof(T) == 8) {
            const int ncb = (g.NK + CE<N, T>::B - 1) / CE<N, T>::B;
            const int ngroups = ncb * nky * 3;
            static int grid = 0;
            if (grid == 0) grid = resident_blocks(k_inv_zp<N, T>, CE<N, T>::THREADS, L.ncu);
            hipLaunchKernelGGL((k_inv_zp<N, T>), dim3(std::min(grid, ngroups)),
                               dim3(CE<N, T>::THREADS), 0, L.st, g, F, Wb, tw, ncb, ngroups);
        } else {
            const int ncb = (g.NK + CE<N, T>::B - 1) / CE<N, T>::B;
            hipLaunchKernelGGL((k_inv_z<N, T>), dim3(ncb, nky, 3), dim3(CE<N, T>::THREADS), 0,
                               L.st, g, F, Wb, tw);
        }
    }
    {
        Timed tm(TC_STEP_INV_Y);
        launch_y<N, 1, T>(L, g, Wb, tw, ROT ? 6 : 3);
    }
    {
        Timed tm(TC_STEP_X);
        const T norm = (T)(1.0 / ((double)N * (double)N * (double)N));
        // the deviatoric form's paired rows: FP64 up to 512^3 (1.45 -> 1.38 ms
        // there at two wavefronts per SIMD, where k_xp runs three); at 1024^3
        // the paired kernel needs all 256 registers (one wavefront per SIMD,
        // 17.2 -> 22.3 ms), and in FP32 it gained nothing
        constexpr bool PAIR = FORM == 2 && sizeof(T) == 8 && N <= 512;
        // 1024^3: half-length row transforms inside one wavefront
        constexpr bool HALF = N == 1024 && !ROT;
        // blocks <= n*n, the size of dmax
        const int64_t nb = PAIR   ? (int64_t)N * N / (2 * XP2<N, T>::ROWS)
                           : HALF ? (int64_t)N * N / XH<N, T>::ROWS
                                  : (int64_t)N * N / XP<N, T>::ROWS;
        // FP64 twiddles in both forms (the FP32 form's recurrence in FP64)
        if constexpr (HALF)
            hipLaunchKernelGGL((k_xh<N, T, FORM == 2>), dim3(nb), dim3(XH<N, T>::THREADS), 0,
                               L.st, g, Wb, L.tw64, norm, L.dmax);
        else if constexpr (ROT)
            hipLaunchKernelGGL((k_xp_rot<N, T>), dim3(nb), dim3(XP<N, T>::THREADS), 0, L.st, g,
                               Wb, L.tw64, norm, L.dmax);
        else if constexpr (PAIR)
            hipLaunchKernelGGL((k_xp_pair<N, T>), dim3(nb), dim3(XP2<N, T>::THREADS), 0, L.st, g,
                               Wb, L.tw64, norm, L.dmax);
        else
            hipLaunchKernelGGL((k_xp<N, T, FORM == 2>), dim3(nb), dim3(XP<N, T>::THREADS), 0,
                               L.st, g, Wb, L.tw64, norm, L.dmax);
        launch_max_fold(L.dmax, nb, L.dscal, L.st);
    }
    {
        Timed tm(TC_STEP_FWD_Y);
        launch_y<N, -1, T>(L, g, Wb, tw, NFWD);
    }
    {
        Timed tm(TC_STEP_FWD_Z);
        Upd u;
        for (int i = 0; i < 3; ++i) {
            u.F[i] = L.F[i + 1];
            u.R[i] = L.R[i + 1];
        }
        u.eshell = L.dshell;
        u.energy_out = L.dscal + 1;
        u.dtp = dtp;
        u.nu = L.nu;
        u.famp = L.famp;
        u.kfmax = L.kfmax;
        u.forced = forced ? 1 : 0;
        u.first = first ? 1 : 0;
        const int ncbz = col_blocks<T, CZ<N, T>::B>(g.NK);
        hipLaunchKernelGGL((k_fwd_z_update<N, T, FORM>), dim3(ncbz, nky),
                           dim3(CZ<N, T>::THREADS), 0, L.st, g, Wb, u, L.tw64);
    }
}

}  // namespace

bool fused_step_supported(bool) {
    Lib& L = lib();
    return L.dealias == 0 && L.nles == 0 && L.n >= 16 && L.kmax >= 1;
}

void fused_step(const double* dtp, bool first, bool mixed) {
    Lib& L = lib();
    dispatch_step(L.n, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        auto go = [&](auto TT) {
            using T = decltype(TT);
            const cpx<T>* tw = mixed ? reinterpret_cast<const cpx<T>*>(L.tw32)
                                     : reinterpret_cast<const cpx<T>*>(L.tw64);
            if (L.nlform == 1) run<N, T, 1>(L, dtp, first, tw);
            else if (L.nlform == 2) run<N, T, 2>(L, dtp, first, tw);
            else run<N, T, 0>(L, dtp, first, tw);
        };
        if (mixed) go(float{});
        else go(double{});
    });
}

}  // namespace hit3d
