// Fused velocity right-hand side of the phase-shift step (rhs_velocity.f90:
// 150-410 with force_velocity, m_force.f90:101-182, the Euler/AB2 update,
// main.f90:156-164, and pressure, pressure.f90:13-47), FP64.
//
// The reference forms the RHS in two halves: the products of the physical
// velocity, and the products of the velocity shifted by half a cell
// (exp(i theta), theta = (kx+ky+kz) dx/2), shifted back.  Here each half is
// three passes over HBM:
//   x  k_xprod6: physical u, v, w rows -> the six products' half spectra
//      (real-to-complex through an N/2-point transform each; the unshifted
//      half also takes the Courant max);
//   y  the forward column pass of each product (in place);
//   z  k_zps: the forward z transform of all six products of a column at
//      once and the assembly -1/2 i k_j P_ij into the RHS:
//      HALF (unshifted): r = -1/2 i k_j P_ij - nu k^2 u, the band forcing,
//        zero where ialias >= 2; the fields take the AB2 term of the old RHS
//        (u' = u - 1/2 dt rhs_old) and rhs_old takes r;
//      FINAL (shifted): the products shifted back, r = rhs_old - 1/2 i k_j
//        P_ij, zero where ialias >= 2, then u = project(u' + c dt r) with
//        c = 1 (Euler) or 3/2 (AB2), and rhs_old = r.
// Splitting the update this way needs no slot beyond the six products:
// the half RHS waits in rhs_old, whose old value is already folded into
// the fields.  The shifted physical velocity must exist before HALF changes
// the fields (the caller keeps it in three slots).
//
// The scalars (rhs_scalars.f90:144-271, 387-389, with their Euler/AB2
// update) go the same way, three products phi u_j per half: k_xprod3, the
// column passes, and k_zsc, whose shifted half waits in wrk(3+s) and whose
// unshifted half adds it, the diffusion, the mean-gradient source and the
// update.
#include <stdexcept>

#include "fft.hpp"
#include "host.hpp"
#include "passes.hpp"

namespace hit3d {

namespace {

struct Q6 {
    double* p[6];
};

// ------------------------------------------------------------------ x pass
// One row of N reals per M = N/2 point complex transform (z_m = x_2m +
// i x_2m+1), TPF threads per row, every row inside one wavefront: the
// exchanges synchronise per wavefront and the real-to-complex split pairs
// entry k (lane t, slot s) with M-k (lane (TPF-t) mod TPF, slot E-1-s, or
// (E-s) mod E for t = 0) by a lane shuffle.  Twiddles: a per-block LDS
// table of each thread's radix-8 twiddles where M is a power of 8 (M = 64,
// 512), else one table load per butterfly and its powers (TWREC).  As block
// barriers with twiddle loads per butterfly and the split through LDS, the
// 1024^3 passes ran one barrier per exchange across four rows.
template <int N>
struct XR {
    static constexpr int M = N / 2;
    static constexpr int E = 8;
    static constexpr int TPF = M / E;
    static constexpr int ROWS = (256 / TPF < N * N) ? 256 / TPF : N * N;
    static constexpr int THREADS = ROWS * TPF;
    static constexpr bool CACHE = M >= 64 && M == (1 << (3 * stage_index(M, 8)));
    using TC = std::conditional_t<CACHE, TwLds<double, M, E>, void>;
    static constexpr int TWSIZE = CACHE ? TwLds<double, M, E>::SIZE : 1;
    static_assert(TPF <= 64 && 64 % TPF == 0, "a row inside one wavefront");
};

// The per-block twiddle table (first row's threads fill it; every thread
// of the block must call this: it synchronises the block).
template <int N>
__device__ __forceinline__ void fill_row_twiddles(cpx<double>* twl, int t, int r,
                                                  const cpx<double>* __restrict__ tw) {
    using C = XR<N>;
    if constexpr (C::CACHE) {
        if (r == 0) {
            TwCache<double, C::M, C::E> tcv;
            fill_twiddles<C::M, C::E>(tcv, t, tw + C::M);
#pragma unroll
            for (int si = 0; si < TwCache<double, C::M, C::E>::NS; ++si)
#pragma unroll
                for (int m = 1; m < C::E; ++m) twl[(si * C::E + m) * C::TPF + t] = tcv.w[si][m];
        }
        __syncthreads();
    }
}

// e^{-2 pi i k/N} for k = t + s*TPF (TPF = N/16): e^{-2 pi i t/N} times the
// constant e^{-2 pi i s/16}, recomputed at every use (kept, the eight
// factors would hold 32 registers through the row)
__device__ __forceinline__ cpx<double> slot_twiddle(cpx<double> at, int s) {
    constexpr double c[4] = {1.0, 0.92387953251128675612818318939678828682,
                             0.70710678118654752440084436210484903928,
                             0.38268343236508977172845998403039886676};
    constexpr double sn[4] = {0.0, 0.38268343236508977172845998403039886676,
                              0.70710678118654752440084436210484903928,
                              0.92387953251128675612818318939678828682};
    cpx<double> a = at;
    asm volatile("" : "+v"(a.x), "+v"(a.y));
    const int s4 = s & 3;
    const cpx<double> r = s4 == 0 ? a : cmul(a, cpx<double>{c[s4], -sn[s4]});
    return s >= 4 ? mul_mi(r) : r;  // e^{-2 pi i (s+4)/16} = -i e^{-2 pi i s/16}
}

// Forward transform of the row z (x_2m + i x_2m+1) and its half spectrum
// X_k = (Z_k + conj Z_{M-k})/2 - i e^{-2 pi i k/N} (Z_k - conj Z_{M-k})/2,
// k = 0..M, stored at rowp.
template <int N>
__device__ __forceinline__ void row_r2c_store(cpx<double> (&z)[1][XR<N>::E], int t,
                                              cpx<double>* L, const cpx<double>* __restrict__ tw,
                                              const typename XR<N>::TC* tcp, cpx<double> at,
                                              cpx<double>* rowp) {
    using C = XR<N>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    lds_sync<true>();  // the previous transform's readers are done
    multi_fft<M, E, -1, double, 1, 1, padded_size<kSwizzle>(M), kSwizzle, true, true, double,
              typename C::TC>(z, t, L, tw + M, tcp);
    const int partner = (TPF - t) & (TPF - 1);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        // lane 0 pairs with itself (slot (E-s) mod E); selecting between the
        // shuffled value and its own keeps z in registers (a select between
        // two entries of z became a dynamically indexed array in scratch)
        const cpx<double> sh{__shfl(z[0][E - 1 - s].x, partner, TPF),
                             __shfl(z[0][E - 1 - s].y, partner, TPF)};
        const cpx<double> zm = conj(t == 0 ? z[0][(E - s) & (E - 1)] : sh);
        const cpx<double> zk = z[0][s];
        const cpx<double> o = mul_mi(cmul(slot_twiddle(at, s), zk - zm));
        st_g<kStream<N>>(&rowp[t + s * TPF], scale((zk + zm) + o, 0.5));
    }
    if (t == 0) st_g<kStream<N>>(&rowp[M], cpx<double>{z[0][0].x - z[0][0].y, 0.0});
}

// DEV: the five products of D_ij = u_i u_j - delta_ij w w (uu - ww, uv, uw,
// vv - ww, vw; D33 = 0) instead of the six u_i u_j: the divergence differs
// by the gradient of ww, which the final projection removes (the
// phase-shift mask keeps it a gradient: it zeroes whole wavevectors), and
// k_zps assembles k_j D_ij from them exactly as from the first five u_i u_j.
template <int N, bool COURANT, bool DEV = false>
__global__ __launch_bounds__(XR<N>::THREADS, 2) void k_xprod6(const double* u, const double* v,
                                                           const double* w, Q6 out, int64_t P,
                                                           const cpx<double>* __restrict__ tw,
                                                           double* __restrict__ blockmax) {
    using C = XR<N>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    __shared__ cpx<double> lds[C::ROWS][padded_size<kSwizzle>(M)];
    __shared__ cpx<double> twl[C::TWSIZE];
    __shared__ double redm[C::THREADS / 64 + 1];
    const int t = threadIdx.x % TPF;
    const int r = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + r;
    // a row of N reals as M complex pairs (x_2j, x_2j+1)
    cpx<double> a[3][E];
    const double* src[3] = {u, v, w};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const cpx<double>* rp = reinterpret_cast<const cpx<double>*>(src[c]) + row * P;
#pragma unroll
        for (int s = 0; s < E; ++s) a[c][s] = ld_g<kStream<N>>(&rp[t + s * TPF]);
    }
    fill_row_twiddles<N>(twl, t, r, tw);
    TwLds<double, M, E> twv{twl, t};
    const typename C::TC* tcp = nullptr;
    if constexpr (C::CACHE) tcp = &twv;
    const cpx<double> at = tw[N + t];
    double m = 0.0;
    if constexpr (COURANT) {  // rhs_velocity.f90:31-32
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const double p0 = (fabs(a[0][s].x) + fabs(a[1][s].x)) + fabs(a[2][s].x);
            const double p1 = (fabs(a[0][s].y) + fabs(a[1][s].y)) + fabs(a[2][s].y);
            m = p0 > m ? p0 : m;
            m = p1 > m ? p1 : m;
        }
    }
    // products uu, uv, uw, vv, vw, ww (rhs_velocity.f90:160-170); DEV: uu - ww,
    // uv, uw, vv - ww, vw
    constexpr int NP = DEV ? 5 : 6;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int i = p < 3 ? 0 : p < 5 ? 1 : 2;
        const int j = p < 3 ? p : p < 5 ? p - 2 : 2;
        cpx<double> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            z[0][s] = cpx<double>{a[i][s].x * a[j][s].x, a[i][s].y * a[j][s].y};
            if (DEV && (p == 0 || p == 3)) {
                // ww recomputed for vv - ww (kept from uu - ww, it held
                // registers across three transforms and spilled)
                cpx<double> c = a[2][s];
                if (p == 3) asm volatile("" : "+v"(c.x), "+v"(c.y));
                z[0][s] = cpx<double>{z[0][s].x - c.x * c.x, z[0][s].y - c.y * c.y};
            }
        }
        double* op = out.p[0];
        if (p == 1) op = out.p[1];
        if (p == 2) op = out.p[2];
        if (p == 3) op = out.p[3];
        if (p == 4) op = out.p[4];
        if (p == 5) op = out.p[5];
        row_r2c_store<N>(z, t, lds[r], tw, tcp, at, reinterpret_cast<cpx<double>*>(op) + row * P);
    }
    if constexpr (COURANT) {
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
}

// ------------------------------------------------------------------ z pass
// Columns (kx = 0..N/2, ky) of the six products, B columns per block (4 at
// N = 1024, whose two halves of a 128-byte segment run on one XCD).
template <int N>
struct ZC {
    static constexpr int E = 8;
    static constexpr int TPF = N / E;
    static constexpr int B = N >= 1024 ? 4 : 8;
    static constexpr int THREADS = B * TPF;
};

// Paired half-segment blocks (B = 4 at 1024^3) load through the cache, so
// that the partner block finds the other half of each line in L2 (streaming
// loads fetched the lines again: the 2/3-rule z/update kernel moved 1.30x its
// bytes that way).
template <int N>
inline constexpr bool kZLoadStream = kStream<N> && ZC<N>::B * 16 >= 128;

template <int N>
__device__ __forceinline__ void zc_coords(int& bx, int& by) {
    if constexpr (ZC<N>::B * 16 >= 128) {
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

struct PsUpd {
    double* F[3];
    double* R[3];
    const double* eshell;  // band energy per shell 1..kfmax (device)
    double* energy_out;
    const double* dtp;
    const double* tab;     // cos, sin of (kx+ky+kz) dx/2 at 2(kx+ky+kz+N)
    double nu, famp;
    int kfmax, forced, first;
};

template <int N, bool FINAL, bool DEV = false>
__global__ __launch_bounds__(ZC<N>::THREADS) void k_zps(Q6 Q, int64_t P, PsUpd u,
                                                        const cpx<double>* __restrict__ tw) {
    using C = ZC<N>;
    constexpr int E = C::E, B = C::B, TPF = C::TPF;
    constexpr int NC = N / 2 + 1;
    __shared__ cpx<double> lds[N * B];
    int bx, by;
    zc_coords<N>(bx, by);
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = bx * B + c;
    const bool valid = kx < NC;
    const int jy = by;
    const int ky = wavenum(jy, N);
    const double kxd = kx, kyd = ky;
    // t_i = k_j P_ij, rhs_velocity.f90:190-196 (shifted back first in FINAL,
    // :282-398)
    cpx<double> acc[3][E];
#pragma unroll
    for (int s = 0; s < E; ++s) acc[0][s] = acc[1][s] = acc[2][s] = cpx<double>{0.0, 0.0};
    constexpr int NP = DEV ? 5 : 6;
#pragma unroll 1
    for (int p = 0; p < NP; ++p) {
        const double* qp = Q.p[0];
        if (p == 1) qp = Q.p[1];
        if (p == 2) qp = Q.p[2];
        if (p == 3) qp = Q.p[3];
        if (p == 4) qp = Q.p[4];
        if (p == 5) qp = Q.p[5];
        // layout after the x and y passes: (z, ky, kx)
        const cpx<double>* src = reinterpret_cast<const cpx<double>*>(qp) + (int64_t)jy * P + kx;
        cpx<double> v[E];
#pragma unroll
        for (int s = 0; s < E; ++s)
            v[s] = valid ? ld_g<kZLoadStream<N>>(&src[(int64_t)(t + s * TPF) * N * P])
                         : cpx<double>{0.0, 0.0};
        const cpx<double>* twp = tw + N;
        asm volatile("" : "+s"(twp));
        block_fft<N, E, -1, double, B, 0, true, double>(v, t, lds + c, twp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int e = t + s * TPF;
            cpx<double> x = v[s];
            if constexpr (FINAL) {  // shift back by exp(-i theta)
                const int ti = kx + ky + wavenum(e, N) + N;
                const double cs = u.tab[2 * ti], sn = u.tab[2 * ti + 1];
                x = cpx<double>{x.x * cs + x.y * sn, x.y * cs - x.x * sn};
            }
            const double kzd = wavenum(e, N);
            // p: 0 uu, 1 uv, 2 uw, 3 vv, 4 vw, 5 ww (DEV: 0 uu - ww, 3 vv - ww,
            // no 5)
            if (p == 0) acc[0][s] = scale(x, kxd);
            if (p == 1) {
                acc[0][s] = acc[0][s] + scale(x, kyd);
                acc[1][s] = scale(x, kxd);
            }
            if (p == 2) {
                acc[0][s] = acc[0][s] + scale(x, kzd);
                acc[2][s] = scale(x, kxd);
            }
            if (p == 3) acc[1][s] = acc[1][s] + scale(x, kyd);
            if (p == 4) {
                acc[1][s] = acc[1][s] + scale(x, kzd);
                acc[2][s] = acc[2][s] + scale(x, kyd);
            }
            if (p == 5) acc[2][s] = acc[2][s] + scale(x, kzd);
        }
    }
    double fac = 0.0;
    if (!FINAL && u.forced) {  // m_force.f90:153-161
        double en = 0.0;
        for (int i = 0; i < u.kfmax; ++i) en += u.eshell[i];
        fac = u.famp / en;
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *u.energy_out = en;
    }
    if (!valid) return;
    const double dt = *u.dtp;
    // In chunks of H slots, every fields / rhs_old load of the chunk is
    // issued before its first store: the compiler cannot tell the six
    // arrays apart, and a store between loads would cost one memory round
    // trip per load.  (FINAL keeps more values live: smaller chunks.)
    constexpr int H = FINAL ? 2 : 4;
#pragma unroll
    for (int h = 0; h < E / H; ++h) {
        cpx<double> f[3][H], ro[3][H];
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int e = t + (h * H + q) * TPF;
            const int64_t off = ((int64_t)jy * N + e) * P + kx;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                f[cc][q] = ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.F[cc])[off]);
                ro[cc][q] = ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.R[cc])[off]);
            }
        }
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = h * H + q;
            const int e = t + s * TPF;
            const int kz = wavenum(e, N);
            const double kzd = kz;
            const int64_t off = ((int64_t)jy * N + e) * P + kx;
            const bool zero = alias_count(kx, ky, kz, N) >= 2;  // rhs_velocity.f90:215,387
            if constexpr (!FINAL) {
                const double lap = -u.nu * (kxd * kxd + kyd * kyd + kzd * kzd);
                const bool fnode = u.forced && kx < N / 2 && [&] {
                    const int sh = shell_of(kx * kx + ky * ky + kz * kz);
                    return sh > 0 && sh <= u.kfmax;
                }();
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const cpx<double> fc = f[cc][q];
                    // r = 1/2 i (-t) + lap u  (rhs_velocity.f90:219-227)
                    cpx<double> r = {0.5 * acc[cc][s].y + lap * fc.x,
                                     -0.5 * acc[cc][s].x + lap * fc.y};
                    if (zero) r = {0.0, 0.0};
                    if (fnode) r = {r.x + fac * fc.x, r.y + fac * fc.y};
                    const cpx<double> fo =
                        u.first ? fc : cpx<double>{fc.x - 0.5 * dt * ro[cc][q].x,
                                                   fc.y - 0.5 * dt * ro[cc][q].y};
                    st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[cc])[off], fo);
                    st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.R[cc])[off], r);
                }
            } else {
                const double cdt = (u.first ? 1.0 : 1.5) * dt;
                cpx<double> fn[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    // r += -1/2 i t  (rhs_velocity.f90:282-398), masked (:400-404)
                    cpx<double> r = {ro[cc][q].x + 0.5 * acc[cc][s].y,
                                     ro[cc][q].y - 0.5 * acc[cc][s].x};
                    if (zero) r = {0.0, 0.0};
                    fn[cc] = {f[cc][q].x + cdt * r.x, f[cc][q].y + cdt * r.y};
                    st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.R[cc])[off], r);
                }
                // projection, pressure.f90:21-43
                const double k2 = kxd * kxd + kyd * kyd + kzd * kzd;
                const double dre = kxd * fn[0].x + kyd * fn[1].x + kzd * fn[2].x;
                const double dim = kxd * fn[0].y + kyd * fn[1].y + kzd * fn[2].y;
                const double lapl = k2 == 0.0 ? 9e20 : k2;
                const double pr = dre / lapl, pi = dim / lapl;
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[0])[off],
                                 cpx<double>{fn[0].x - pr * kxd, fn[0].y - pi * kxd});
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[1])[off],
                                 cpx<double>{fn[1].x - pr * kyd, fn[1].y - pi * kyd});
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F[2])[off],
                                 cpx<double>{fn[2].x - pr * kzd, fn[2].y - pi * kzd});
            }
        }
    }
}

// ------------------------------------------------------------ scalars
// x pass: a physical scalar row phi and the velocity rows -> the half
// spectra of phi u, phi v, phi w (out may alias phi: each row is read
// before it is written).
template <int N>
__global__ __launch_bounds__(XR<N>::THREADS) void k_xprod3(const double* phi, const double* u,
                                                           const double* v, const double* w,
                                                           Q6 out, int64_t P,
                                                           const cpx<double>* __restrict__ tw) {
    using C = XR<N>;
    constexpr int M = C::M, E = C::E, TPF = C::TPF;
    __shared__ cpx<double> lds[C::ROWS][padded_size<kSwizzle>(M)];
    __shared__ cpx<double> twl[C::TWSIZE];
    const int t = threadIdx.x % TPF;
    const int r = threadIdx.x / TPF;
    const int64_t row = (int64_t)blockIdx.x * C::ROWS + r;
    cpx<double> ph[E];
    {
        const cpx<double>* rp = reinterpret_cast<const cpx<double>*>(phi) + row * P;
#pragma unroll
        for (int s = 0; s < E; ++s) ph[s] = ld_g<kStream<N>>(&rp[t + s * TPF]);
    }
    fill_row_twiddles<N>(twl, t, r, tw);
    TwLds<double, M, E> twv{twl, t};
    const typename C::TC* tcp = nullptr;
    if constexpr (C::CACHE) tcp = &twv;
    const cpx<double> at = tw[N + t];
    const double* vel[3] = {u, v, w};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const cpx<double>* vp = reinterpret_cast<const cpx<double>*>(vel[p]) + row * P;
        cpx<double> z[1][E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const cpx<double> b = ld_g<kStream<N>>(&vp[t + s * TPF]);
            z[0][s] = cpx<double>{ph[s].x * b.x, ph[s].y * b.y};
        }
        double* op = out.p[0];
        if (p == 1) op = out.p[1];
        if (p == 2) op = out.p[2];
        row_r2c_store<N>(z, t, lds[r], tw, tcp, at, reinterpret_cast<cpx<double>*>(op) + row * P);
    }
}

struct ScUpd {
    double* F;          // the scalar field
    double* R;          // its rhs_old
    double* H;          // the RHS of the shifted half (HALF writes, FINAL reads)
    const double* U1;   // the mean-gradient source -u (type 0) or null
    const double* tab;
    const double* dtp;
    double pe;
    int first;
};

// z pass of the scalar products phi u_j (rhs_scalars.f90:144-262):
// HALF (shifted): the products shifted back, H = -1/2 i k_j Q_j, zero where
// ialias >= 2 (:200); FINAL (unshifted): r = H - 1/2 i k_j Q_j - pe k^2 phi
// where ialias < 2 (:240-262), minus u for a type-0 scalar (:387-389), then
// the Euler/AB2 update of the scalar (main.f90:137-146): phi += dt r (or
// dt (3/2 r - 1/2 rhs_old)), rhs_old = r.
template <int N, bool FINAL>
__global__ __launch_bounds__(ZC<N>::THREADS) void k_zsc(Q6 Q, int64_t P, ScUpd u,
                                                        const cpx<double>* __restrict__ tw) {
    using C = ZC<N>;
    constexpr int E = C::E, B = C::B, TPF = C::TPF;
    constexpr int NC = N / 2 + 1;
    __shared__ cpx<double> lds[N * B];
    int bx, by;
    zc_coords<N>(bx, by);
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const int kx = bx * B + c;
    const bool valid = kx < NC;
    const int jy = by;
    const int ky = wavenum(jy, N);
    const double kxd = kx, kyd = ky;
    cpx<double> acc[E];
#pragma unroll
    for (int s = 0; s < E; ++s) acc[s] = cpx<double>{0.0, 0.0};
#pragma unroll 1
    for (int p = 0; p < 3; ++p) {
        const double* qp = Q.p[0];
        if (p == 1) qp = Q.p[1];
        if (p == 2) qp = Q.p[2];
        const cpx<double>* src = reinterpret_cast<const cpx<double>*>(qp) + (int64_t)jy * P + kx;
        cpx<double> v[E];
#pragma unroll
        for (int s = 0; s < E; ++s)
            v[s] = valid ? ld_g<kZLoadStream<N>>(&src[(int64_t)(t + s * TPF) * N * P])
                         : cpx<double>{0.0, 0.0};
        const cpx<double>* twp = tw + N;
        asm volatile("" : "+s"(twp));
        block_fft<N, E, -1, double, B, 0, true, double>(v, t, lds + c, twp);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int e = t + s * TPF;
            cpx<double> x = v[s];
            if constexpr (!FINAL) {  // shift back by exp(-i theta)
                const int ti = kx + ky + wavenum(e, N) + N;
                const double cs = u.tab[2 * ti], sn = u.tab[2 * ti + 1];
                x = cpx<double>{x.x * cs + x.y * sn, x.y * cs - x.x * sn};
            }
            const double k = p == 0 ? kxd : p == 1 ? kyd : (double)wavenum(e, N);
            acc[s] = acc[s] + scale(x, k);
        }
    }
    if (!valid) return;
    const double dt = *u.dtp;
    constexpr int H = 4;  // loads of a chunk before its stores, as in k_zps
#pragma unroll
    for (int h = 0; h < E / H; ++h) {
        cpx<double> f[H], hh[H], ro[H], s1[H];
        if constexpr (FINAL) {
#pragma unroll
            for (int q = 0; q < H; ++q) {
                const int e = t + (h * H + q) * TPF;
                const int64_t off = ((int64_t)jy * N + e) * P + kx;
                f[q] = ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.F)[off]);
                hh[q] = ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.H)[off]);
                ro[q] = u.first ? cpx<double>{0.0, 0.0}
                                : ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.R)[off]);
                s1[q] = u.U1 ? ld_g<kZLoadStream<N>>(&reinterpret_cast<const cpx<double>*>(u.U1)[off])
                             : cpx<double>{0.0, 0.0};
            }
        }
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = h * H + q;
            const int e = t + s * TPF;
            const int kz = wavenum(e, N);
            const double kzd = kz;
            const int64_t off = ((int64_t)jy * N + e) * P + kx;
            const bool zero = alias_count(kx, ky, kz, N) >= 2;
            // -1/2 i t
            const cpx<double> half = {0.5 * acc[s].y, -0.5 * acc[s].x};
            if constexpr (!FINAL) {
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.H)[off],
                                 zero ? cpx<double>{0.0, 0.0} : half);
            } else {
                const double k2 = kxd * kxd + kyd * kyd + kzd * kzd;
                cpx<double> r = hh[q];
                if (!zero)
                    r = {(hh[q].x + half.x) - u.pe * k2 * f[q].x,
                         (hh[q].y + half.y) - u.pe * k2 * f[q].y};
                r = {r.x - s1[q].x, r.y - s1[q].y};
                const cpx<double> fn =
                    u.first ? cpx<double>{f[q].x + dt * r.x, f[q].y + dt * r.y}
                            : cpx<double>{f[q].x + dt * (1.5 * r.x - 0.5 * ro[q].x),
                                          f[q].y + dt * (1.5 * r.y - 0.5 * ro[q].y)};
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.R)[off], r);
                st_g<kStream<N>>(&reinterpret_cast<cpx<double>*>(u.F)[off], fn);
            }
        }
    }
}

template <typename F>
void dispatch_ps(int n, F&& f) {
    switch (n) {
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        case 128: f(std::integral_constant<int, 128>{}); break;
        case 256: f(std::integral_constant<int, 256>{}); break;
        case 512: f(std::integral_constant<int, 512>{}); break;
        case 1024: f(std::integral_constant<int, 1024>{}); break;
        default: throw std::runtime_error("fused phase-shift step: unsupported grid size");
    }
}

template <int N, bool DEV>
void half(Lib& L, const double* const uvw[3], bool final_half, bool first, const double* dtp) {
    using X = XR<N>;
    using Z = ZC<N>;
    const int64_t P = L.g.P;
    Q6 q;
    for (int i = 0; i < 6; ++i) q.p[i] = L.W[i + 1];
    {
        Timed tm(TC_FFT_X);
        const int64_t nb = (int64_t)N * N / X::ROWS;  // <= n*n, the size of dmax
        if (!final_half)
            hipLaunchKernelGGL((k_xprod6<N, true, DEV>), dim3(nb), dim3(X::THREADS), 0, L.st,
                               uvw[0], uvw[1], uvw[2], q, P, L.tw64, L.dmax);
        else
            hipLaunchKernelGGL((k_xprod6<N, false, DEV>), dim3(nb), dim3(X::THREADS), 0, L.st,
                               uvw[0], uvw[1], uvw[2], q, P, L.tw64, L.dmax);
        if (!final_half) launch_max_fold(L.dmax, nb, L.dscal, L.st);
    }
    for (int i = 1; i <= (DEV ? 5 : 6); ++i) pass_y(L.W[i], -1);
    PsUpd u;
    for (int i = 0; i < 3; ++i) {
        u.F[i] = L.F[i + 1];
        u.R[i] = L.R[i + 1];
    }
    u.eshell = L.dshell;
    u.energy_out = L.dscal + 1;
    u.dtp = dtp;
    u.tab = L.ps_tab;
    u.nu = L.nu;
    u.famp = L.famp;
    u.kfmax = L.kfmax;
    u.forced = L.flow_type == 1 ? 1 : 0;
    u.first = first ? 1 : 0;
    int ncb = (N / 2 + 1 + Z::B - 1) / Z::B;
    if (Z::B * 16 < 128) ncb += ncb & 1;
    Timed tm(TC_FFT_Z);
    if (!final_half)
        hipLaunchKernelGGL((k_zps<N, false, DEV>), dim3(ncb, N), dim3(Z::THREADS), 0, L.st, q, P,
                           u, L.tw64);
    else
        hipLaunchKernelGGL((k_zps<N, true, DEV>), dim3(ncb, N), dim3(Z::THREADS), 0, L.st, q, P,
                           u, L.tw64);
}

template <int N>
void scalar_half(Lib& L, const double* phi, const double* const vel[3], double* const q[3],
                 bool final_half, const ScUpd& u) {
    using X = XR<N>;
    using Z = ZC<N>;
    const int64_t P = L.g.P;
    Q6 qq{};
    for (int i = 0; i < 3; ++i) qq.p[i] = q[i];
    {
        Timed tm(TC_FFT_X);
        const int64_t nb = (int64_t)N * N / X::ROWS;
        hipLaunchKernelGGL((k_xprod3<N>), dim3(nb), dim3(X::THREADS), 0, L.st, phi, vel[0], vel[1],
                           vel[2], qq, P, L.tw64);
    }
    for (int i = 0; i < 3; ++i) pass_y(q[i], -1);
    int ncb = (N / 2 + 1 + Z::B - 1) / Z::B;
    if (Z::B * 16 < 128) ncb += ncb & 1;
    Timed tm(TC_FFT_Z);
    if (!final_half)
        hipLaunchKernelGGL((k_zsc<N, false>), dim3(ncb, N), dim3(Z::THREADS), 0, L.st, qq, P, u,
                           L.tw64);
    else
        hipLaunchKernelGGL((k_zsc<N, true>), dim3(ncb, N), dim3(Z::THREADS), 0, L.st, qq, P, u,
                           L.tw64);
}

}  // namespace

bool fused_ps_scalars_supported() {
    Lib& L = lib();
    if (!fused_ps_supported() || L.ns < 1 || L.nw < 3 + L.ns + 2) return false;
    for (int s = 0; s < L.ns; ++s)
        if (L.stype[s] > 300) return false;  // the type-311 reaction takes the unfused path
    return true;
}

// rhs_scalars (phase shift, rhs_scalars.f90:144-271, 387-389) and the
// scalars' Euler/AB2 update in fused passes.  Leaves the shifted physical
// velocity in lib().keep (three slots) and the physical velocity in
// wrk(1:3), as rhs_scalars does (:229-232), for the velocity RHS.
void fused_ps_scalars(bool first, const double* dtp) {
    Lib& L = lib();
    auto& F = L.F;
    auto& W = L.W;
    if (L.keep.size() != 3) throw std::runtime_error("internal: shifted-velocity slots missing");
    for (int c = 0; c < 3; ++c) fft_inverse_from(F[1 + c], L.keep[c], +1);
    dispatch_ps(L.n, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        const double* vs[3] = {L.keep[0], L.keep[1], L.keep[2]};
        // shifted halves into wrk(3+s), products in wrk(1:3)
        for (int s = 1; s <= L.ns; ++s) {
            fft_inverse_from(F[3 + s], W[0], +1);
            ScUpd u{};
            u.H = W[3 + s];
            u.tab = L.ps_tab;
            u.dtp = dtp;
            double* q[3] = {W[1], W[2], W[3]};
            scalar_half<N>(L, W[0], vs, q, false, u);
        }
        // the physical velocity in wrk(1:3), then the unshifted halves,
        // products in wrk(n1), wrk(n2) and wrk(0)
        const double* src[3] = {F[1], F[2], F[3]};
        double** dst[3] = {&W[1], &W[2], &W[3]};
        for (int c = 0; c < 3; ++c) fft_inverse_from(src[c], *dst[c], 0);
        const double* vu[3] = {W[1], W[2], W[3]};
        const int n1 = L.ns + 4, n2 = L.ns + 5;
        for (int s = 1; s <= L.ns; ++s) {
            fft_inverse_from(F[3 + s], W[0], 0);
            ScUpd u{};
            u.F = F[3 + s];
            u.R = L.R[3 + s];
            u.H = W[3 + s];
            u.U1 = L.stype[s - 1] == 0 ? F[1] : nullptr;
            u.tab = L.ps_tab;
            u.dtp = dtp;
            u.pe = L.pe[s - 1];
            u.first = first ? 1 : 0;
            double* q[3] = {W[n1], W[n2], W[0]};
            scalar_half<N>(L, W[0], vu, q, true, u);
        }
    });
}

bool fused_ps_supported() {
    Lib& L = lib();
    return L.dealias == 1 && L.nles == 0 && L.les_model == 0 && L.n >= 32;
}

// The velocity part of a phase-shift step after rhs_scalars: wrk(1:3) holds
// the physical velocity and `shifted` the physical velocity shifted by
// exp(i theta) (three slots, not wrk(1:6)).  Leaves fields(1:3) updated and
// projected, rhs_old(1:3) the RHS with forcing, the Courant max in
// dscal[0] and the band energy in dscal[1]; wrk(1:6) is scratch.
void fused_ps_velocity(double* const shifted[3], bool first, const double* dtp) {
    Lib& L = lib();
    if (L.flow_type == 1) {
        if (L.kfmax < 1 || L.kfmax > L.kmax) throw std::runtime_error("force_velocity: kfmax range");
        Timed t(TC_REDUCE);
        launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kfmax, L.dshell, L.st);
    }
    dispatch_ps(L.n, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        const double* un[3] = {L.W[1], L.W[2], L.W[3]};
        const double* sh[3] = {shifted[0], shifted[1], shifted[2]};
        if (L.nlform == 2) {
            half<N, true>(L, un, false, first, dtp);
            half<N, true>(L, sh, true, first, dtp);
        } else {
            half<N, false>(L, un, false, first, dtp);
            half<N, false>(L, sh, true, first, dtp);
        }
    });
}

}  // namespace hit3d
