// Pointwise kernels and deterministic FP64 reductions of the time step
// (unfused forms, one kernel per reference loop nest).
#include <stdexcept>

#include "spectral.hpp"

namespace hit3d {

namespace {

constexpr int TB = 256;

__host__ int grid_for(int64_t total) {
    int64_t b = (total + TB - 1) / TB;
    if (b > 65536) b = 65536;
    return (int)(b < 1 ? 1 : b);
}

struct SIdx {
    int kx, kyv, kzv, j, k;
    int64_t off;
};

// Decode a linear index over the valid spectral entries (kx <= n/2).
__device__ __forceinline__ SIdx sdecode(const Geo& g, int64_t idx) {
    const int H = g.n / 2 + 1;
    SIdx s;
    s.kx = (int)(idx % H);
    const int64_t r = idx / H;
    s.j = (int)(r % g.n);
    s.k = (int)(r / g.n);
    s.kzv = wavenum(s.j, g.n);
    s.kyv = wavenum(s.k, g.n);
    s.off = r * g.P + s.kx;
    return s;
}

__device__ __forceinline__ int64_t nspec(const Geo& g) { return (int64_t)g.n * g.n * (g.n / 2 + 1); }

__device__ __forceinline__ const cpx<double>* C(const double* p) {
    return reinterpret_cast<const cpx<double>*>(p);
}
__device__ __forceinline__ cpx<double>* C(double* p) { return reinterpret_cast<cpx<double>*>(p); }
// access to entry off of a slot in the per-step passes that touch each
// entry once: streaming from 64^3 up
__device__ __forceinline__ cpx<double> ldc(const Geo& g, const double* p, int64_t off) {
    return streams(g.n) ? ld_nt(&C(p)[off]) : C(p)[off];
}
__device__ __forceinline__ void stc(const Geo& g, double* p, int64_t off, cpx<double> v) {
    if (streams(g.n)) st_nt(&C(p)[off], v);
    else C(p)[off] = v;
}
__device__ __forceinline__ void std_nt(const Geo& g, double* p, double v) {
    if (streams(g.n)) __builtin_nontemporal_store(v, p);
    else *p = v;
}

#define GRID_STRIDE(i, total) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (total); i += (int64_t)gridDim.x * blockDim.x)

// fixed-order block sum (wave shuffles, then wave partials in order)
__device__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) tot += sh[i];
    }
    return tot;  // valid in thread 0
}

__device__ __forceinline__ int isqrt_floor(int64_t x) {
    int64_t s = (int64_t)sqrt((double)x);
    while (s * s > x) --s;
    while ((s + 1) * (s + 1) <= x) ++s;
    return (int)s;
}

// ---------------------------------------------------------------------------

__global__ void k_derivative(Geo g, const double* in, double* out, int axis) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const double k = axis == 1 ? (double)s.kx : (axis == 2 ? (double)s.kyv : (double)s.kzv);
        cpx<double> a = C(in)[s.off];
        C(out)[s.off] = cpx<double>{-(k * a.y), k * a.x};
    }
}

__global__ void k_truncate(Geo g, double* slot, int kmax) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        if (s.kx > kmax || abs(s.kyv) > kmax || abs(s.kzv) > kmax) C(slot)[s.off] = {0.0, 0.0};
    }
}

__global__ void k_pressure(Geo g, double* u, double* v, double* w) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        cpx<double> a = C(u)[s.off], b = C(v)[s.off], c = C(w)[s.off];
        const double dre = kx * a.x + ky * b.x + kz * c.x;
        const double dim = kx * a.y + ky * b.y + kz * c.y;
        double lap = kx * kx + ky * ky + kz * kz;
        if (lap == 0.0) lap = 9e20;
        const double pr = dre / lap, pi = dim / lap;
        C(u)[s.off] = {a.x - pr * kx, a.y - pi * kx};
        C(v)[s.off] = {b.x - pr * ky, b.y - pi * ky};
        C(w)[s.off] = {c.x - pr * kz, c.y - pi * kz};
    }
}

__global__ void k_advance(Geo g, double* f, const double* r, double* ro, double dt,
                          const double* dtp, int first) {
    if (dtp) dt = *dtp;
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        cpx<double> a = ldc(g, f, s.off), b = ldc(g, r, s.off);
        if (first) {
            stc(g, ro, s.off, b);
            stc(g, f, s.off, {a.x + dt * b.x, a.y + dt * b.y});
        } else {
            cpx<double> o = ldc(g, ro, s.off);
            stc(g, f, s.off, {a.x + dt * (1.5 * b.x - 0.5 * o.x), a.y + dt * (1.5 * b.y - 0.5 * o.y)});
            stc(g, ro, s.off, b);
        }
    }
}

// zero fill by a kernel (also inside captured graphs, where memset nodes
// are avoided)
__global__ void k_fill_zero(double* p, int64_t count) {
    GRID_STRIDE(i, count) p[i] = 0.0;
}
// Plain streaming copy of n16 16-byte words (one nontemporal dwordx4 load
// and store per thread and iteration, grid-stride): the HBM rate an ideal
// pass that reads and writes each byte once reaches on this part, the
// yardstick the step's kernels are set beside.  Unrolling four words per
// thread measured slower (4.2-5.6 against 5.0-6.0 TB/s over 1024-65536
// blocks on a 1.1 GB slot).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(const u32x4* __restrict__ src,
                                                     u32x4* __restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ void k_narrow(const double* __restrict__ src, float* __restrict__ dst, int64_t count) {
    GRID_STRIDE(i, count) dst[i] = (float)src[i];
}
__global__ void k_widen(const float* __restrict__ src, double* __restrict__ dst, int64_t count) {
    GRID_STRIDE(i, count) dst[i] = (double)src[i];
}

struct Ptr6 {
    double* p[6];
};
struct CPtr6 {
    const double* p[6];
};
struct CPtr3 {
    const double* p[3];
};
struct Ptr3 {
    double* p[3];
};

// AB2/Euler update of the three velocity components followed by the
// projection, per entry (k_advance then k_pressure without the round trip)
__global__ void k_advance_project(Geo g, Ptr3 f, CPtr3 r, Ptr3 ro, double dt, const double* dtp,
                                  int first) {
    if (dtp) dt = *dtp;
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        // every load before the first store (the arrays may alias as far as
        // the compiler knows)
        cpx<double> u[3], b[3], o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u[c] = ldc(g, f.p[c], s.off);
            b[c] = ldc(g, r.p[c], s.off);
            if (!first) o[c] = ldc(g, ro.p[c], s.off);
        }
        cpx<double> a[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (first)
                a[c] = {u[c].x + dt * b[c].x, u[c].y + dt * b[c].y};
            else
                a[c] = {u[c].x + dt * (1.5 * b[c].x - 0.5 * o[c].x),
                        u[c].y + dt * (1.5 * b[c].y - 0.5 * o[c].y)};
            stc(g, ro.p[c], s.off, b[c]);
        }
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        const double dre = kx * a[0].x + ky * a[1].x + kz * a[2].x;
        const double dim = kx * a[0].y + ky * a[1].y + kz * a[2].y;
        double lap = kx * kx + ky * ky + kz * kz;
        if (lap == 0.0) lap = 9e20;
        const double pr = dre / lap, pi = dim / lap;
        stc(g, f.p[0], s.off, {a[0].x - pr * kx, a[0].y - pi * kx});
        stc(g, f.p[1], s.off, {a[1].x - pr * ky, a[1].y - pi * ky});
        stc(g, f.p[2], s.off, {a[2].x - pr * kz, a[2].y - pi * kz});
    }
}


__global__ void k_products(Geo g, const double* u, const double* v, const double* w, Ptr6 out,
                           double* blockmax) {
    __shared__ double sh[TB / 64];
    const int64_t total = (int64_t)g.n * g.n * g.n;
    double m = 0.0;
    GRID_STRIDE(i, total) {
        const int64_t x = i % g.n, row = i / g.n;
        const int64_t o = row * 2 * g.P + x;
        const double a = u[o], b = v[o], c = w[o];
        const double t = (fabs(a) + fabs(b)) + fabs(c);
        m = t > m ? t : m;
        const double p0 = a * a, p1 = a * b, p2 = a * c, p3 = b * b, p4 = b * c, p5 = c * c;
        std_nt(g, &out.p[0][o], p0);
        std_nt(g, &out.p[1][o], p1);
        std_nt(g, &out.p[2][o], p2);
        std_nt(g, &out.p[3][o], p3);
        std_nt(g, &out.p[4][o], p4);
        std_nt(g, &out.p[5][o], p5);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) mm = sh[i] > mm ? sh[i] : mm;
        blockmax[blockIdx.x] = mm;
    }
}

// out[0] = max of count non-negative partials (one block of 1024 threads)
__global__ __launch_bounds__(1024) void k_max_fold(const double* part, int64_t count, double* out) {
    __shared__ double sh[16];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += 1024) m = part[i] > m ? part[i] : m;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double q = __shfl_down(m, o, 64);
        m = q > m ? q : m;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int i = 0; i < 16; ++i) mm = sh[i] > mm ? sh[i] : mm;
        out[0] = mm;
    }
}

__global__ void k_rhs_assemble(Geo g, CPtr6 P, CPtr3 U, Ptr3 out, double nu, int phase_shift) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        const int ia = alias_count(s.kx, s.kyv, s.kzv, g.n);
        if (ia > (phase_shift ? 1 : 0)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) stc(g, out.p[c], s.off, {0.0, 0.0});
            continue;
        }
        cpx<double> p[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) p[q] = ldc(g, P.p[q], s.off);
        cpx<double> u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = ldc(g, U.p[c], s.off);
        const double lap = -nu * (kx * kx + ky * ky + kz * kz);
        const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        const double h = phase_shift ? 0.5 : 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cpx<double> a = p[idx[c][0]], b = p[idx[c][1]], d = p[idx[c][2]];
            const double tre = -(kx * a.x + ky * b.x + kz * d.x);
            const double tim = -(kx * a.y + ky * b.y + kz * d.y);
            stc(g, out.p[c], s.off, {-(h * tim) + lap * u[c].x, h * tre + lap * u[c].y});
        }
    }
}

__global__ void k_phys_mul(Geo g, const double* a, const double* b, double* out) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) {
        const int64_t x = i % g.n, row = i / g.n;
        const int64_t o = row * 2 * g.P + x;
        out[o] = a[o] * b[o];
    }
}

// sub (may be null): subtracted afterwards at every entry (the mean-gradient
// source, rhs_scalars.f90:387-389)
__global__ void k_scalar_rhs(Geo g, CPtr3 Q, const double* F, double* out, double pe,
                             const double* sub) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        cpx<double> r = {0.0, 0.0};
        if (alias_count(s.kx, s.kyv, s.kzv, g.n) == 0) {
            const double kx = s.kx, ky = s.kyv, kz = s.kzv;
            const cpx<double> a = ldc(g, Q.p[0], s.off), b = ldc(g, Q.p[1], s.off), c = ldc(g, Q.p[2], s.off);
            const cpx<double> f = ldc(g, F, s.off);
            const double r1 = kx * a.y + ky * b.y + kz * c.y;
            const double r2 = kx * a.x + ky * b.x + kz * c.x;
            const double k2 = kx * kx + ky * ky + kz * kz;
            r = {r1 - pe * k2 * f.x, -r2 - pe * k2 * f.y};
        }
        if (sub) {
            const cpx<double> u = ldc(g, sub, s.off);
            r = {r.x - u.x, r.y - u.y};
        }
        stc(g, out, s.off, r);
    }
}

__global__ void k_scalar_rhs_ps(Geo g, CPtr3 Q, const double* F, double* out, double pe, int acc,
                                const double* sub) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const bool keep = alias_count(s.kx, s.kyv, s.kzv, g.n) <= 1;
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        const cpx<double> a = ldc(g, Q.p[0], s.off), b = ldc(g, Q.p[1], s.off), c = ldc(g, Q.p[2], s.off);
        const double r1 = kx * a.y + ky * b.y + kz * c.y;
        const double r2 = kx * a.x + ky * b.x + kz * c.x;
        if (!acc) {
            stc(g, out, s.off, keep ? cpx<double>{0.5 * r1, -0.5 * r2} : cpx<double>{0.0, 0.0});
        } else {
            cpx<double> o = ldc(g, out, s.off);
            if (keep) {
                const cpx<double> f = ldc(g, F, s.off);
                const double k2 = kx * kx + ky * ky + kz * kz;
                o = {o.x + 0.5 * r1 - pe * k2 * f.x, o.y - 0.5 * r2 - pe * k2 * f.y};
            }
            if (sub) {
                const cpx<double> u = ldc(g, sub, s.off);
                o = {o.x - u.x, o.y - u.y};
            }
            stc(g, out, s.off, o);
        }
    }
}

__global__ void k_shell_sums(Geo g, const double* u, const double* v, const double* w, int s1,
                             double fac, double* out) {
    __shared__ double sh[TB / 64];
    const int s = s1 + blockIdx.x;
    const int side = 2 * s + 1;
    const int64_t lo_k2 = (int64_t)s * s - s + 1, hi_k2 = (int64_t)s * s + s;
    const int half = g.n / 2;
    double acc = 0.0;
    for (int p = threadIdx.x; p < side * side; p += blockDim.x) {
        const int ky = p / side - s, kz = p % side - s;
        if (ky <= -half || ky > half || kz <= -half || kz > half) continue;
        const int64_t q = (int64_t)ky * ky + (int64_t)kz * kz;
        if (q > hi_k2) continue;
        const int64_t lo2 = lo_k2 - q;
        int kx0 = 0;
        if (lo2 > 0) {
            kx0 = isqrt_floor(lo2);
            if ((int64_t)kx0 * kx0 < lo2) ++kx0;
        }
        int kx1 = isqrt_floor(hi_k2 - q);
        if (kx1 > half - 1) kx1 = half - 1;
        const int jy = ky < 0 ? ky + g.n : ky, jz = kz < 0 ? kz + g.n : kz;
        const int64_t rowoff = ((int64_t)jy * g.n + jz) * g.P;
        for (int kx = kx0; kx <= kx1; ++kx) {
            const cpx<double> a = C(u)[rowoff + kx], b = C(v)[rowoff + kx], c = C(w)[rowoff + kx];
            double e1 = fac * (a.x * a.x + b.x * b.x + c.x * c.x);
            double e2 = fac * (a.y * a.y + b.y * b.y + c.y * c.y);
            if (kx == 0) {
                e1 *= 0.5;
                e2 *= 0.5;
            }
            acc += e1;
            acc += e2;
        }
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

__global__ void k_force_apply(Geo g, CPtr3 f, Ptr3 r, const double* e, int kfmax, double famp,
                              double* energy_out) {
    __shared__ double fac_sh;
    if (threadIdx.x == 0) {
        double en = 0.0;
        for (int i = 0; i < kfmax; ++i) en += e[i];
        fac_sh = famp / en;
        *energy_out = en;
    }
    __syncthreads();
    const double fac = fac_sh;
    const int side = 2 * kfmax + 1;
    const int cnt = side * side * (kfmax + 1);
    const int half = g.n / 2;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        const int kx = p % (kfmax + 1);
        const int q = p / (kfmax + 1);
        const int ky = q / side - kfmax, kz = q % side - kfmax;
        if (kx >= half || ky <= -half || ky > half || kz <= -half || kz > half) continue;
        const int sh = shell_of(kx * kx + ky * ky + kz * kz);
        if (sh <= 0 || sh > kfmax) continue;
        const int jy = ky < 0 ? ky + g.n : ky, jz = kz < 0 ? kz + g.n : kz;
        const int64_t off = ((int64_t)jy * g.n + jz) * g.P + kx;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cpx<double> a = C(f.p[c])[off];
            cpx<double> b = C(r.p[c])[off];
            C(r.p[c])[off] = {b.x + fac * a.x, b.y + fac * a.y};
        }
    }
}

__global__ void k_enstrophy_plane(Geo g, const double* u, const double* v, const double* w,
                                  double* partial) {
    __shared__ double sh[TB / 64];
    const int k = blockIdx.x;
    const int H = g.n / 2 + 1;
    const double ky = wavenum(k, g.n);
    double acc = 0.0;
    for (int p = threadIdx.x; p < g.n * H; p += blockDim.x) {
        const int kx = p % H, j = p / H;
        const double kz = wavenum(j, g.n), kxd = kx;
        const int64_t off = ((int64_t)k * g.n + j) * g.P + kx;
        const cpx<double> a = C(u)[off], b = C(v)[off], c = C(w)[off];
        // |k x U|^2 (the factor i does not change the modulus)
        const double x1r = ky * c.x - kz * b.x, x1i = ky * c.y - kz * b.y;
        const double x2r = kz * a.x - kxd * c.x, x2i = kz * a.y - kxd * c.y;
        const double x3r = kxd * b.x - ky * a.x, x3i = kxd * b.y - ky * a.y;
        double t = x1r * x1r + x1i * x1i + x2r * x2r + x2i * x2i + x3r * x3r + x3i * x3i;
        if (kx != 0 && kx != g.n / 2) t *= 2.0;
        acc += t;
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[k] = tot;
}

__global__ void k_sum_serial(const double* partial, int cnt, double scale, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < cnt; ++i) s += partial[i];
        *out = s * scale;
    }
}

__global__ void k_ialias(Geo g, int8_t* out) {
    const int64_t total = (int64_t)g.n * g.n * (g.n + 2);
    GRID_STRIDE(i, total) {
        const int ii = (int)(i % (g.n + 2));
        const int64_t r = i / (g.n + 2);
        const int j = (int)(r % g.n), k = (int)(r / g.n);
        out[i] = (int8_t)alias_count(ii / 2, wavenum(k, g.n), wavenum(j, g.n), g.n);
    }
}

__global__ void k_init_curl(Geo g, const double* w1, const double* w2, const double* w3,
                            double* f1, double* f2, double* f3, int kmax) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const int sh = shell_of(s.kx * s.kx + s.kyv * s.kyv + s.kzv * s.kzv);
        if (sh > 0 && sh <= kmax) {
            const double kx = s.kx, ky = s.kyv, kz = s.kzv;
            const cpx<double> a = C(w1)[s.off], b = C(w2)[s.off], c = C(w3)[s.off];
            C(f1)[s.off] = {-(ky * c.y - kz * b.y), ky * c.x - kz * b.x};
            C(f2)[s.off] = {-(kz * a.y - kx * c.y), kz * a.x - kx * c.x};
            C(f3)[s.off] = {-(kx * b.y - ky * a.y), kx * b.x - ky * a.x};
        } else {
            C(f1)[s.off] = {0.0, 0.0};
            C(f2)[s.off] = {0.0, 0.0};
            C(f3)[s.off] = {0.0, 0.0};
        }
    }
}

__global__ void k_init_rescale(Geo g, double* f, const double* ratio, int kmax) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const int sh = shell_of(s.kx * s.kx + s.kyv * s.kyv + s.kzv * s.kzv);
        cpx<double> a = C(f)[s.off];
        if (sh > 0 && sh <= kmax && ratio[sh] >= 0.0)
            C(f)[s.off] = {a.x * ratio[sh], a.y * ratio[sh]};
        else
            C(f)[s.off] = {0.0, 0.0};
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int64_t nspec_h(const Geo& g) { return (int64_t)g.n * g.n * (g.n / 2 + 1); }
static int64_t nphys_h(const Geo& g) { return (int64_t)g.n * g.n * g.n; }

void launch_derivative(const Geo& g, const double* in, double* out, int axis, hipStream_t s) {
    hipLaunchKernelGGL(k_derivative, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, in, out, axis);
}
void launch_truncate(const Geo& g, double* slot, int kmax, hipStream_t s) {
    hipLaunchKernelGGL(k_truncate, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, slot, kmax);
}
void launch_pressure(const Geo& g, double* u, double* v, double* w, hipStream_t s) {
    hipLaunchKernelGGL(k_pressure, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, u, v, w);
}
void launch_advance(const Geo& g, double* f, const double* r, double* ro, double dt,
                    const double* dtp, bool first, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, f, r, ro, dt, dtp,
                       first ? 1 : 0);
}
void launch_advance_project(const Geo& g, double* const f[3], const double* const r[3],
                            double* const ro[3], double dt, const double* dtp, bool first,
                            hipStream_t s) {
    Ptr3 pf, pro;
    CPtr3 pr;
    for (int c = 0; c < 3; ++c) {
        pf.p[c] = f[c];
        pr.p[c] = r[c];
        pro.p[c] = ro[c];
    }
    hipLaunchKernelGGL(k_advance_project, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, pf, pr, pro,
                       dt, dtp, first ? 1 : 0);
}
void launch_fill_zero(double* p, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(count)), dim3(TB), 0, s, p, count);
}
void launch_stream_copy(const void* src, void* dst, int64_t bytes, int blocks, hipStream_t s) {
    if (bytes % 16 != 0 || blocks < 1) throw std::runtime_error("stream copy: 16-byte words");
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s,
                       static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), bytes / 16);
}
void launch_narrow(const double* src, float* dst, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(count)), dim3(TB), 0, s, src, dst, count);
}
void launch_widen(const float* src, double* dst, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_widen, dim3(grid_for(count)), dim3(TB), 0, s, src, dst, count);
}
void launch_max_fold(const double* part, int64_t count, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_max_fold, dim3(1), dim3(1024), 0, s, part, count, out);
}
void launch_products(const Geo& g, const double* u, const double* v, const double* w,
                     double* const out[6], double* blockmax, double* maxout, hipStream_t s) {
    Ptr6 o;
    for (int i = 0; i < 6; ++i) o.p[i] = out[i];
    const int nb = grid_for(nphys_h(g));  // <= 65536, the smallest size of blockmax
    hipLaunchKernelGGL(k_products, dim3(nb), dim3(TB), 0, s, g, u, v, w, o, blockmax);
    launch_max_fold(blockmax, nb, maxout, s);
}
void launch_rhs_assemble(const Geo& g, const double* const P[6], const double* const U[3],
                         double* const out[3], double nu, bool phase_shift, hipStream_t s) {
    CPtr6 p;
    CPtr3 u;
    Ptr3 o;
    for (int i = 0; i < 6; ++i) p.p[i] = P[i];
    for (int i = 0; i < 3; ++i) {
        u.p[i] = U[i];
        o.p[i] = out[i];
    }
    hipLaunchKernelGGL(k_rhs_assemble, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, p, u, o, nu,
                       phase_shift ? 1 : 0);
}
void launch_phys_mul(const Geo& g, const double* a, const double* b, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_phys_mul, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a, b, out);
}
void launch_scalar_rhs(const Geo& g, const double* const Q[3], const double* F, double* out,
                       double pe, const double* sub, hipStream_t s) {
    CPtr3 q;
    for (int i = 0; i < 3; ++i) q.p[i] = Q[i];
    hipLaunchKernelGGL(k_scalar_rhs, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, q, F, out, pe,
                       sub);
}
void launch_scalar_rhs_ps(const Geo& g, const double* const Q[3], const double* F, double* out,
                          double pe, bool accumulate, const double* sub, hipStream_t s) {
    CPtr3 q;
    for (int i = 0; i < 3; ++i) q.p[i] = Q[i];
    hipLaunchKernelGGL(k_scalar_rhs_ps, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, q, F, out,
                       pe, accumulate ? 1 : 0, accumulate ? sub : nullptr);
}
void launch_shell_sums(const Geo& g, const double* u, const double* v, const double* w, int s1,
                       int ns, double* out, hipStream_t s) {
    if (ns <= 0) return;
    const double n3 = (double)g.n * (double)g.n * (double)g.n;
    const double fac = 1.0 / (n3 * n3);
    hipLaunchKernelGGL(k_shell_sums, dim3(ns), dim3(TB), 0, s, g, u, v, w, s1, fac, out);
}
void launch_force_apply(const Geo& g, const double* const f[3], double* const r[3],
                        const double* e, int kfmax, double famp, double* energy_out,
                        hipStream_t s) {
    CPtr3 ff;
    Ptr3 rr;
    for (int i = 0; i < 3; ++i) {
        ff.p[i] = f[i];
        rr.p[i] = r[i];
    }
    hipLaunchKernelGGL(k_force_apply, dim3(1), dim3(TB), 0, s, g, ff, rr, e, kfmax, famp,
                       energy_out);
}
void launch_enstrophy(const Geo& g, const double* u, const double* v, const double* w,
                      double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_enstrophy_plane, dim3(g.n), dim3(TB), 0, s, g, u, v, w, partial);
    const double n3 = (double)g.n * (double)g.n * (double)g.n;
    hipLaunchKernelGGL(k_sum_serial, dim3(1), dim3(64), 0, s, partial, g.n, 1.0 / (n3 * n3), out);
}
void launch_ialias(const Geo& g, int8_t* out, hipStream_t s) {
    const int64_t total = (int64_t)g.n * g.n * (g.n + 2);
    hipLaunchKernelGGL(k_ialias, dim3(grid_for(total)), dim3(TB), 0, s, g, out);
}
void launch_init_curl(const Geo& g, const double* w1, const double* w2, const double* w3,
                      double* f1, double* f2, double* f3, int kmax, hipStream_t s) {
    hipLaunchKernelGGL(k_init_curl, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, w1, w2, w3, f1,
                       f2, f3, kmax);
}
void launch_init_rescale(const Geo& g, double* f, const double* ratio, int kmax, hipStream_t s) {
    hipLaunchKernelGGL(k_init_rescale, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, f, ratio,
                       kmax);
}

// ---------------------------------------------------------------------------
// scalar statistics, divergence
// ---------------------------------------------------------------------------
namespace {

__global__ void k_shell_sums1(Geo g, const double* f, int s1, double fac, double* out) {
    __shared__ double sh[TB / 64];
    const int s = s1 + blockIdx.x;
    const int side = 2 * s + 1;
    const int64_t lo_k2 = (int64_t)s * s - s + 1, hi_k2 = (int64_t)s * s + s;
    const int half = g.n / 2;
    double acc = 0.0;
    for (int p = threadIdx.x; p < side * side; p += blockDim.x) {
        const int ky = p / side - s, kz = p % side - s;
        if (ky <= -half || ky > half || kz <= -half || kz > half) continue;
        const int64_t q = (int64_t)ky * ky + (int64_t)kz * kz;
        if (q > hi_k2) continue;
        const int64_t lo2 = lo_k2 - q;
        int kx0 = 0;
        if (lo2 > 0) {
            kx0 = isqrt_floor(lo2);
            if ((int64_t)kx0 * kx0 < lo2) ++kx0;
        }
        int kx1 = isqrt_floor(hi_k2 - q);
        if (kx1 > half - 1) kx1 = half - 1;
        const int jy = ky < 0 ? ky + g.n : ky, jz = kz < 0 ? kz + g.n : kz;
        const int64_t rowoff = ((int64_t)jy * g.n + jz) * g.P;
        for (int kx = kx0; kx <= kx1; ++kx) {
            const cpx<double> a = C(f)[rowoff + kx];
            double e1 = fac * (a.x * a.x), e2 = fac * (a.y * a.y);
            if (kx == 0) {
                e1 *= 0.5;
                e2 *= 0.5;
            }
            acc += e1;
            acc += e2;
        }
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

// per-z-plane sums of a physical slot: with b and c given, sum((a^2+b^2)+c^2);
// otherwise sum(a) and sum(a^2).  x < n only.
__global__ void k_phys_plane_sums(Geo g, const double* a, const double* b, const double* c,
                                  double* partial) {
    __shared__ double sh[TB / 64];
    const int z = blockIdx.x;
    double s0 = 0.0, s1 = 0.0;
    for (int p = threadIdx.x; p < g.n * g.n; p += blockDim.x) {
        const int x = p % g.n, y = p / g.n;
        const int64_t o = ((int64_t)z * g.n + y) * 2 * g.P + x;
        const double va = a[o];
        if (b) {
            const double vb = b[o], vc = c[o];
            s0 += (va * va + vb * vb) + vc * vc;
        } else {
            s0 += va;
            s1 += va * va;
        }
    }
    const double t0 = block_sum(s0, sh);
    const double t1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        partial[2 * z] = t0;
        partial[2 * z + 1] = t1;
    }
}

__global__ void k_phys_plane_m23(Geo g, const double* a, double* partial) {
    __shared__ double sh[TB / 64];
    const int z = blockIdx.x;
    double s2 = 0.0, s3 = 0.0;
    for (int p = threadIdx.x; p < g.n * g.n; p += blockDim.x) {
        const int x = p % g.n, y = p / g.n;
        const double va = a[((int64_t)z * g.n + y) * 2 * g.P + x];
        const double q = va * va;
        s2 += q;
        s3 += q * va;
    }
    const double t2 = block_sum(s2, sh);
    const double t3 = block_sum(s3, sh);
    if (threadIdx.x == 0) {
        partial[2 * z] = t2;
        partial[2 * z + 1] = t3;
    }
}

__global__ void k_sum2_serial(const double* partial, int cnt, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < cnt; ++i) {
            s0 += partial[2 * i];
            s1 += partial[2 * i + 1];
        }
        out[0] = s0;
        out[1] = s1;
    }
}

// order-preserving map of doubles onto unsigned keys
__device__ __forceinline__ unsigned long long dkey(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void k_minmax_init(unsigned long long* keys) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        keys[0] = ~0ull;
        keys[1] = 0ull;
    }
}

// min and max of (a + b) + c over x < n (b, c optional); when sum is given,
// the combined value is stored there (physical layout, pad entries untouched).
__global__ void k_minmax(Geo g, const double* a, const double* b, const double* c, double* sum,
                         unsigned long long* keys) {
    __shared__ double shl[TB / 64], shh[TB / 64];
    const int64_t total = (int64_t)g.n * g.n * g.n;
    double lo = 1.0e308, hi = -1.0e308;
    GRID_STRIDE(i, total) {
        const int64_t x = i % g.n, row = i / g.n;
        const int64_t o = row * 2 * g.P + x;
        double v = a[o];
        if (b) v = (v + b[o]) + c[o];
        if (sum) sum[o] = v;
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double l = __shfl_down(lo, o, 64), h = __shfl_down(hi, o, 64);
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        shl[wv] = lo;
        shh[wv] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            lo = shl[i] < lo ? shl[i] : lo;
            hi = shh[i] > hi ? shh[i] : hi;
        }
        atomicMin(keys, dkey(lo));
        atomicMax(keys + 1, dkey(hi));
    }
}

}  // namespace

void launch_shell_sums1(const Geo& g, const double* f, int s1, int ns, double* out, hipStream_t s) {
    if (ns <= 0) return;
    const double n3 = (double)g.n * (double)g.n * (double)g.n;
    hipLaunchKernelGGL(k_shell_sums1, dim3(ns), dim3(TB), 0, s, g, f, s1, 1.0 / (n3 * n3), out);
}
void launch_phys_sums(const Geo& g, const double* a, const double* b, const double* c,
                      double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_phys_plane_sums, dim3(g.n), dim3(TB), 0, s, g, a, b, c, partial);
    hipLaunchKernelGGL(k_sum2_serial, dim3(1), dim3(64), 0, s, partial, g.n, out);
}
void launch_phys_moments23(const Geo& g, const double* a, double* partial, double* out,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_phys_plane_m23, dim3(g.n), dim3(TB), 0, s, g, a, partial);
    hipLaunchKernelGGL(k_sum2_serial, dim3(1), dim3(64), 0, s, partial, g.n, out);
}
void launch_minmax(const Geo& g, const double* a, const double* b, const double* c, double* sum,
                   unsigned long long* keys, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, s, keys);
    hipLaunchKernelGGL(k_minmax, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a, b, c, sum, keys);
}

// ---------------------------------------------------------------------------
// scalar reactions
// ---------------------------------------------------------------------------
namespace {

__global__ void k_reaction(Geo g, const double* a, double* out, int rtype, double rrate,
                           const double* mean_src) {
    const double mean = mean_src[0] / ((double)g.n * (double)g.n * (double)g.n);
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) {
        const int64_t o = (i / g.n) * 2 * g.P + i % g.n;
        const double w = a[o];
        const double base = rrate * (1.0 - w * w);
        out[o] = rtype == 1 ? base : (rtype == 2 ? base * w : base * (w - mean));
    }
}

__global__ void k_double_delta(Geo g, double* a) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) {
        const int64_t o = (i / g.n) * 2 * g.P + i % g.n;
        a[o] = copysign(1.0, a[o]) * 0.9;
    }
}

__global__ void k_shell_cut(Geo g, double* f, int kcut) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const int sh = shell_of(s.kx * s.kx + s.kyv * s.kyv + s.kzv * s.kzv);
        if (sh == 0 || sh >= kcut) C(f)[s.off] = {0.0, 0.0};
    }
}

__global__ void k_sphere_cut(Geo g, double* f, int kmax) {
    const double rk2 = (double)kmax * (double)kmax;
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        if (kx * kx + ky * ky + kz * kz > rk2) C(f)[s.off] = {0.0, 0.0};
    }
}

__global__ void k_pack_f32(Geo g, const double* a, float* out) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) out[i] = (float)a[(i / g.n) * 2 * g.P + i % g.n];
}

}  // namespace

void launch_sphere_cut(const Geo& g, double* slot, int kmax, hipStream_t s) {
    hipLaunchKernelGGL(k_sphere_cut, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, slot, kmax);
}
void launch_pack_f32(const Geo& g, const double* phys, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_f32, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, phys, out);
}

void launch_double_delta(const Geo& g, double* a, hipStream_t s) {
    hipLaunchKernelGGL(k_double_delta, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a);
}
void launch_shell_cut(const Geo& g, double* slot, int kcut, hipStream_t s) {
    hipLaunchKernelGGL(k_shell_cut, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, slot, kcut);
}

void launch_reaction(const Geo& g, const double* a, double* out, int rtype, double rrate,
                     const double* mean_src, hipStream_t s) {
    hipLaunchKernelGGL(k_reaction, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a, out, rtype,
                       rrate, mean_src);
}

// ---------------------------------------------------------------------------
// Smagorinsky model
// ---------------------------------------------------------------------------
namespace {

__global__ void k_strain(Geo g, CPtr3 U, Ptr3 out, int group) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        const cpx<double> u = C(U.p[0])[s.off], v = C(U.p[1])[s.off], w = C(U.p[2])[s.off];
        if (group == 1) {
            C(out.p[0])[s.off] = {-(kx * u.y), kx * u.x};
            C(out.p[1])[s.off] = {-0.5 * (ky * u.y + kx * v.y), 0.5 * (ky * u.x + kx * v.x)};
            C(out.p[2])[s.off] = {-0.5 * (kz * u.y + kx * w.y), 0.5 * (kz * u.x + kx * w.x)};
        } else {
            C(out.p[0])[s.off] = {-(ky * v.y), ky * v.x};
            C(out.p[1])[s.off] = {-0.5 * (kz * v.y + ky * w.y), 0.5 * (kz * v.x + ky * w.x)};
            C(out.p[2])[s.off] = {-(kz * w.y), kz * w.x};
        }
    }
}

__global__ void k_tv_accum(Geo g, const double* a, const double* b, const double* c, double* tv,
                           int group, double coef) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) {
        const int64_t x = i % g.n, row = i / g.n;
        const int64_t o = row * 2 * g.P + x;
        const double A = a[o], Bv = b[o], Cv = c[o];
        if (group == 1) {
            double t = A * A;
            t = t + 2.0 * (Bv * Bv);
            t = t + 2.0 * (Cv * Cv);
            tv[o] = t;
        } else {
            double t = tv[o];
            t = t + A * A;
            t = t + 2.0 * (Bv * Bv);
            t = t + Cv * Cv;
            tv[o] = sqrt(2.0 * t) * coef;
        }
    }
}

__global__ void k_tau(Geo g, double* a, const double* tv) {
    const int64_t total = (int64_t)g.n * g.n * g.n;
    GRID_STRIDE(i, total) {
        const int64_t x = i % g.n, row = i / g.n;
        const int64_t o = row * 2 * g.P + x;
        a[o] = -2.0 * tv[o] * a[o];
    }
}

__global__ void k_les_vel_accum(Geo g, CPtr3 T, Ptr3 R, int group) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        if (alias_count(s.kx, s.kyv, s.kzv, g.n) != 0) continue;
        const double kx = s.kx, ky = s.kyv, kz = s.kzv;
        const cpx<double> a = C(T.p[0])[s.off], b = C(T.p[1])[s.off], c = C(T.p[2])[s.off];
        if (group == 1) {  // a = tau11, b = tau12, c = tau13
            cpx<double> r = C(R.p[0])[s.off];
            C(R.p[0])[s.off] = {r.x + kx * a.y + ky * b.y + kz * c.y, r.y - kx * a.x - ky * b.x - kz * c.x};
            r = C(R.p[1])[s.off];
            C(R.p[1])[s.off] = {r.x + kx * b.y, r.y - kx * b.x};
            r = C(R.p[2])[s.off];
            C(R.p[2])[s.off] = {r.x + kx * c.y, r.y - kx * c.x};
        } else {  // a = tau22, b = tau23, c = tau33
            cpx<double> r = C(R.p[1])[s.off];
            C(R.p[1])[s.off] = {r.x + ky * a.y + kz * b.y, r.y - ky * a.x - kz * b.x};
            r = C(R.p[2])[s.off];
            C(R.p[2])[s.off] = {r.x + ky * b.y + kz * c.y, r.y - ky * b.x - kz * c.x};
        }
    }
}

__global__ void k_mask_alias(Geo g, double* slot, int maxcount) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        if (alias_count(s.kx, s.kyv, s.kzv, g.n) > maxcount) C(slot)[s.off] = {0.0, 0.0};
    }
}

__global__ void k_add(Geo g, double* out, const double* a, int only_unaliased) {
    GRID_STRIDE(i, nspec(g)) {
        SIdx s = sdecode(g, i);
        if (only_unaliased && alias_count(s.kx, s.kyv, s.kzv, g.n) != 0) continue;
        const cpx<double> o = C(out)[s.off], b = C(a)[s.off];
        C(out)[s.off] = {o.x + b.x, o.y + b.y};
    }
}

}  // namespace

void launch_strain(const Geo& g, const double* const U[3], double* const out[3], int group,
                   hipStream_t s) {
    CPtr3 u;
    Ptr3 o;
    for (int i = 0; i < 3; ++i) {
        u.p[i] = U[i];
        o.p[i] = out[i];
    }
    hipLaunchKernelGGL(k_strain, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, u, o, group);
}
void launch_tv_accum(const Geo& g, const double* a, const double* b, const double* c, double* tv,
                     int group, double coef, hipStream_t s) {
    hipLaunchKernelGGL(k_tv_accum, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a, b, c, tv,
                       group, coef);
}
void launch_tau(const Geo& g, double* a, const double* tv, hipStream_t s) {
    hipLaunchKernelGGL(k_tau, dim3(grid_for(nphys_h(g))), dim3(TB), 0, s, g, a, tv);
}
void launch_les_vel_accum(const Geo& g, const double* const tau[3], double* const rhs[3],
                          int group, hipStream_t s) {
    CPtr3 t;
    Ptr3 r;
    for (int i = 0; i < 3; ++i) {
        t.p[i] = tau[i];
        r.p[i] = rhs[i];
    }
    hipLaunchKernelGGL(k_les_vel_accum, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, t, r, group);
}
void launch_mask_alias(const Geo& g, double* slot, int maxcount, hipStream_t s) {
    hipLaunchKernelGGL(k_mask_alias, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, slot, maxcount);
}
void launch_add(const Geo& g, double* out, const double* a, bool only_unaliased, hipStream_t s) {
    hipLaunchKernelGGL(k_add, dim3(grid_for(nspec_h(g))), dim3(TB), 0, s, g, out, a,
                       only_unaliased ? 1 : 0);
}

}  // namespace hit3d
