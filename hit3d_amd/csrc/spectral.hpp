// Pointwise spectral/physical kernels and FP64 reductions of the time step.
#pragma once
#include "common.hpp"

namespace hit3d {

struct Geo {
    int n;        // grid points per direction
    int64_t P;    // row pitch in complex entries
    int64_t nc;   // complex entries per slot = n*n*P
};

// Physical-space per-row views use 2P reals per row.

// x_fftw.f90:637-682: out = i * k_axis * in (axis 1=x, 2=y, 3=z).
void launch_derivative(const Geo& g, const double* in, double* out, int axis, hipStream_t s);
// dealias_all.f90:19-32: zero entries with any |k_i| > kmax.
void launch_truncate(const Geo& g, double* slot, int kmax, hipStream_t s);
// pressure.f90:13-47.
void launch_pressure(const Geo& g, double* u, double* v, double* w, hipStream_t s);
// main.f90:137-146,156-164: f += dt*r (first) or f += dt*(1.5 r - 0.5 ro); ro = r.
// dt is *dtp when dtp is not null (a device value, so a captured graph
// replays with the current step size).
void launch_advance(const Geo& g, double* f, const double* r, double* ro, double dt,
                    const double* dtp, bool first, hipStream_t s);
// The velocity update of main.f90:156-164 and then pressure.f90:13-47 in
// one pass over the three components.
void launch_advance_project(const Geo& g, double* const f[3], const double* const r[3],
                            double* const ro[3], double dt, const double* dtp, bool first,
                            hipStream_t s);
// p[0..count) = 0 by a kernel.
void launch_fill_zero(double* p, int64_t count, hipStream_t s);
// streaming copy of `bytes` (a multiple of 16) with `blocks` workgroups of 256
void launch_stream_copy(const void* src, void* dst, int64_t bytes, int blocks, hipStream_t s);
// dst[i] = src[i] rounded to FP32 / widened to FP64, i < count (no overlap).
void launch_narrow(const double* src, float* dst, int64_t count, hipStream_t s);
void launch_widen(const float* src, double* dst, int64_t count, hipStream_t s);
// rhs_velocity.f90:31-32,56-72: max(|u|+|v|+|w|) into *maxout (per-block
// maxima staged in blockmax, >= 65536 doubles) and the six products u_i u_j
// (written to out[0..5], which may alias u, v, w).
void launch_products(const Geo& g, const double* u, const double* v, const double* w,
                     double* const out[6], double* blockmax, double* maxout, hipStream_t s);
// out[0] = max(part[0..count)), part non-negative.
void launch_max_fold(const double* part, int64_t count, double* out, hipStream_t s);
// rhs_velocity.f90:85-142 (half = false) or :183-232 (half = true):
// r_i = mask * (c * i * t_i - nu k^2 U_i), t_i = -k_j P_ij; c = 1 or 1/2;
// mask: alias_count == 0 (2/3 rule) or <= 1 (phase shift).
void launch_rhs_assemble(const Geo& g, const double* const P[6], const double* const U[3],
                         double* const out[3], double nu, bool phase_shift, hipStream_t s);
// physical out = a * b (x < n only).
void launch_phys_mul(const Geo& g, const double* a, const double* b, double* out, hipStream_t s);
// rhs_scalars.f90:88-121: r = mask0 * (-i k.Q - pe k^2 F); Q[3] = (u phi, v phi, w phi)^.
// sub (may be null): afterwards out -= sub at every entry (the mean-gradient
// source of rhs_scalars.f90:387-389).
void launch_scalar_rhs(const Geo& g, const double* const Q[3], const double* F, double* out,
                       double pe, const double* sub, hipStream_t s);
// rhs_scalars.f90:196-219 (phase-shifted half, Q already shifted back):
// out = mask1 * (-1/2 i k.Q)   and :253-269 (unshifted half):
// out += mask1 * (-1/2 i k.Q - pe k^2 F), then out -= sub (if not null) at
// every entry (rhs_scalars.f90:387-389).
void launch_scalar_rhs_ps(const Geo& g, const double* const Q[3], const double* F, double* out,
                          double pe, bool accumulate, const double* sub, hipStream_t s);
// Deterministic FP64 shell sums of (|u|^2+|v|^2+|w|^2)/n^6 over entries
// kx < n/2, kx = 0 halved, for shells s1..s1+ns-1 (m_stats.f90:207-228).
void launch_shell_sums(const Geo& g, const double* u, const double* v, const double* w, int s1,
                       int ns, double* out, hipStream_t s);
// m_force.f90:153-173: energy = sum(e[0..kfmax)), r += famp/energy * f on
// forced entries (0 < shell <= kfmax, kx < n/2).  *energy_out on device.
void launch_force_apply(const Geo& g, const double* const f[3], double* const r[3],
                        const double* e, int kfmax, double famp, double* energy_out,
                        hipStream_t s);
// Enstrophy <|omega|^2> by Parseval over the half spectrum, deterministic.
// partial must hold n doubles.
void launch_enstrophy(const Geo& g, const double* u, const double* v, const double* w,
                      double* partial, double* out, hipStream_t s);
// ialias in the reference layout (n, n, n+2) int8 [ky][kz][i].
void launch_ialias(const Geo& g, int8_t* out, hipStream_t s);
// Init curl with the FP64 shell mask (init_velocity.f90:123-147).
void launch_init_curl(const Geo& g, const double* w1, const double* w2, const double* w3,
                      double* f1, double* f2, double* f3, int kmax, hipStream_t s);
// init_velocity.f90:224-237: multiply by ratio[shell] (shell from the float32
// rule), zero elsewhere; ratio has kmax+1 entries (index 0 unused, <0 = zero).
void launch_init_rescale(const Geo& g, double* f, const double* ratio, int kmax, hipStream_t s);

// Scalar statistics (m_stats.f90:295-440) and the divergence check
// (pressure.f90:55-93).
// Shell sums of |phi|^2/n^6 (one field), kx = 0 halved, kx < n/2.
void launch_shell_sums1(const Geo& g, const double* f, int s1, int ns, double* out, hipStream_t s);
// Deterministic physical-space sums over x < n: with b and c non-null,
// out[0] = sum((a^2+b^2)+c^2); otherwise out[0] = sum(a), out[1] = sum(a^2).
// partial must hold 2n doubles.
void launch_phys_sums(const Geo& g, const double* a, const double* b, const double* c,
                      double* partial, double* out, hipStream_t s);
// out[0] = sum(a^2), out[1] = sum(a^3) over x < n, deterministic (partial: 2n).
void launch_phys_moments23(const Geo& g, const double* a, double* partial, double* out,
                           hipStream_t s);
// min and max over x < n of a (b, c null) or (a + b) + c, into keys[0..1] as
// order-preserving unsigned keys; the combined field is stored into sum when
// sum is non-null.
void launch_minmax(const Geo& g, const double* a, const double* b, const double* c, double* sum,
                   unsigned long long* keys, hipStream_t s);

// Scalar reaction rate in physical space (rhs_scalars.f90:405-456 and the
// type-311 form :352-356), x < n: out = r (1 - a^2) [rtype 1],
// r (1 - a^2) a [2], r (1 - a^2) (a - mean) [3] with mean = Re F(0,0,0)/n^3
// read from the spectral slot mean_src.  out may alias a.
void launch_reaction(const Geo& g, const double* a, double* out, int rtype, double rrate,
                     const double* mean_src, hipStream_t s);

// init_scalar_spectrum type 3 (init_scalars.f90:239-262): physical
// a = sign(1, a) * 0.9 (x < n); spectral zeroing of shells 0 and >= kcut
// (float32 shell rule, all kx).
void launch_double_delta(const Geo& g, double* a, hipStream_t s);
void launch_shell_cut(const Geo& g, double* slot, int kcut, hipStream_t s);

// io_write_4 (io_write_4.f90:16-37): zero entries with kx^2+ky^2+kz^2 > kmax^2;
// and the float32 copy of a physical slot, packed (x fastest, no padding).
void launch_sphere_cut(const Geo& g, double* slot, int kmax, hipStream_t s);
void launch_pack_f32(const Geo& g, const double* phys, float* out, hipStream_t s);

// Smagorinsky model (m_les.f90:902-1215).
// Strain components of group 1 (S11, S12, S13) or group 2 (S22, S23, S33)
// from the spectral velocity (m_les.f90:1154-1196).
void launch_strain(const Geo& g, const double* const U[3], double* const out[3], int group,
                   hipStream_t s);
// nu_t accumulation over physical strain components: group 1 sets
// tv = a^2 + 2b^2 + 2c^2, group 2 adds a^2 + 2b^2 + c^2 and finishes
// tv = sqrt(2 tv) * coef (m_les.f90:1170-1212).
void launch_tv_accum(const Geo& g, const double* a, const double* b, const double* c, double* tv,
                     int group, double coef, hipStream_t s);
// physical a = -2 tv a (m_les.f90:930-934).
void launch_tau(const Geo& g, double* a, const double* tv, hipStream_t s);
// rhs += -i k_j tau_ij on entries with alias_count == 0 for one group
// (m_les.f90:937-960 group 1, :987-1015 group 2).
void launch_les_vel_accum(const Geo& g, const double* const tau[3], double* const rhs[3],
                          int group, hipStream_t s);
// zero entries with alias_count > maxcount.
void launch_mask_alias(const Geo& g, double* slot, int maxcount, hipStream_t s);
// out += a on all entries (only_unaliased = false) or on alias_count == 0.
void launch_add(const Geo& g, double* out, const double* a, bool only_unaliased, hipStream_t s);

}  // namespace hit3d
