/*
 * hit3d.h -- C ABI of the MI355X-native HIT3DP time step (libhit3d_hip.so).
 *
 * Drop-in boundary for the reference's Fortran subroutines.  Conventions:
 *   - every integer argument is an int64_t passed by reference and every real
 *     a double passed by reference, as a Fortran caller built with
 *     -fdefault-integer-8 -fdefault-real-8 passes its default kinds
 *     (reference Makefile:11,29,53); Fortran binds them with bind(C) and
 *     VALUE-less dummies (the binding module is hit3d_amd/fortran/hit3d_hip.f90);
 *   - every entry point returns 0 on success and a non-zero status on error;
 *     hit3d_last_error() copies the message.  The reference has no error
 *     codes: its fatal paths call my_exit(-1) (my_exit.f90:23-62), which the
 *     Fortran wrapper does on a non-zero status;
 *   - the state the reference keeps in modules (fields, wrk(0:nw), rhs_old,
 *     m_fields.f90:5,21; m_work.f90:16-22,102-109) lives on the GPU.  Host
 *     copies in the reference layout (nx+2, ny, nz) are exchanged only by
 *     hit3d_put / hit3d_get, at the points where host code reads or writes
 *     them (init, restart, output).  Array ids: 1 = fields (slots 1..nf),
 *     2 = wrk (slots 0..nw), 3 = rhs_old (slots 1..nf).
 *   - one process, one GPU, one HIP stream; calls are synchronous from the
 *     caller's point of view where they return host values.
 */
#ifndef HIT3D_H
#define HIT3D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* m_fields_init + m_work_init + x_fftw_allocate(1) + x_fftw_init
 * (m_fields.f90:11-33, m_work.f90:30-110, x_fftw.f90:81-317): allocates
 * fields(3+ns+nles), wrk(0:max(6,3+ns+nles+2)), rhs_old(3+ns+nles) on the
 * device (zeroed), builds twiddle and phase-shift tables.  nx = ny = nz_all,
 * a power of two in [8, 1024].  dealias: 0 = 2/3 rule, 1 = phase shift. */
int hit3d_init(const int64_t *nx, const int64_t *ny, const int64_t *nz_all,
               const int64_t *n_scalars, const int64_t *n_les, const int64_t *dealias);

/* Physical parameters read by read_input_file (m_parameters.f90:274-451):
 * nu = 1/RE, flow_type, kfmax, famp; per scalar: pe(n) = nu/sc(n) and
 * scalar_type(n) (arrays of n_scalars entries, may be NULL when 0). */
int hit3d_set_physics(const double *nu, const int64_t *flow_type, const int64_t *kfmax,
                      const double *famp, const double *pe, const int64_t *scalar_type);

/* Reaction rates reac_sc(1:n_scalars) (m_parameters.f90:445) for scalar types
 * >= 100, whose hundreds digit selects the rate (add_reaction,
 * rhs_scalars.f90:405-456: 1 KPP, 2 symmetric bistable, 3 self-adjusting
 * bistable; with the phase shift only type 311, :340-377).  Defaults to 0. */
int hit3d_set_reaction(const double *reac_sc);

/* x_fftw_allocate(-1) (x_fftw.f90:134-142): frees everything. */
int hit3d_finalize(void);

/* Host <-> device staging of one slot in the reference layout, pitch nx+2
 * reals (sync points of the drop-in; no reference counterpart). */
int hit3d_put(const int64_t *array_id, const int64_t *slot, const double *host);
int hit3d_get(const int64_t *array_id, const int64_t *slot, double *host);

/* dst(:,:,:,dslot) = src(:,:,:,sslot), the whole-slot array assignments the
 * reference writes inline, e.g. wrk(:,:,:,1:3) = fields(:,:,:,1:3)
 * (rhs_scalars.f90:34, m_stats.f90:254-256). */
int hit3d_copy(const int64_t *dst_array, const int64_t *dst_slot, const int64_t *src_array,
               const int64_t *src_slot);

/* xFFT3d(flag, n) (x_fftw.f90:852-1085): in-place 3-D real FFT of wrk slot n;
 * flag=+1 forward (physical (x,y,z) -> spectral (kx,kz,ky)), -1 inverse with
 * norm 1/(nx*ny*nz_all). */
int hit3d_xfft3d(const int64_t *flag, const int64_t *n);

/* xFFT3d_fields(flag, n) (x_fftw.f90:1098-1282): the same on fields slot n. */
int hit3d_xfft3d_fields(const int64_t *flag, const int64_t *n);

/* x_derivative(n, axis, nto) (x_fftw.f90:637-682): wrk(nto) = i k_axis wrk(n);
 * axis 1 = 'x', 2 = 'y', 3 = 'z' (the Fortran wrapper maps the character). */
int hit3d_x_derivative(const int64_t *n, const int64_t *axis, const int64_t *nto);

/* dealias_all (dealias_all.f90:1-35). */
int hit3d_dealias_all(void);

/* rhs_scalars (rhs_scalars.f90:1-402, with the reactions of hit3d_set_reaction
 * and the Smagorinsky scalar term of hit3d_set_les; LES-only scalars n_les > 0
 * are rejected): leaves the physical velocity in wrk(1:3) and, when scalars
 * are active (*int_scalars != 0), the scalar RHS in wrk(4:3+ns). */
int hit3d_rhs_scalars(const int64_t *int_scalars);

/* rhs_velocity (rhs_velocity.f90:1-419, with the Smagorinsky term of
 * hit3d_set_les, :414-416): reads the
 * physical velocity in wrk(1:3), writes the spectral RHS to wrk(1:3) and
 * returns courant = max(|u|+|v|+|w|)*dt/dx (:31-38). */
int hit3d_rhs_velocity(const double *dt, double *courant);

/* force_velocity (m_force.f90:101-182): band energy of shells 1..kfmax and
 * wrk(1:3) += famp/energy * fields(1:3) on the forced entries. */
int hit3d_force_velocity(double *energy);

/* The Euler/AB2 updates the driver writes inline (main.f90:137-146 for
 * which = 2, scalars 4..3+ns+nles; main.f90:156-164 for which = 1, velocity):
 * *first_order != 0: rhs_old = wrk; fields += dt*rhs_old; then *first_order = 0.
 * else fields += dt*(1.5 wrk - 0.5 rhs_old); rhs_old = wrk. */
int hit3d_advance(const int64_t *which, int64_t *first_order, const double *dt);

/* pressure (pressure.f90:1-51): projection of fields(1:3). */
int hit3d_pressure(void);

/* One hydro step of the driver loop (main.f90:124-167) fused on the GPU:
 * rhs_scalars, scalar update, rhs_velocity, force_velocity (flow_type 1),
 * velocity update, pressure.  fov/fos as in main.f90:71-72 (updated).
 * precision: 0 = FP64 throughout; 1 = FP32 transforms and nonlinear term
 * with FP64 state, update and reductions; the mixed form stores
 * rhs_old(1:3) in FP32, so after it hit3d_get of rhs_old returns
 * FP32-rounded values (converted to FP64 on that read and before any other
 * step).  Returns courant and, when forced, the band energy.  Leaves wrk
 * undefined. */
int hit3d_step(const double *dt, int64_t *fov, int64_t *fos, const int64_t *int_scalars,
               const int64_t *precision, double *courant, double *energy);

/* LES model selection (m_les.f90:69-104, dispatch :636-690, :700-895):
 * les_model 0 = none, 1 = Smagorinsky, 11 = the learned closure below (not a
 * reference model number), nu_t = (c_smag dx)^2 sqrt(2 S_ij S_ij)
 * (les_get_turb_visc_smag, m_les.f90:1142-1215) applied to velocities
 * (les_rhsv_turb_visc :902-1018) and scalars (m_les_rhss_turb_visc1
 * :1041-1091).  The reference hard-codes c_smag = 0.18 (:1148).  Needs the
 * 2/3 rule, as the reference forces (m_parameters.f90:415-420). */
int hit3d_set_les(const int64_t *les_model, const double *c_smag);

/* Learned subgrid-stress closure (no counterpart in the reference; it stands
 * in for the eddy viscosity of m_les.f90:1142-1215 in coarse LES, with
 * Smagorinsky as its fallback).  A pointwise MLP 10 -> 64 -> 64 -> 6 with GELU
 * maps the standardised inputs (du_i/dx_j in i-major order, then the filter
 * width) to the standardised deviatoric stress (tau_11, tau_12, tau_13,
 * tau_22, tau_23, tau_33); products run on bfloat16 matrix cores with FP32
 * accumulation.  Parameters are one float array of 6288 entries: W1[16][64]
 * (input rows 10..15 zero), b1[64], W2[64][64], b2[64], W3[64][16] (output
 * columns 6..15 zero), b3[16]; W[k][n] multiplies input k into output n.
 *
 * set_params also resets the Adam state; set_norm gives the input/output
 * means and scales (10, 10, 6, 6 doubles; scales > 0).  set_data uploads a
 * standardised training set X[count][10], T[count][6] (floats).  train runs
 * `steps` Adam steps (lr, betas 0.9/0.999, eps 1e-8) on consecutive batches
 * of `batch` samples (a multiple of 32; indices wrap), the loss being the
 * mean squared error over batch x 6 outputs; losses[i] receives the loss of
 * step i before its update.  grads returns the gradient of that loss over
 * the count samples given (a multiple of 32) without updating; forward the
 * network's standardised outputs.  hit3d_set_les(11, ...) selects the
 * closure for the LES right-hand side (velocity only); a step whose network
 * output is not finite everywhere uses the Smagorinsky stress instead and is
 * counted by hit3d_mlp_fallbacks. */
int hit3d_mlp_set_params(const float *params);
int hit3d_mlp_get_params(float *params);
int hit3d_mlp_set_norm(const double *xmean, const double *xstd, const double *ymean,
                       const double *ystd);
int hit3d_mlp_set_data(const int64_t *count, const float *X, const float *T);
int hit3d_mlp_train(const int64_t *steps, const int64_t *batch, const double *lr, double *losses);
int hit3d_mlp_grads(const int64_t *count, const float *X, const float *T, float *grads,
                    double *loss);
int hit3d_mlp_forward(const int64_t *count, const float *X, float *Y);
int hit3d_mlp_fallbacks(int64_t *count);
/* The LES stress of les_model 11 as a blend (no counterpart in the
 * reference): tau = (1 - weight) tau_network + weight tau_Smagorinsky, the
 * second -2 (c_smag dx)^2 |S| S_ij with the constant of hit3d_set_les,
 * before the backscatter clearing; weight in [0, 1], 0 (the network alone)
 * by default. */
int hit3d_mlp_set_mix(const double *weight);

/* hit3d_mlp_grads without any device allocation: no counterpart in the
 * reference (the closure has none); for callers that cannot allocate --
 * code under HIP graph capture, tight training loops, harnesses that forbid
 * hipMalloc.  Same kernels and results as hit3d_mlp_grads, bit for bit.
 * The samples' staging, the gradient partials and the reduction live in one
 * static device array of 8.2 MiB per GPU, held until the process ends;
 * count must be a multiple of 32 and at most 8192 (larger counts are
 * rejected with an error).  All calls on one GPU share that array: calls
 * that overlap on one GPU (other threads, other libraries' streams) race. */
int hit3d_mlp_grads_static(const int64_t *count, const float *X, const float *T, float *grads,
                           double *loss);

/* Training data for the closure.  box_filter: top-hat average over ratio^3
 * blocks of the physical velocity in wrk(1:3) (as hit3d_fields_to_parts
 * leaves it) onto the grid m = nx/ratio: ubar[3][m^3] and the subgrid stress
 * tau[6][m^3] = avg(u_i u_j) - ubar_i ubar_j (11,12,13,22,23,33), coarse
 * index (z*m + y)*m + x; grids up to 512.  sgs_features: the closure's raw
 * inputs at every point of this grid from fields(1:3): X[p][10] =
 * (du_i/dx_j, delta), p = (z*nx + y)*nx + x. */
int hit3d_box_filter(const int64_t *ratio, double *ubar, double *tau);
int hit3d_sgs_features(const double *delta, float *X);

/* Form of the nonlinear term in hit3d_step's fused pipelines:
 * 0 (default) the divergence form -i k_j (u_i u_j)^ of rhs_velocity.f90:56-102;
 * 1 the rotational form (u x omega)^, omega = i k x u-hat (6 inverse and 3
 * forward transforms; 2/3 rule only); 2 five products instead of six: with
 * the 2/3 rule the divergence of the trace-free tensor u_i u_j - delta_ij
 * |u|^2/3 (3 inverse and 5 forward transforms), with the phase shift (the
 * fused velocity step of rhs_velocity.f90:150-410) that of u_i u_j - delta_ij
 * w w in both halves.  Forms 1 and 2 have no reference counterpart: they
 * differ from form 0 by a gradient (i k (|u|^2/2)^, i k (|u|^2/3)^,
 * i k (w w)^, masked like the rest of the RHS), which the projection after
 * the update removes, so the fields agree to rounding while rhs_old holds
 * the form's own RHS.  A step that cannot run fused (2/3 rule with active
 * scalars or LES, fusion off) fails while form 1 or 2 is selected.  The
 * unfused entry points (hit3d_rhs_velocity) keep the divergence form. */
int hit3d_set_nonlinear(const int64_t *form);

/* Selects how hit3d_step runs: *on != 0 (default) the fused pipeline where
 * it applies (2/3 rule without active scalars), 0 the unfused sequence of
 * the reference subroutines above (no reference counterpart; for A/B
 * checks).  The mixed form needs the fused pipeline. */
int hit3d_set_fusion(const int64_t *on);

/* Selects whether hit3d_step replays captured HIP graphs: *on != 0
 * (default) captures a step's kernels once per combination of slot pointers
 * and step flags (after running that combination once directly) and replays
 * it, cutting the launch cost of small grids; 0 launches every kernel
 * directly.  Results are identical.  Steps with timing enabled or with
 * les_model 11 always launch directly.  No reference counterpart. */
int hit3d_set_graphs(const int64_t *on);
/* Number of step graphs held now and of steps run from one so far. */
int hit3d_graph_stats(int64_t *captured, int64_t *replays);

/* Statistics of m_stats.f90:195-289: e_spec(1:kmax) (FP64 shell sums) and
 * the mean enstrophy <|omega|^2>. */
int hit3d_stats(double *e_spec, double *enstrophy);

/* stat_scalars with get_scalar_spectra (m_stats.f90:295-437), for every
 * scalar n = 1..n_scalars into element n-1 of each array (any may be NULL):
 * sc_diss = 2 pe <|grad phi|^2>, mean, variance, sc_min, sc_max (physical
 * space, the reference's moments(3+n,1:2)), and es_spec(kmax, n_scalars) =
 * 2 x the FP64 shell sums of |phi|^2/n^6 (written to es_NN.gp by the
 * reference).  Uses wrk(0:3) as work space like the reference. */
int hit3d_stat_scalars(double *sc_diss, double *mean, double *variance, double *sc_min,
                       double *sc_max, double *es_spec);

/* Lagrangian particles (m_particles.f90).  Positions are grid units, 1-based,
 * in [1, n+1) per coordinate, in the reference's xyzp(3, np) order (x, y, z
 * of particle 1 first); uvwp(3, np) holds the velocity interpolated at the
 * previous move, or -999 before the first (Euler) move.
 *
 * hit3d_particles_set replaces the device copy (particles_init and
 * particles_restart_read_binary end with these arrays, :300-327, :720-740);
 * uvwp NULL means -999 everywhere.  Positions outside [1, n+1) are rejected. */
int hit3d_particles_set(const int64_t *np, const double *xyzp, const double *uvwp);
/* Copies the device particles back (either pointer may be NULL). */
int hit3d_particles_get(double *xyzp, double *uvwp);
/* fields_to_parts (m_fields.f90:209-290) on one GPU: wrk(1:3) = the physical
 * velocity of fields(1:3), the array rhs_scalars leaves there
 * (rhs_scalars.f90:32-39). */
int hit3d_fields_to_parts(void);
/* particles_move (m_particles.f90:749-791) without the locally averaged
 * velocity option: interpolate wrk(1:3) at every particle (scheme 0
 * trilinear :815-1009, 1 CINT tricubic :1017-1230 with cint_3d :1451-1508),
 * Euler or AB2 in grid units with step dt, periodic wrap in x, y and z (the
 * slab exchange of :1236-1300 on one rank).  Particles keep their order. */
int hit3d_particles_move(const double *dt, const int64_t *scheme);

/* Velocity-derivative skewness, the mean over i of
 * <(du_i/dx_i)^3> / <(du_i/dx_i)^2>^(3/2) in physical space (no reference
 * counterpart; a statistic for comparing the precision forms).  Uses wrk(0). */
int hit3d_skewness(double *skew);

/* divergence (pressure.f90:53-93): min and max over the grid of
 * du/dx + dv/dy + dw/dz in physical space; leaves wrk(1:3) = fields(1:3),
 * wrk(4:6) = the three derivatives in physical space and wrk(0) = their sum. */
int hit3d_divergence(double *dmin, double *dmax);

/* init_velocity (init_velocity.f90:1-246): seeded random solenoidal field
 * from RANDu (RANDu.f:12-52) with seed rn1, shell-rescaled to the spectrum
 * isp_type (0, 1, 3 as the reference; 2 = k^ir_exp exp(-(ir_exp/2)(k/peak)^2),
 * an addition) into fields(1:3).  isp_type -1 = Taylor-Green (:252-293). */
int hit3d_init_velocity(const double *rn1, const int64_t *isp_type, const int64_t *ir_exp,
                        const double *peak);

/* init_scalars (init_scalars.f90:6-383) for every scalar, by the last two
 * digits of its type (hit3d_set_physics): 0 zero field; 1-9 init_scalar_spectrum
 * (type 1 or 3 exponential, 2 von Karman spectrum with peak peak_sc(n); 3 then
 * made double-delta +-0.9 and smoothed) from the reference generator seeded
 * with rn2, continuing from the generator state left by hit3d_init_velocity
 * (or by the previous spectrum scalar) as the reference's module state does;
 * 11 tanh slab in z, 12 two tanh slabs in x, 13 sin waves in x.  Ends by
 * zeroing every mode with ialias > 0 (:27-36). */
int hit3d_init_scalars(const double *rn2, const double *peak_sc);

/* init_scalar(n) (init_scalars.f90:47-79) for one scalar, without the final
 * ialias mask: what restart_read_parallel calls for scalars the restart file
 * lacks (restart_io.f90:426-429). */
int hit3d_init_scalar(const int64_t *n_scalar, const double *rn2, const double *peak);

/* io_write_4 for one field (io_write_4.f90:16-50): fields(n_field) with every
 * mode of k^2 > kmax^2 removed, transformed to physical space and rounded to
 * float32 into out(nx, ny, nz) (x fastest, no padding) -- the tmp4 array
 * write_tmp4 writes after its int32 (nx, ny, nz) header
 * (write_tmp4.f90:14-30).  Uses wrk(0) as work space. */
int hit3d_snapshot(const int64_t *n_field, float *out);

/* The turbulent viscosity of les_model 1 as the last LES right-hand side
 * left it, rounded to float32 into out(nx, ny, nz): the tmp4 = turb_visc
 * that io_write_4 writes as nu_t.<ext> when LES is on (io_write_4.f90:73-79).
 * Other models hold none (an error). */
int hit3d_turb_visc(float *out);

/* The generator's saved iy (RANDu.f:22-24) as the library holds it after its
 * last initialisation (0 before any). */
int hit3d_rng_state(int64_t *iy);

/* The Gaussian field of init_scalar_spectrum (init_scalars.f90:105-133) from
 * a generator whose saved iy is *iy_prev, in the reference layout; *iy_next
 * receives the saved iy afterwards.  Host-only, like hit3d_init_gaussian. */
int hit3d_scalar_gaussian(const int64_t *nx, const double *rn2, const int64_t *iy_prev,
                          double *out, int64_t *iy_next);

/* The saved iy after init_velocity's draws on a fresh generator
 * (random(-rn1) and 6 (nx+2) nx^2 draws, init_velocity.f90:67-103). */
int hit3d_velocity_rng_state(const int64_t *nx, const double *rn1, int64_t *iy);

/* Draws of the reference generator (RANDu.f:12-52): a fresh generator,
 * random(-seed), `skip` draws discarded, then `count` draws into out. */
int hit3d_random_draws(const double *seed, const int64_t *skip, const int64_t *count,
                       double *out);

/* The Gaussian field sqrt(-2 log a) sin(2 pi b) of velocity component c
 * (1..3) before its transform (init_velocity.f90:87-113), in the reference
 * layout, computed by the same host code hit3d_init_velocity uses. */
int hit3d_init_gaussian(const int64_t *nx, const double *rn1, const int64_t *c, double *out);

/* ialias (x_fftw.f90:298-310) as int8 in the reference layout (nx+2,ny,nz). */
int hit3d_get_ialias(int8_t *out);

/* Per-kernel-class GPU time, measured with HIP events on the library's
 * stream while enabled (no reference counterpart).  hit3d_timing_read fills
 * up to *cap entries of total milliseconds and launch counts; the order of
 * classes is given by hit3d_timing_names (comma-separated). */
int hit3d_timing_enable(const int64_t *on);
int hit3d_timing_read(const int64_t *cap, double *ms, int64_t *count);
int hit3d_timing_names(char *buf, const int64_t *len);

/* Waits for all GPU work of the library. */
int hit3d_synchronize(void);

/* Copies the last error message (NUL-terminated, truncated to len). */
int hit3d_last_error(char *buf, const int64_t *len);

#ifdef __cplusplus
}
#endif
#endif
