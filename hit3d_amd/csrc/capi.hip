// C ABI of the library (include/hit3d.h): device-resident state and the
// entry points that replace the reference's x_fftw / rhs_* / pressure /
// m_force / dealias_all subroutines.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "guard.hpp"
#include "hit3d.h"
#include "host.hpp"
#include "init_host.hpp"
#include "passes.hpp"

using namespace hit3d;

extern "C" {

int hit3d_init(const int64_t* nx, const int64_t* ny, const int64_t* nz_all,
               const int64_t* n_scalars, const int64_t* n_les, const int64_t* dealias) {
    return guarded([&] {
        // a previous init, complete or not, leaves nothing held
        free_all();
        Lib& L = lib();
        if (*nx != *ny || *nx != *nz_all)
            throw std::runtime_error("only cubic grids nx = ny = nz_all are supported");
        if (!supported_n((int)*nx))
            throw std::runtime_error("grid size must be a power of two in [8, 1024]");
        if (*n_scalars < 0 || *n_les < 0) throw std::runtime_error("negative field count");
        if (*dealias != 0 && *dealias != 1) throw std::runtime_error("dealias must be 0 or 1");
        L.n = (int)*nx;
        L.g.n = L.n;
        L.g.P = row_pitch(L.n);
        L.g.nc = (int64_t)L.n * L.n * L.g.P;
        L.ns = (int)*n_scalars;
        L.nles = (int)*n_les;
        L.nf = 3 + L.ns + L.nles;
        L.nw = std::max(6, 3 + L.ns + L.nles + 2);
        L.dealias = (int)*dealias;
        L.kmax = L.dealias == 0 ? L.n / 3 : (int)std::floor((double)L.n / 3.0 * std::sqrt(2.0));
        const double pi = 4.0 * std::atan(1.0);
        L.dx = 2.0 * pi / (double)L.n;
        check(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking), "hipStreamCreate");
        {
            int dev = 0;
            check(hipGetDevice(&dev), "hipGetDevice");
            check(hipDeviceGetAttribute(&L.ncu, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
        }
        static_scratch_setup();
        const size_t sb = slot_bytes();
        L.F.assign(L.nf + 1, nullptr);
        L.R.assign(L.nf + 1, nullptr);
        L.W.assign(L.nw + 1, nullptr);
        for (int i = 1; i <= L.nf; ++i) {
            L.F[i] = dalloc(sb);
            L.R[i] = dalloc(sb);
        }
        for (int i = 0; i <= L.nw; ++i) L.W[i] = dalloc(sb);
        L.spare = dalloc(sb);
        // twiddles exp(-2 pi i j / len) for every power of two len <= kMaxLen
        std::vector<cpx<double>> t64(2 * kMaxLen);
        std::vector<cpx<float>> t32(2 * kMaxLen);
        for (int len = 1; len <= kMaxLen; len *= 2)
            for (int j = 0; j < len; ++j) {
                const long double a = -2.0L * 3.14159265358979323846264338327950288L * j / len;
                t64[len + j] = {(double)cosl(a), (double)sinl(a)};
                t32[len + j] = {(float)cosl(a), (float)sinl(a)};
            }
        L.tw64 = reinterpret_cast<cpx<double>*>(dalloc(t64.size() * sizeof(cpx<double>)));
        L.tw32 = reinterpret_cast<cpx<float>*>(dalloc(t32.size() * sizeof(cpx<float>)));
        check(hipMemcpy(L.tw64, t64.data(), t64.size() * sizeof(cpx<double>), hipMemcpyHostToDevice),
              "hipMemcpy");
        check(hipMemcpy(L.tw32, t32.data(), t32.size() * sizeof(cpx<float>), hipMemcpyHostToDevice),
              "hipMemcpy");
        // phase-shift factors cos/sin(half*(kx+ky+kz)*dx), rhs_velocity.f90:248-249
        const int ntab = 3 * L.n;
        std::vector<double> tab(2 * ntab);
        for (int t = 0; t < ntab; ++t) {
            const double ks = (double)(t - L.n);
            const double a = 0.5 * ks * L.dx;
            tab[2 * t] = std::cos(a);
            tab[2 * t + 1] = std::sin(a);
        }
        L.ps_tab = dalloc(tab.size() * sizeof(double));
        check(hipMemcpy(L.ps_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy");
        L.dshell = dalloc((size_t)(L.n + 2) * sizeof(double) * 2);
        L.dpart = dalloc((size_t)L.n * 2 * sizeof(double));
        L.dmax = dalloc(std::max<size_t>(65536, (size_t)L.n * L.n) * sizeof(double));
        L.dscal = dalloc(16 * sizeof(double));
        check(hipHostMalloc((void**)&L.hpin, 16 * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        L.nu = 0.0;
        L.ready = true;
    });
}

int hit3d_set_physics(const double* nu, const int64_t* flow_type, const int64_t* kfmax,
                      const double* famp, const double* pe, const int64_t* scalar_type) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        clear_graphs();  // captured steps hold the old parameters as kernel arguments
        L.nu = *nu;
        L.flow_type = (int)*flow_type;
        L.kfmax = (int)*kfmax;
        L.famp = *famp;
        // m_force_init (m_force.f90:13-97) reads KFMAX for forced runs only
        if (L.flow_type == 1 && L.kfmax > L.kmax) throw std::runtime_error("kfmax exceeds kmax");
        L.pe.assign(L.ns, 0.0);
        L.stype.assign(L.ns, 0);
        L.reac.assign(L.ns, 0.0);
        for (int s = 0; s < L.ns; ++s) {
            if (!pe || !scalar_type) throw std::runtime_error("scalar parameters missing");
            L.pe[s] = pe[s];
            L.stype[s] = scalar_type[s];
            const int64_t rtype = L.stype[s] / 100;
            if (L.stype[s] < 0 || rtype > 3)
                throw std::runtime_error("scalar type: reaction digit must be 0..3");
            if (L.dealias == 1 && L.stype[s] > 300 && L.stype[s] != 311)
                throw std::runtime_error(
                    "The scalar type has a reaction rate that is not supported yet (phase "
                    "shifting supports 311 only)");
        }
    });
}

int hit3d_set_reaction(const double* reac_sc) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if ((int)L.reac.size() != L.ns) throw std::runtime_error("set_reaction: set_physics first");
        clear_graphs();
        for (int s = 0; s < L.ns; ++s) L.reac[s] = reac_sc[s];
    }, false);
}

int hit3d_finalize(void) {
    return guarded([&] { free_all(); });
}

int hit3d_put(const int64_t* array_id, const int64_t* slot, const double* host) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double* d = slot_ref(*array_id, *slot);
        const size_t w = (size_t)(L.n + 2) * sizeof(double);
        check(hipMemcpy2DAsync(d, (size_t)L.g.P * 2 * sizeof(double), host, w, w,
                               (size_t)L.n * L.n, hipMemcpyHostToDevice, L.st),
              "hipMemcpy2DAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_get(const int64_t* array_id, const int64_t* slot, double* host) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double* d = slot_ref(*array_id, *slot);
        const size_t w = (size_t)(L.n + 2) * sizeof(double);
        check(hipMemcpy2DAsync(host, w, d, (size_t)L.g.P * 2 * sizeof(double), w,
                               (size_t)L.n * L.n, hipMemcpyDeviceToHost, L.st),
              "hipMemcpy2DAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_copy(const int64_t* dst_array, const int64_t* dst_slot, const int64_t* src_array,
               const int64_t* src_slot) {
    return guarded([&] {
        require_ready();
        double* d = slot_ref(*dst_array, *dst_slot);
        double* s = slot_ref(*src_array, *src_slot);
        if (d != s) copy_slot(d, s);
    });
}

int hit3d_xfft3d(const int64_t* flag, const int64_t* n) {
    return guarded([&] {
        require_ready();
        double*& s = slot_ref(2, *n);
        if (*flag == 1) fft_forward(s);
        else if (*flag == -1) fft_inverse(s);
        else throw std::runtime_error("xFFT3d: flag must be +1 or -1");
    });
}

int hit3d_xfft3d_fields(const int64_t* flag, const int64_t* n) {
    return guarded([&] {
        require_ready();
        double*& s = slot_ref(1, *n);
        if (*flag == 1) fft_forward(s);
        else if (*flag == -1) fft_inverse(s);
        else throw std::runtime_error("xFFT3d_fields: flag must be +1 or -1");
    });
}

int hit3d_x_derivative(const int64_t* n, const int64_t* axis, const int64_t* nto) {
    return guarded([&] {
        require_ready();
        if (*axis < 1 || *axis > 3) throw std::runtime_error("x_derivative: wrong value of axis");
        Lib& L = lib();
        double* a = slot_ref(2, *n);
        double* b = slot_ref(2, *nto);
        Timed t(TC_POINTWISE);
        launch_derivative(L.g, a, b, (int)*axis, L.st);
    });
}

int hit3d_dealias_all(void) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        Timed t(TC_POINTWISE);
        for (int c = 1; c <= 3 + L.ns; ++c) launch_truncate(L.g, L.F[c], L.kmax, L.st);
    });
}

int hit3d_rhs_scalars(const int64_t* int_scalars) {
    return guarded([&] {
        require_ready();
        do_rhs_scalars(*int_scalars != 0);
    });
}

int hit3d_rhs_velocity(const double* dt, double* courant) {
    return guarded([&] {
        require_ready();
        do_rhs_velocity();
        const double m = read_courant_max();
        if (courant) *courant = m * (*dt) / lib().dx;
    });
}

int hit3d_force_velocity(double* energy) {
    return guarded([&] {
        require_ready();
        do_force(energy);
    });
}

int hit3d_advance(const int64_t* which, int64_t* first_order, const double* dt) {
    return guarded([&] {
        require_ready();
        do_advance((int)*which, *first_order != 0, *dt);
        *first_order = 0;
    });
}

int hit3d_pressure(void) {
    return guarded([&] {
        require_ready();
        do_pressure();
    });
}

int hit3d_step(const double* dt, int64_t* fov, int64_t* fos, const int64_t* int_scalars,
               const int64_t* precision, double* courant, double* energy) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const bool mixed = *precision == 1;
        if (*precision != 0 && *precision != 1)
            throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        const bool scal = (*int_scalars != 0) && L.ns > 0;
        const bool fused = !scal && L.fuse && L.les_model == 0 && fused_step_supported(mixed);
        if (!fused && mixed) throw std::runtime_error("mixed precision needs the fused path");
        // the deviatoric form also runs in the fused phase-shift velocity step
        const bool ps_dev = L.nlform == 2 && L.dealias == 1 && L.fuse && fused_ps_supported();
        if (!fused && L.nlform != 0 && !ps_dev)
            throw std::runtime_error(
                "the rotational and deviatoric forms run in the fused steps only (2/3 rule: no "
                "active scalars, no LES; the deviatoric form also with the phase shift; fusion "
                "on)");
        const bool first_v = *fov != 0, first_s = *fos != 0;
        // the step size goes to the device (dscal[12]) ahead of the step's
        // work, so a captured step replays with the current dt
        L.hpin[12] = *dt;
        check(hipMemcpyAsync(L.dscal + 12, L.hpin + 12, sizeof(double), hipMemcpyHostToDevice,
                             L.st),
              "hipMemcpyAsync");
        // The out-of-place passes swap work slots with the spare; fields and
        // rhs_old never move.  wrk is scratch of the step (undefined after
        // it), so the work-slot labels are put back afterwards: every step
        // then starts from the same pointers and one captured graph serves
        // them all.
        // scratch the step allocates on first use, before the pointer snapshot
        if (L.les_model == 11) mlp_alloc_scratch();
        if (!fused) batch_spares();
        if ((scal || (L.fuse && fused_ps_supported())) && L.dealias == 1) keep_slots();
        // the mixed step keeps rhs_old(1:3) in FP32, every other step in FP64
        rhs_old_format(mixed);
        const std::vector<void*> s0 = slot_state();
        // the learned closure checks its output on the host mid-step
        if (L.graph_mode && !L.timing && L.les_model != 11)
            step_graph(fused, mixed, scal, first_v, first_s);
        else
            step_body(fused, mixed, scal, first_v, first_s);
        restore_work_slots(s0);
        *fov = 0;
        if (scal) *fos = 0;
        // Courant max (dscal[0], order-free bits of a non-negative double)
        // and band energy (dscal[1]) back in one copy
        check(hipMemcpyAsync(L.hpin, L.dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        double cmax;
        std::memcpy(&cmax, L.hpin, sizeof(cmax));
        if (energy && L.flow_type == 1) *energy = L.hpin[1];
        if (courant) *courant = cmax * (*dt) / L.dx;
    });
}

int hit3d_stats(double* e_spec, double* enstrophy) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        {
            Timed t(TC_REDUCE);
            launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kmax, L.dshell, L.st);
            launch_enstrophy(L.g, L.F[1], L.F[2], L.F[3], L.dpart, L.dscal + 2, L.st);
        }
        if (e_spec)
            check(hipMemcpyAsync(e_spec, L.dshell, (size_t)L.kmax * sizeof(double),
                                 hipMemcpyDeviceToHost, L.st),
                  "hipMemcpyAsync");
        check(hipMemcpyAsync(L.hpin + 2, L.dscal + 2, sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        if (enstrophy) *enstrophy = L.hpin[2];
    });
}

int hit3d_particles_set(const int64_t* np, const double* xyzp, const double* uvwp) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int64_t cnt = *np;
        if (cnt < 0) throw std::runtime_error("particles_set: negative particle count");
        const double hi = (double)(L.n + 1);
        for (int64_t i = 0; i < 3 * cnt; ++i)
            if (!(xyzp[i] >= 1.0 && xyzp[i] < hi))
                throw std::runtime_error("particles_set: position outside [1, n+1) grid units");
        if (cnt > L.np_cap) {
            if (L.xyzp) check(hipFree(L.xyzp), "hipFree");
            if (L.uvwp) check(hipFree(L.uvwp), "hipFree");
            L.xyzp = L.uvwp = nullptr;
            L.np_cap = 0;
            L.xyzp = static_cast<double*>(dalloc((size_t)cnt * 3 * sizeof(double)));
            L.uvwp = static_cast<double*>(dalloc((size_t)cnt * 3 * sizeof(double)));
            L.np_cap = cnt;
        }
        L.np = cnt;
        if (cnt == 0) return;
        const size_t b = (size_t)cnt * 3 * sizeof(double);
        check(hipMemcpyAsync(L.xyzp, xyzp, b, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        if (uvwp) {
            check(hipMemcpyAsync(L.uvwp, uvwp, b, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        } else {  // particles_init: uvwp = -999 (m_particles.f90:327)
            std::vector<double> fill((size_t)cnt * 3, -999.0);
            check(hipMemcpyAsync(L.uvwp, fill.data(), b, hipMemcpyHostToDevice, L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        }
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_particles_get(double* xyzp, double* uvwp) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const size_t b = (size_t)L.np * 3 * sizeof(double);
        if (b == 0) return;
        if (xyzp) check(hipMemcpyAsync(xyzp, L.xyzp, b, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        if (uvwp) check(hipMemcpyAsync(uvwp, L.uvwp, b, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_fields_to_parts(void) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        for (int c = 1; c <= 3; ++c) {
            copy_slot(L.W[c], L.F[c]);
            fft_inverse(L.W[c]);
        }
    });
}

int hit3d_particles_move(const double* dt, const int64_t* scheme) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*scheme != 0 && *scheme != 1)
            throw std::runtime_error("particles_tracking_scheme must be 0 (trilinear) or 1 (CINT)");
        if (*scheme == 1 && L.n < 4) throw std::runtime_error("CINT needs n >= 4");
        Timed t(TC_PARTICLES);
        launch_particles_move(L.g, L.W[1], L.W[2], L.W[3], L.xyzp, L.uvwp, L.np, *dt, L.dx,
                              (int)*scheme, L.st);
    });
}

int hit3d_stat_scalars(double* sc_diss, double* mean, double* variance, double* sc_min,
                       double* sc_max, double* es_spec) {
    return guarded([&] {
        require_ready();
        do_stat_scalars(sc_diss, mean, variance, sc_min, sc_max, es_spec);
    });
}

int hit3d_skewness(double* skew) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double tot = 0.0;
        for (int c = 1; c <= 3; ++c) {
            {
                Timed t(TC_POINTWISE);
                launch_derivative(L.g, L.F[c], L.W[0], c, L.st);
            }
            fft_inverse(L.W[0]);
            {
                Timed t(TC_REDUCE);
                launch_phys_moments23(L.g, L.W[0], L.dpart, L.dscal + 4, L.st);
            }
            check(hipMemcpyAsync(L.hpin + 4, L.dscal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                 L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            const double n3 = (double)L.n * L.n * L.n;
            const double m2 = L.hpin[4] / n3, m3 = L.hpin[5] / n3;
            tot += m3 / std::pow(m2, 1.5);
        }
        *skew = tot / 3.0;
    });
}

int hit3d_divergence(double* dmin, double* dmax) {
    return guarded([&] {
        require_ready();
        do_divergence(dmin, dmax);
    });
}

int hit3d_init_velocity(const double* rn1, const int64_t* isp_type, const int64_t* ir_exp,
                        const double* peak) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int n = L.n;
        const size_t hsz = (size_t)(n + 2) * n * n;
        std::vector<double> host(hsz);
        if (*isp_type == -1) {  // init_velocity.f90:252-293
            for (int c = 0; c < 3; ++c) {
                taylor_green_slot(n, c, host.data());
                const int64_t a = 2, s = c + 1;
                if (hit3d_put(&a, &s, host.data()) != 0) throw std::runtime_error(L.err);
                fft_forward(L.W[c + 1]);
                copy_slot(L.F[c + 1], L.W[c + 1]);
            }
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            return;
        }
        // init_velocity.f90:67-120: uniforms, Box-Muller, forward transforms
        for (int c = 0; c < 3; ++c) {
            gaussian_slot(n, *rn1, c, host.data());
            const int64_t a = 2, s = c + 1;
            if (hit3d_put(&a, &s, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.W[c + 1]);
        }
        L.rng_iy = velocity_rng_iy(n, *rn1);
        // :123-149 curl on shells 1..kmax
        launch_init_curl(L.g, L.W[1], L.W[2], L.W[3], L.F[1], L.F[2], L.F[3], L.kmax, L.st);
        // :156-186 shell spectrum
        launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kmax, L.dshell, L.st);
        std::vector<double> e(L.kmax);
        check(hipMemcpyAsync(e.data(), L.dshell, L.kmax * sizeof(double), hipMemcpyDeviceToHost,
                             L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        // :192-237 target spectrum and per-shell ratio
        std::vector<double> tgt = target_spectrum(L.kmax, (int)*isp_type, *peak, (double)*ir_exp);
        std::vector<double> ratio(L.kmax + 1, -1.0);
        for (int k = 1; k <= L.kmax; ++k)
            if (e[k - 1] > 0.0) ratio[k] = std::sqrt(tgt[k - 1] / e[k - 1]);
        check(hipMemcpyAsync(L.dshell, ratio.data(), ratio.size() * sizeof(double),
                             hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        for (int c = 1; c <= 3; ++c) launch_init_rescale(L.g, L.F[c], L.dshell, L.kmax, L.st);
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_init_scalar(const int64_t* n_scalar, const double* rn2, const double* peak) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int s = (int)*n_scalar;
        if ((int)L.stype.size() != L.ns) throw std::runtime_error("init_scalar: set_physics first");
        if (s < 1 || s > L.ns) throw std::runtime_error("init_scalar: scalar number out of range");
        const int n = L.n;
        std::vector<double> host((size_t)(n + 2) * n * n);
        const int64_t st = L.stype[s - 1];
        const int ic = (int)(st - (st / 100) * 100);
        double* F = L.F[3 + s];
        if (ic == 0) {
            zero_slot(F);
        } else if (ic < 10) {  // init_scalar_spectrum, :84-270
            if (st < 1 || st > 3)
                throw std::runtime_error("INIT_SCALARS: WRONG INITIAL SPECTRUM TYPE");
            L.rng_iy = scalar_gaussian_slot(n, *rn2, L.rng_iy, host.data());
            const int64_t a = 2, w3 = 3;
            if (hit3d_put(&a, &w3, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.W[3]);
            {
                Timed t(TC_REDUCE);
                launch_shell_sums1(L.g, L.W[3], 1, L.kmax, L.dshell, L.st);
            }
            std::vector<double> e(L.kmax);
            check(hipMemcpyAsync(e.data(), L.dshell, L.kmax * sizeof(double),
                                 hipMemcpyDeviceToHost, L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            const std::vector<double> tgt = scalar_target_spectrum(L.kmax, st, *peak);
            std::vector<double> ratio(L.kmax + 1, -1.0);
            for (int k = 1; k <= L.kmax; ++k)
                if (e[k - 1] > 0.0) ratio[k] = std::sqrt(tgt[k - 1] / e[k - 1]);
            check(hipMemcpyAsync(L.dshell, ratio.data(), ratio.size() * sizeof(double),
                                 hipMemcpyHostToDevice, L.st),
                  "hipMemcpyAsync");
            copy_slot(F, L.W[3]);
            {
                Timed t(TC_POINTWISE);
                launch_init_rescale(L.g, F, L.dshell, L.kmax, L.st);
            }
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            if (st == 3) {  // double-delta PDF, :239-262
                copy_slot(L.W[0], F);
                fft_inverse(L.W[0]);
                {
                    Timed t(TC_POINTWISE);
                    launch_double_delta(L.g, L.W[0], L.st);
                }
                fft_forward(L.W[0]);
                {
                    Timed t(TC_POINTWISE);
                    launch_shell_cut(L.g, L.W[0], L.kmax * 2 / 3, L.st);
                }
                copy_slot(F, L.W[0]);
            }
        } else {  // init_scalar_space, :278-383
            scalar_space_slot(n, ic, *peak, host.data());
            const int64_t a = 1, slot = 3 + s;
            if (hit3d_put(&a, &slot, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.F[3 + s]);  // may exchange the slot with the spare buffer
            if (ic == 11 || ic == 12)  // fields(1,1,1,3+n) = 0
                check(hipMemsetAsync(L.F[3 + s], 0, sizeof(double), L.st), "hipMemsetAsync");
        }
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_init_scalars(const double* rn2, const double* peak_sc) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        for (int64_t s = 1; s <= L.ns; ++s)  // init_scalar, init_scalars.f90:20-22
            if (hit3d_init_scalar(&s, rn2, peak_sc + (s - 1)) != 0) throw std::runtime_error(L.err);
        Timed t(TC_POINTWISE);  // :27-36 zero ialias > 0 on every scalar
        for (int s = 1; s <= L.ns; ++s) launch_mask_alias(L.g, L.F[3 + s], 0, L.st);
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_snapshot(const int64_t* n_field, float* out) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int64_t f = *n_field;
        if (f < 1 || f > L.nf) throw std::runtime_error("snapshot: field number out of range");
        copy_slot(L.W[0], L.F[f]);
        {
            Timed t(TC_POINTWISE);
            launch_sphere_cut(L.g, L.W[0], L.kmax, L.st);
        }
        fft_inverse(L.W[0]);
        float* stage = reinterpret_cast<float*>(L.spare);  // free between transforms
        {
            Timed t(TC_POINTWISE);
            launch_pack_f32(L.g, L.W[0], stage, L.st);
        }
        check(hipMemcpyAsync(out, stage, sizeof(float) * (size_t)L.n * L.n * L.n,
                             hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_turb_visc(float* out) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (L.les_model != 1 || !L.tv)
            throw std::runtime_error("turb_visc: only les_model 1 holds a turbulent viscosity");
        float* stage = reinterpret_cast<float*>(L.spare);  // free between transforms
        {
            Timed t(TC_POINTWISE);
            launch_pack_f32(L.g, L.tv, stage, L.st);
        }
        check(hipMemcpyAsync(out, stage, sizeof(float) * (size_t)L.n * L.n * L.n,
                             hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_rng_state(int64_t* iy) {
    return guarded([&] { *iy = lib().rng_iy; }, false);
}

int hit3d_scalar_gaussian(const int64_t* nx, const double* rn2, const int64_t* iy_prev,
                          double* out, int64_t* iy_next) {
    return guarded([&] {
        if (*nx < 2) throw std::runtime_error("bad grid size");
        const int64_t iy = scalar_gaussian_slot((int)*nx, *rn2, *iy_prev, out);
        if (iy_next) *iy_next = iy;
    }, false);
}

int hit3d_velocity_rng_state(const int64_t* nx, const double* rn1, int64_t* iy) {
    return guarded([&] {
        if (*nx < 2) throw std::runtime_error("bad grid size");
        *iy = velocity_rng_iy((int)*nx, *rn1);
    }, false);
}

int hit3d_random_draws(const double* seed, const int64_t* skip, const int64_t* count,
                       double* out) {
    return guarded([&] {
        if (*skip < 0 || *count < 0) throw std::runtime_error("negative draw count");
        random_draws(*seed, *skip, *count, out);
    }, false);
}

int hit3d_init_gaussian(const int64_t* nx, const double* rn1, const int64_t* c, double* out) {
    return guarded([&] {
        if (*c < 1 || *c > 3) throw std::runtime_error("component must be 1..3");
        if (*nx < 2) throw std::runtime_error("bad grid size");
        gaussian_slot((int)*nx, *rn1, (int)*c - 1, out);
    }, false);
}

int hit3d_get_ialias(int8_t* out) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const size_t bytes = (size_t)L.n * L.n * (L.n + 2);
        DevBuf d(bytes);
        launch_ialias(L.g, d.as<int8_t>(), L.st);
        check(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_set_les(const int64_t* les_model, const double* c_smag) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*les_model != 0 && *les_model != 1 && *les_model != 11)
            throw std::runtime_error(
                "only les_model 0 (none), 1 (Smagorinsky) and 11 (learned closure) are supported");
        if (*les_model == 11 && !(L.mlp.have_params && L.mlp.have_norm))
            throw std::runtime_error("les_model 11: set the closure's parameters and scaling first");
        if (*les_model == 11 && L.ns > 0)
            throw std::runtime_error("les_model 11 closes the velocity only (no scalars)");
        if (*les_model != 0 && L.dealias != 0)
            throw std::runtime_error("LES needs the 2/3 rule (m_parameters.f90:415-420)");
        clear_graphs();
        L.les_model = (int)*les_model;
        L.c_smag = *c_smag;
        if (L.les_model != 0 && !L.tv) L.tv = dalloc(slot_bytes());
        if (L.les_model == 11) L.mlp.delta = L.dx;  // filter width of the LES grid
    });
}

int hit3d_set_nonlinear(const int64_t* form) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*form < 0 || *form > 2)
            throw std::runtime_error(
                "nonlinear form must be 0 (divergence, the reference's), 1 (rotational) or 2 "
                "(deviatoric)");
        if (*form == 1 && L.dealias != 0)
            throw std::runtime_error("the rotational form is built for the 2/3 rule (dealias 0)");
        clear_graphs();  // captured steps hold the form in their kernels
        L.nlform = (int)*form;
    }, false);
}

int hit3d_set_fusion(const int64_t* on) {
    return guarded([&] {
        require_ready();
        clear_graphs();
        lib().fuse = *on != 0;
    });
}

int hit3d_set_graphs(const int64_t* on) {
    return guarded([&] {
        require_ready();
        clear_graphs();
        lib().graph_mode = *on != 0;
    });
}

int hit3d_graph_stats(int64_t* captured, int64_t* replays) {
    return guarded([&] {
        require_ready();
        if (captured) *captured = (int64_t)lib().graphs.size();
        if (replays) *replays = lib().graph_replays;
    });
}

int hit3d_timing_enable(const int64_t* on) {
    return guarded([&] {
        Lib& L = lib();
        collect_timers();
        L.timing = *on != 0;
        for (int i = 0; i < TC_COUNT; ++i) {
            L.tms[i] = 0.0;
            L.tcnt[i] = 0;
        }
    });
}

int hit3d_timing_read(const int64_t* cap, double* ms, int64_t* count) {
    return guarded([&] {
        Lib& L = lib();
        collect_timers();
        for (int i = 0; i < TC_COUNT && i < *cap; ++i) {
            ms[i] = L.tms[i];
            count[i] = L.tcnt[i];
        }
    });
}

int hit3d_timing_names(char* buf, const int64_t* len) {
    return guarded([&] {
        std::strncpy(buf, kTimerNames, (size_t)*len);
        if (*len > 0) buf[*len - 1] = '\0';
    }, false);
}

int hit3d_synchronize(void) {
    return guarded([&] {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipGetLastError(), "kernel launch");
    });
}

int hit3d_last_error(char* buf, const int64_t* len) {
    if (*len <= 0) return 0;
    std::strncpy(buf, lib().err.c_str(), (size_t)*len);
    buf[*len - 1] = '\0';
    return 0;
}

}  // extern "C"
