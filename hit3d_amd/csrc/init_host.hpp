// Host-side parts of the seeded initial condition (init_velocity.f90).
#pragma once
#include <stdint.h>

#include <vector>

namespace hit3d {

// Draws of the reference generator (RANDu.f:12-52) from a fresh state:
// random(-seed), `skip` draws discarded, `count` draws to out.
void random_draws(double seed, int64_t skip, int64_t count, double* out);

// Gaussian field of velocity slot c (0..2) in the reference layout
// (pitch n+2): sqrt(-2 log a) sin(2 pi b) with a the draws of slot c and b
// those of slot c+3 of the six-slot fill (init_velocity.f90:87-113).
void gaussian_slot(int n, double rn1, int c, double* out);

// The generator's saved iy after init_velocity's draws (random(-rn1) and
// 6 (n+2) n^2 draws, one rank, fresh process).
int64_t velocity_rng_iy(int n, double rn1);

// init_scalar_spectrum (init_scalars.f90:84-133): random(-rn2) with the
// saved iy iy_prev, 2 (n+2) n^2 draws filling wrk1 then wrk2, and
// out = sqrt(-2 log wrk1) sin(2 pi wrk2) in the reference layout.  Returns
// the saved iy afterwards.
int64_t scalar_gaussian_slot(int n, double rn2, int64_t iy_prev, double* out);

// init_scalar_space (init_scalars.f90:278-383): physical scalar of type
// 11 (tanh slab in z), 12 (two tanh slabs in x) or 13 (sin waves in x), in
// the reference layout (pad entries zero).
void scalar_space_slot(int n, int ic_type, double peak, double* out);

// Scalar target spectrum (init_scalars.f90:170-196) normalised to unit sum:
// type 0 k^-5/3, 1 and 3 (k/k0)^3/k0 exp(-3k/k0), 2 von Karman.
std::vector<double> scalar_target_spectrum(int kmax, int64_t sc_type, double peak);

// Taylor-Green component c in the reference layout (init_velocity.f90:271-282).
void taylor_green_slot(int n, int c, double* out);

// Target spectrum normalised to unit sum (init_velocity.f90:193-220, type 2 added).
std::vector<double> target_spectrum(int kmax, int isp_type, double peak, double ir_exp);

}  // namespace hit3d
