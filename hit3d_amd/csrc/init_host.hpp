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

// Taylor-Green component c in the reference layout (init_velocity.f90:271-282).
void taylor_green_slot(int n, int c, double* out);

// Target spectrum normalised to unit sum (init_velocity.f90:193-220, type 2 added).
std::vector<double> target_spectrum(int kmax, int isp_type, double peak, double ir_exp);

}  // namespace hit3d
