// Lagrangian particle tracking on the physical velocity (m_particles.f90).
#pragma once
#include "spectral.hpp"

namespace hit3d {

// Interpolate (u, v, w) -- physical slots -- at each particle (scheme 0:
// trilinear, m_particles.f90:815-1009; 1: tricubic CINT, :1017-1230), then
// advance xyzp (grid units, 1-based, (3, np) order) by Euler when
// uvwp(1) == -999 and AB2 otherwise, store the new velocity in uvwp and wrap
// the position periodically in x, y and z.
void launch_particles_move(const Geo& g, const double* u, const double* v, const double* w,
                           double* xyzp, double* uvwp, int64_t np, double dt, double dx,
                           int scheme, hipStream_t s);

}  // namespace hit3d
