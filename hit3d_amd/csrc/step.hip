// Fused time step (filled in below); the unfused sequence in capi.hip is the
// fallback for configurations the fused pipeline does not cover.
#include "state.hpp"

namespace hit3d {

bool fused_step_supported(bool) { return false; }

void fused_step(double, bool, bool, double*) {}

}  // namespace hit3d
