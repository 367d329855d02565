// Learned subgrid-stress closure: a pointwise MLP 10 -> 64 -> 64 -> 6 (GELU)
// on bfloat16 matrix cores with FP32 accumulation, trained with Adam.  It has
// no counterpart in the reference; it stands in for the eddy viscosity of
// m_les.f90:1142-1215, which stays the fallback.
//
// Parameters live in one FP32 array (master copy) with the layout below; the
// matrix units read a bfloat16 copy of it.  W1 is stored [16][64] (input
// rows 10..15 zero), W3 [64][16] (output columns 6..15 zero).
#pragma once
#include <stdint.h>

#include "spectral.hpp"

namespace hit3d {

namespace mlp {
constexpr int NIN = 10, NH = 64, NOUT = 6, KIN = 16, KOUT = 16;
constexpr int OW1 = 0, OB1 = OW1 + KIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * KOUT, NPARAM = OB3 + KOUT;  // 6288
constexpr int TILE = 64;  // samples per workgroup
}  // namespace mlp

struct MlpNorm {
    float xmean[mlp::NIN], xstd[mlp::NIN], ymean[mlp::NOUT], ystd[mlp::NOUT];
};

// One forward/backward pass over `tiles` tiles of 64 samples starting at
// sample `start` (indices wrap modulo count): per-tile partial gradients
// (NPARAM floats each) and per-tile loss partials.  nb_total is the batch
// size the loss is averaged over.  X[count][10], T[count][6] standardised.
void launch_mlp_train_tiles(const float* X, const float* T, int64_t count, int64_t start, int tiles,
                            int64_t nb_total, const uint16_t* wb, const float* p, float* partial,
                            float* lossp, hipStream_t s);
// grads[k] = sum over tiles of partial[t][k] and *loss = sum of lossp, in a
// fixed order (deterministic): kReduceGroups contiguous tile groups, each in
// tile order, then the groups in order; red holds kReduceGroups * (NPARAM + 1)
// floats of group sums.
constexpr int kReduceGroups = 64;
void launch_mlp_reduce(const float* partial, const float* lossp, int tiles, float* red,
                       float* grads, float* loss, hipStream_t s);
// Adam step t (1-based) on the FP32 master copy, refreshing the bfloat16 copy.
void launch_mlp_adam(float* p, uint16_t* wb, const float* g, float* m, float* v, int t, float lr,
                     hipStream_t s);
// bfloat16 copy of the master parameters.
void launch_mlp_to_bf16(const float* p, uint16_t* wb, hipStream_t s);
// Forward pass of count standardised samples: Y[count][6] (standardised).
void launch_mlp_forward(const float* X, int64_t count, const uint16_t* wb, const float* p, float* Y,
                        hipStream_t s);
// Inference at every point of the grid from the nine physical gradient
// slots du_i/dx_j (order 11,12,13,21,...,33) and the filter width: writes
// the deviatoric stress tau_11, tau_12, tau_13, tau_22, tau_23, tau_33 into
// six physical slots; *bad counts non-finite outputs.  mix > 0 blends in
// the Smagorinsky stress -2 cd2 |S| S with that weight.
void launch_mlp_les(const Geo& g, const double* const grad[9], double delta, const MlpNorm& nrm,
                    const uint16_t* wb, const float* p, double* const tau[6], unsigned int* bad,
                    double mix, double cd2, hipStream_t s);
// Features of every grid point: X[p][10] = (du_i/dx_j, delta), raw.
void launch_sgs_features(const Geo& g, const double* const grad[9], double delta, float* X,
                         hipStream_t s);
// Box filter of ratio r (top-hat over r^3 blocks) of the physical velocity:
// ubar[3][m^3] and tau[6][m^3] = avg(u_i u_j) - ubar_i ubar_j, m = n / r,
// coarse index (z*m + y)*m + x.
void launch_box_filter(const Geo& g, const double* u, const double* v, const double* w, int r,
                       double* ubar, double* tau, hipStream_t s);

}  // namespace hit3d
