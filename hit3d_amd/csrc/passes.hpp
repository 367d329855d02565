// Host-side launchers of the FFT axis passes (fft_passes.hip).
#pragma once
#include "common.hpp"

namespace hit3d {

// Twiddle tables: for every power of two L <= kMaxLen the entries
// exp(-2 pi i j / L), j < L, stored at offset L of one buffer.
constexpr int kMaxLen = 4096;

bool supported_n(int n);

// x pass, forward: each of the n*n rows (n reals, pitch 2P) -> n/2+1 complex
// values in place (real-to-complex via a half-length complex transform).
template <typename T>
void launch_x_r2c(int n, T* slot, int64_t P, const cpx<T>* tw, hipStream_t s);

// x pass, inverse: n/2+1 complex values per row -> n reals times `norm`.
template <typename T>
void launch_x_c2r(int n, T* slot, int64_t P, const cpx<T>* tw, T norm, hipStream_t s);

// Column pass: n complex values per column at stride s_in (s_out), for the
// n/2+1 columns kx of each of the n groups a; element e of column kx of
// group a is read at in[a*a_in + kx + e*s_in] and written at
// out[a*a_out + kx + e*s_out].  dir = -1 forward, +1 inverse.
// ph = +1 (inverse pass only): the spectral input (kx = column, kz =
// element, ky = group) is multiplied by exp(+i theta) on load; ph = -1
// (forward pass only): the spectral output by exp(-i theta) on store;
// theta = (kx+ky+kz) dx/2 from tab (cos, sin per kx+ky+kz+n).  FP64 only.
// acc (forward pass with ph = -1 only): instead of storing the result v,
// add -1/2 i k_ax[j] v to each non-null r[j] (same spectral layout as out),
// ax 1 = kx, 2 = ky, 3 = kz; where mask[j], entries with alias_count >= 2
// are set to zero instead (the phase-shift mask, rhs_velocity.f90:400-404,
// applied by the last accumulation into a target).
struct ColAcc {
    cpx<double>* r[2] = {nullptr, nullptr};
    int ax[2] = {0, 0};
    bool mask[2] = {false, false};
};
template <typename T>
void launch_cols(int n, int dir, const cpx<T>* in, cpx<T>* out, int64_t a_in, int64_t a_out,
                 int64_t s_in, int64_t s_out, const cpx<T>* tw, hipStream_t s, int ph = 0,
                 const double* tab = nullptr, ColAcc acc = ColAcc{});

// The same passes on nb <= 3 FP64 slots in one launch (slot i of the batch
// in the launch's z dimension); no phase or accumulate options.
void launch_x_r2c_batch(int n, double* const* slots, int nb, int64_t P, const cpx<double>* tw,
                        hipStream_t s);
void launch_x_c2r_batch(int n, double* const* slots, int nb, int64_t P, const cpx<double>* tw,
                        double norm, hipStream_t s);
void launch_cols_batch(int n, int dir, const cpx<double>* const* in, cpx<double>* const* out,
                       int nb, int64_t a_in, int64_t a_out, int64_t s_in, int64_t s_out,
                       const cpx<double>* tw, hipStream_t s);

// x pass, forward, of the pointwise product of two physical slots:
// out = r2c(a * b) row by row (out may alias a or b).
template <typename T>
void launch_x_r2c_mul(int n, const T* a, const T* b, T* out, int64_t P, const cpx<T>* tw,
                      hipStream_t s);

}  // namespace hit3d
