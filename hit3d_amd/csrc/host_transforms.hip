// Host sequencing of the 3-D transforms (xFFT3d, x_fftw.f90:852-1085) and
// their fused product / phase-shift / accumulate forms.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"
#include "passes.hpp"

namespace hit3d {

// ---------------------------------------------------------------------------
// transforms
// ---------------------------------------------------------------------------
void pass_y(double* slot, int dir) {
    Lib& L = lib();
    auto* c = reinterpret_cast<cpx<double>*>(slot);
    const int64_t n = L.n, P = L.g.P;
    Timed t(TC_FFT_Y);
    launch_cols<double>(L.n, dir, c, c, n * P, n * P, P, P, L.tw64, L.st);
}

// y and z passes of a forward transform whose x pass is done; ph = -1
// multiplies the spectral result by exp(-i theta) (the phase shift back).
void fft_forward_yz(double*& slot, int ph) {
    Lib& L = lib();
    const int64_t n = L.n, P = L.g.P;
    pass_y(slot, -1);
    {
        Timed t(TC_FFT_Z);
        launch_cols<double>(L.n, -1, reinterpret_cast<cpx<double>*>(slot),
                            reinterpret_cast<cpx<double>*>(L.spare), P, n * P, n * P, P, L.tw64,
                            L.st, ph, L.ps_tab);
    }
    std::swap(slot, L.spare);
}

void fft_forward(double*& slot, int ph) {
    Lib& L = lib();
    {
        Timed t(TC_FFT_X);
        launch_x_r2c<double>(L.n, slot, L.g.P, L.tw64, L.st);
    }
    fft_forward_yz(slot, ph);
}

// dst = forward transform of the physical product a*b (dst may be a or b),
// times exp(-i theta) when ph = -1
void fft_forward_product(const double* a, const double* b, double*& dst, int ph) {
    Lib& L = lib();
    {
        Timed t(TC_FFT_X);
        launch_x_r2c_mul<double>(L.n, a, b, dst, L.g.P, L.tw64, L.st);
    }
    fft_forward_yz(dst, ph);
}

// Forward transform of the physical product a*b, shifted back by exp(-i theta),
// accumulated into the RHS as rhs_j += -1/2 i k_axj (a*b)^ for the given
// targets instead of being stored; `work` is overwritten (x and y passes).
void fft_forward_product_acc(const double* a, const double* b, double* work, double* r0, int ax0,
                             bool mask0, double* r1, int ax1, bool mask1) {
    Lib& L = lib();
    const int64_t n = L.n, P = L.g.P;
    {
        Timed t(TC_FFT_X);
        launch_x_r2c_mul<double>(L.n, a, b, work, P, L.tw64, L.st);
    }
    pass_y(work, -1);
    ColAcc acc;
    acc.r[0] = reinterpret_cast<cpx<double>*>(r0);
    acc.ax[0] = ax0;
    acc.r[1] = reinterpret_cast<cpx<double>*>(r1);
    acc.ax[1] = ax1;
    acc.mask[0] = mask0;
    acc.mask[1] = mask1;
    Timed t(TC_FFT_Z);
    launch_cols<double>(L.n, -1, reinterpret_cast<cpx<double>*>(work),
                        reinterpret_cast<cpx<double>*>(L.spare), P, n * P, n * P, P, L.tw64, L.st,
                        -1, L.ps_tab, acc);
}

// dst = inverse transform of src (src unchanged unless it is dst), the
// spectral input multiplied by exp(+i theta) first when ph = +1
void fft_inverse_from(const double* src, double*& dst, int ph) {
    Lib& L = lib();
    const int64_t n = L.n, P = L.g.P;
    {
        Timed t(TC_FFT_Z);
        launch_cols<double>(L.n, 1, reinterpret_cast<const cpx<double>*>(src),
                            reinterpret_cast<cpx<double>*>(L.spare), n * P, P, P, n * P, L.tw64,
                            L.st, ph, L.ps_tab);
    }
    std::swap(dst, L.spare);
    pass_y(dst, 1);
    {
        Timed t(TC_FFT_X);
        const double norm = 1.0 / ((double)n * (double)n * (double)n);
        launch_x_c2r<double>(L.n, dst, P, L.tw64, norm, L.st);
    }
}

void fft_inverse(double*& slot) { fft_inverse_from(slot, slot, 0); }

// Batched forms for up to three slots: one launch per pass instead of one per
// slot, which is what the launch-bound small grids gain from (64^3 LES step
// 0.42 -> 0.27 ms); from 256^3 up the per-slot passes are faster (a 256^3
// phase-shift step took 0.4 ms longer batched), so only N <= 128 batches.
// The batches need two spares beyond the usual one, allocated on first use.
bool batch_spares() {
    Lib& L = lib();
    if (L.n > 128) return false;
    if (L.spares_b.empty()) {
        L.spares_b.assign(2, nullptr);
        for (auto& p : L.spares_b) p = dalloc(slot_bytes());
    }
    return true;
}

void fft_forward_batch(double** const* slots, int nb) {
    Lib& L = lib();
    if (nb == 1 || !batch_spares()) {
        for (int i = 0; i < nb; ++i) fft_forward(*slots[i]);
        return;
    }
    if (nb > 3) throw std::runtime_error("fft_forward_batch: at most 3 slots");
    const int64_t n = L.n, P = L.g.P;
    double* s[3];
    cpx<double>* c[3];
    cpx<double>* sp[3];
    double* spare[3] = {L.spare, L.spares_b[0], L.spares_b[1]};
    for (int i = 0; i < nb; ++i) {
        s[i] = *slots[i];
        c[i] = reinterpret_cast<cpx<double>*>(s[i]);
        sp[i] = reinterpret_cast<cpx<double>*>(spare[i]);
    }
    {
        Timed t(TC_FFT_X);
        launch_x_r2c_batch(L.n, s, nb, P, L.tw64, L.st);
    }
    {
        Timed t(TC_FFT_Y);
        launch_cols_batch(L.n, -1, c, c, nb, n * P, n * P, P, P, L.tw64, L.st);
    }
    {
        Timed t(TC_FFT_Z);
        launch_cols_batch(L.n, -1, c, sp, nb, P, n * P, n * P, P, L.tw64, L.st);
    }
    for (int i = 0; i < nb; ++i) std::swap(*slots[i], spare[i]);
    L.spare = spare[0];
    for (int i = 1; i < nb; ++i) L.spares_b[i - 1] = spare[i];
}

void fft_inverse_from_batch(const double* const* src, double** const* dst, int nb) {
    Lib& L = lib();
    if (nb == 1 || !batch_spares()) {
        for (int i = 0; i < nb; ++i) fft_inverse_from(src[i], *dst[i], 0);
        return;
    }
    if (nb > 3) throw std::runtime_error("fft_inverse_batch: at most 3 slots");
    const int64_t n = L.n, P = L.g.P;
    const cpx<double>* in[3];
    cpx<double>* sp[3];
    double* spare[3] = {L.spare, L.spares_b[0], L.spares_b[1]};
    for (int i = 0; i < nb; ++i) {
        in[i] = reinterpret_cast<const cpx<double>*>(src[i]);
        sp[i] = reinterpret_cast<cpx<double>*>(spare[i]);
    }
    {
        Timed t(TC_FFT_Z);
        launch_cols_batch(L.n, 1, in, sp, nb, n * P, P, P, n * P, L.tw64, L.st);
    }
    for (int i = 0; i < nb; ++i) std::swap(*dst[i], spare[i]);
    L.spare = spare[0];
    for (int i = 1; i < nb; ++i) L.spares_b[i - 1] = spare[i];
    double* d[3];
    cpx<double>* c[3];
    for (int i = 0; i < nb; ++i) {
        d[i] = *dst[i];
        c[i] = reinterpret_cast<cpx<double>*>(d[i]);
    }
    {
        Timed t(TC_FFT_Y);
        launch_cols_batch(L.n, 1, c, c, nb, n * P, n * P, P, P, L.tw64, L.st);
    }
    Timed t(TC_FFT_X);
    const double norm = 1.0 / ((double)n * (double)n * (double)n);
    launch_x_c2r_batch(L.n, d, nb, P, L.tw64, norm, L.st);
}

void fft_inverse_batch(double** const* slots, int nb) {
    const double* src[3];
    if (nb > 3) throw std::runtime_error("fft_inverse_batch: at most 3 slots");
    for (int i = 0; i < nb; ++i) src[i] = *slots[i];
    fft_inverse_from_batch(src, slots, nb);
}

}  // namespace hit3d
