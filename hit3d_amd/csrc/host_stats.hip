// Host sides of the statistics and checks that reduce on the device:
// scalar statistics (m_stats.f90:295-366) and divergence (pressure.f90:53-93).
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace hit3d {

double key_to_double(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// min and max of (a [+ b + c]) in physical space, stored to sum if given.
void minmax_to_host(const double* a, const double* b, const double* c, double* sum, double* mn,
                    double* mx) {
    Lib& L = lib();
    auto* keys = reinterpret_cast<unsigned long long*>(L.dscal + 8);
    {
        Timed t(TC_REDUCE);
        launch_minmax(L.g, a, b, c, sum, keys, L.st);
    }
    check(hipMemcpyAsync(L.hpin + 8, L.dscal + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    uint64_t k[2];
    std::memcpy(k, L.hpin + 8, sizeof k);
    *mn = key_to_double(k[0]);
    *mx = key_to_double(k[1]);
}

void sums_to_host(const double* a, const double* b, const double* c, double out[2]) {
    Lib& L = lib();
    {
        Timed t(TC_REDUCE);
        launch_phys_sums(L.g, a, b, c, L.dpart, L.dscal + 4, L.st);
    }
    check(hipMemcpyAsync(L.hpin + 4, L.dscal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    out[0] = L.hpin[4];
    out[1] = L.hpin[5];
}

// m_stats.f90:295-366 (stat_scalars) with get_scalar_spectra (:371-437):
// per scalar, 2 pe <|grad phi|^2>, mean, variance, min, max, and the variance
// spectrum.  wrk(0:4) are used as work space like the reference.
void do_stat_scalars(double* diss, double* mean, double* var, double* mn, double* mx, double* es) {
    Lib& L = lib();
    auto& W = L.W;
    const double fac = 1.0 / ((double)L.n * (double)L.n * (double)L.n);
    if ((int)L.pe.size() < L.ns) throw std::runtime_error("stat_scalars: set_physics first");
    for (int n = 1; n <= L.ns; ++n) {
        const double* phi = L.F[3 + n];
        if (es) {
            {
                Timed t(TC_REDUCE);
                launch_shell_sums1(L.g, phi, 1, L.kmax, L.dshell, L.st);
            }
            std::vector<double> e(L.kmax);
            check(hipMemcpyAsync(e.data(), L.dshell, L.kmax * sizeof(double),
                                 hipMemcpyDeviceToHost, L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            for (int k = 0; k < L.kmax; ++k) es[(int64_t)(n - 1) * L.kmax + k] = 2.0 * e[k];
        }
        copy_slot(W[0], phi);
        {
            Timed t(TC_POINTWISE);
            for (int a = 1; a <= 3; ++a) launch_derivative(L.g, W[0], W[a], a, L.st);
        }
        for (int a = 1; a <= 3; ++a) fft_inverse(W[a]);
        double s[2];
        sums_to_host(W[1], W[2], W[3], s);
        if (diss) diss[n - 1] = 2.0 * L.pe[n - 1] * s[0] * fac;
        fft_inverse(W[0]);
        sums_to_host(W[0], nullptr, nullptr, s);
        const double m1 = s[0] * fac;
        if (mean) mean[n - 1] = m1;
        if (var) var[n - 1] = s[1] * fac - m1 * m1;
        double lo, hi;
        minmax_to_host(W[0], nullptr, nullptr, nullptr, &lo, &hi);
        if (mn) mn[n - 1] = lo;
        if (mx) mx[n - 1] = hi;
    }
}

// pressure.f90:53-93: wrk(1:3) = fields(1:3); wrk(4:6) = d/dx_i of them in
// physical space; wrk(0) = their sum; min and max of wrk(0).
void do_divergence(double* dmin, double* dmax) {
    Lib& L = lib();
    auto& W = L.W;
    for (int c = 1; c <= 3; ++c) copy_slot(W[c], L.F[c]);
    {
        Timed t(TC_POINTWISE);
        for (int c = 1; c <= 3; ++c) launch_derivative(L.g, W[c], W[3 + c], c, L.st);
    }
    for (int c = 4; c <= 6; ++c) fft_inverse(W[c]);
    minmax_to_host(W[4], W[5], W[6], W[0], dmin, dmax);
}

}  // namespace hit3d
