// Host side of the seeded initial condition.
//
// The reference generator (RANDu.f:12-52) re-derives its first LCG state
// from the argument on every call, so after random(-seed) every later call
// random(seed) feeds the same constant c = lcg1(nint(seed/987)) into the
// 32-slot shuffle table; only the second LCG (idum2, multiplier ia2 mod im2)
// evolves.  Once every table slot holds c, draw m returns
//     am * mod1(c - idum2_m),   idum2_m = ia2^m * idum2_0 (mod im2),
// with mod1(x) = x + (im1 - 1) when x < 1.  Draws are therefore produced by a
// short exact serial warm-up (until the table is uniform) followed by
// independent chunks, each starting from idum2 advanced by modular
// exponentiation, on all host threads.  Box-Muller uses the C library's
// log/sin/sqrt, the functions a gfortran build of init_velocity.f90:113
// calls, so the Gaussian field is bit-identical to the reference's.
#include "init_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

namespace hit3d {

namespace {

constexpr int64_t IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1;
constexpr int64_t IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774;
constexpr int64_t IR1 = 12211, IR2 = 3791, NTAB = 32, NDIV = 1 + IMM1 / NTAB;

inline double out_of(int64_t iy) {
    const double am = 1.0 / (double)IM1;
    const double rnmx = 1.0 - 1.2e-7;
    const double r = am * (double)iy;
    return r < rnmx ? r : rnmx;
}

inline int64_t lcg1(int64_t x) {  // Schrage: IA1*x mod IM1
    const int64_t k = x / IQ1;
    x = IA1 * (x - k * IQ1) - k * IR1;
    return x < 0 ? x + IM1 : x;
}

inline int64_t lcg2(int64_t x) {  // Schrage: IA2*x mod IM2
    const int64_t k = x / IQ2;
    x = IA2 * (x - k * IQ2) - k * IR2;
    return x < 0 ? x + IM2 : x;
}

int64_t powmod(int64_t b, int64_t e, int64_t m) {
    unsigned __int128 r = 1, x = (unsigned __int128)(b % m);
    while (e > 0) {
        if (e & 1) r = (r * x) % (unsigned __int128)m;
        x = (x * x) % (unsigned __int128)m;
        e >>= 1;
    }
    return (int64_t)r;
}

// Exact generator state, stepped one call at a time.
struct Gen {
    int64_t idum2 = 123456789, iy = 0, iv[NTAB] = {0};
    int64_t c = 0;  // lcg1 of the seed, fed on every non-initialising call

    double call(double idum3) {
        int64_t idum = std::llround(idum3 / 987.0);
        if (idum <= 0) {
            idum = std::max(-idum, (int64_t)1);
            idum2 = idum;
            for (int64_t j = NTAB + 8; j >= 1; --j) {
                idum = lcg1(idum);
                if (j <= NTAB) iv[j - 1] = idum;
            }
        }
        idum = lcg1(idum);
        idum2 = lcg2(idum2);
        const int64_t j = 1 + iy / NDIV;
        iy = iv[j - 1] - idum2;
        iv[j - 1] = idum;
        if (iy < 1) iy += IMM1;
        return out_of(iy);
    }
    bool uniform() const {
        for (int i = 0; i < NTAB; ++i)
            if (iv[i] != c) return false;
        return true;
    }
};

int hw_threads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 4;
    return (int)std::min(h, 16u);
}

// Fills out[i] = f(draw index first+i) for `count` draws with the closed form,
// given idum2 at draw index `base` (the state after `base` draws).
template <typename F>
void closed_form(int64_t c, int64_t idum2_base, int64_t base, int64_t first, int64_t count, F&& f) {
    const int nt = hw_threads();
    const int64_t chunk = (count + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int64_t lo = first + t * chunk;
        const int64_t hi = std::min(first + count, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=, &f] {
            // idum2 after draw lo (draw numbers count from 1 after the init call)
            int64_t s = (int64_t)(((unsigned __int128)powmod(IA2, lo + 1 - base, IM2) *
                                   (unsigned __int128)idum2_base) %
                                  (unsigned __int128)IM2);
            for (int64_t i = lo; i < hi; ++i) {
                int64_t iy = c - s;
                if (iy < 1) iy += IMM1;
                f(i, out_of(iy));
                s = lcg2(s);
            }
        });
    }
    for (auto& x : th) x.join();
}

// Generates draws [first, first+count) (0-based, after random(-seed)) and
// passes each to f(index, value).  iy0 is the generator's saved iy before the
// initialising call (0 in a fresh process; random(-seed) refills the table
// and resets idum2 but keeps iy, RANDu.f:28-37).  Returns iy after draw
// index total-1 (after the initialising call itself when total == 0).
template <typename F>
int64_t draws(double seed, int64_t first, int64_t count, F&& f, int64_t iy0 = 0,
              int64_t total = -1) {
    Gen g;
    g.iy = iy0;
    g.call(-seed);
    g.c = lcg1(std::llround(seed / 987.0));
    if (std::llround(seed / 987.0) <= 0) throw std::runtime_error("seed must be >= 494");
    int64_t iy_total = total == 0 ? g.iy : -1;
    // serial warm-up until the shuffle table is uniform
    int64_t done = 0;  // draws made so far
    const int64_t cap = 1 << 20;
    while (!g.uniform()) {
        const double v = g.call(seed);
        if (done >= first && done < first + count) f(done, v);
        ++done;
        if (done == total) iy_total = g.iy;
        if (done > cap) throw std::runtime_error("random: shuffle table did not converge");
    }
    // after `done` draws idum2 == g.idum2; the next draw (index done) uses lcg2 of it
    if (total > done) {
        const int64_t sv = (int64_t)(((unsigned __int128)powmod(IA2, total - done, IM2) *
                                      (unsigned __int128)g.idum2) %
                                     (unsigned __int128)IM2);
        iy_total = g.c - sv;
        if (iy_total < 1) iy_total += IMM1;
    }
    const int64_t lo = std::max(first, done);
    const int64_t hi = first + count;
    if (hi > lo)  // draw index i uses ia2^(i-done+1)*g.idum2
        closed_form(g.c, g.idum2, done, lo, hi - lo, f);
    return iy_total;
}

}  // namespace

void random_draws(double seed, int64_t skip, int64_t count, double* out) {
    draws(seed, skip, count, [&](int64_t i, double v) { out[i - skip] = v; });
}

void gaussian_slot(int n, double rn1, int c, double* out) {
    const int64_t per = (int64_t)(n + 2) * n * n;
    std::vector<double> a(per);
    draws(rn1, c * per, per, [&](int64_t i, double v) { a[i - c * per] = v; });
    const double pi = 4.0 * std::atan(1.0);
    const double two_pi = 2.0 * pi;
    // b draws of slot c+3, combined directly into out
    draws(rn1, (c + 3) * per, per, [&](int64_t i, double v) {
        const int64_t e = i - (c + 3) * per;
        out[e] = std::sqrt(-2.0 * std::log(a[e])) * std::sin(two_pi * v);
    });
}

int64_t velocity_rng_iy(int n, double rn1) {
    const int64_t per = (int64_t)(n + 2) * n * n;
    return draws(rn1, 0, 0, [](int64_t, double) {}, 0, 6 * per);
}

int64_t scalar_gaussian_slot(int n, double rn2, int64_t iy_prev, double* out) {
    const int64_t per = (int64_t)(n + 2) * n * n;
    std::vector<double> a(per);
    const int64_t iy = draws(rn2, 0, per, [&](int64_t i, double v) { a[i] = v; }, iy_prev,
                             2 * per);
    const double pi = 4.0 * std::atan(1.0);
    const double two_pi = 2.0 * pi;
    draws(rn2, per, per, [&](int64_t i, double v) {
        const int64_t e = i - per;
        out[e] = std::sqrt(-2.0 * std::log(a[e])) * std::sin(two_pi * v);
    }, iy_prev);
    return iy;
}

void scalar_space_slot(int n, int ic_type, double peak, double* out) {
    const double pi = 4.0 * std::atan(1.0);
    const double dx = 2.0 * pi / (double)n;
    const int64_t row = n + 2;
    std::fill(out, out + row * n * n, 0.0);
    if (ic_type == 11) {  // single slab in z, init_scalars.f90:311-330
        const double h = std::max(2. * dx, 2.0 * pi / peak);
        for (int k = 0; k < n; ++k) {
            const double zloc = (double)k * dx;
            const double v = std::tanh((zloc - pi * 0.5) / h) - std::tanh((zloc - pi * 1.5) / h) - 1.0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) out[((int64_t)k * n + j) * row + i] = v;
        }
    } else if (ic_type == 12) {  // two slabs in x, :335-356
        const double h = std::max(8. * dx, pi / 8.0);
        for (int i = 0; i < n; ++i) {
            const double xx = (double)i * dx;
            const double v = std::tanh((xx - pi * 0.25) / h) - std::tanh((xx - pi * 0.75) / h) +
                             std::tanh((xx - pi * 1.25) / h) - std::tanh((xx - pi * 1.75) / h);
            for (int64_t r = 0; r < (int64_t)n * n; ++r) out[r * row + i] = v - 1.0;
        }
    } else if (ic_type == 13) {  // sinusoidal waves in x, :362-372
        const double pk = std::min(peak, (double)(n / 8));
        for (int i = 0; i < n; ++i) {
            const double v = std::sin(pk * dx * (double)i);
            for (int64_t r = 0; r < (int64_t)n * n; ++r) out[r * row + i] = v;
        }
    } else {
        throw std::runtime_error("INIT_SCALARS: UNEXPECTED SCALAR TYPE");
    }
}

std::vector<double> scalar_target_spectrum(int kmax, int64_t sc_type, double peak) {
    std::vector<double> e(kmax);
    const double pi = 4.0 * std::atan(1.0);
    for (int k = 1; k <= kmax; ++k) {
        const double wmag = (double)k;
        const double ratio = wmag / peak;
        double v;
        if (sc_type == 0) v = std::pow(wmag, -5.0 / 3.0);
        else if (sc_type == 1 || sc_type == 3) v = std::pow(ratio, 3) / peak * std::exp(-3.0 * ratio);
        else if (sc_type == 2) {
            const double fac = 2.0 * pi * ratio;
            v = std::pow(fac, 4) / std::pow(1.0 + fac * fac, 3);
        } else {
            throw std::runtime_error("INIT_SCALARS: WRONG INITIAL SPECTRUM TYPE");
        }
        e[k - 1] = v;
    }
    double s = 0.0;
    for (double v : e) s += v;
    for (double& v : e) v /= s;
    return e;
}

void taylor_green_slot(int n, int c, double* out) {
    const double pi = 4.0 * std::atan(1.0);
    const double dx = 2.0 * pi / (double)n;
    for (int k = 0; k < n; ++k) {
        const double zz = (double)k * dx;
        for (int j = 0; j < n; ++j) {
            const double yy = (double)j * dx;
            double* row = out + ((int64_t)k * n + j) * (n + 2);
            for (int i = 0; i < n; ++i) {
                const double xx = (double)i * dx;
                if (c == 0) row[i] = std::sin(xx) * std::cos(yy) * std::cos(zz);
                else if (c == 1) row[i] = -std::cos(xx) * std::sin(yy) * std::cos(zz);
                else row[i] = 0.0;
            }
            row[n] = row[n + 1] = 0.0;
        }
    }
}

std::vector<double> target_spectrum(int kmax, int isp_type, double peak, double ir_exp) {
    std::vector<double> e(kmax);
    const double pi = 4.0 * std::atan(1.0);
    for (int k = 1; k <= kmax; ++k) {
        const double wmag = (double)k;
        const double ratio = wmag / peak;
        double v;
        if (isp_type == 0) v = std::pow(wmag, -5.0 / 3.0);
        else if (isp_type == 1) v = std::pow(ratio, 3) / peak * std::exp(-3.0 * ratio);
        else if (isp_type == 2) v = std::pow(wmag, ir_exp) * std::exp(-0.5 * ir_exp * ratio * ratio);
        else if (isp_type == 3) {
            const double fac = 2.0 * pi * ratio;
            v = std::pow(fac, 4) / std::pow(1.0 + fac * fac, 3);
        } else {
            throw std::runtime_error("wrong initial spectrum type");
        }
        e[k - 1] = v;
    }
    double s = 0.0;
    for (double v : e) s += v;
    for (double& v : e) v /= s;
    return e;
}

}  // namespace hit3d
