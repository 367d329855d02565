// Shared device/host definitions for the HIT3DP pseudospectral step on MI355X.
//
// Device layout of one slot (all slots of a grid share it):
//   row pitch P complex entries, P = N/2+1 rounded up to a multiple of 16, so
//   every row starts on a 256-byte boundary in FP64 (128 bytes in FP32);
//   rows are ordered like the reference's Fortran arrays (x_fftw.f90:321-327):
//     physical  real (x, y, z)       at  (z*N + y)*2P + x          x < N
//     spectral  complex (kx, kz, ky) at  (ky*N + kz)*P + kx        kx <= N/2
//   The host-side (Fortran) arrays use pitch N+2 reals; the C-ABI converts
//   with 2-D copies.  Entries kx in (N/2, P) are padding and are never read
//   by a kernel as data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hit3d {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by +i and -i
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return {-a.y, a.x}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) { return {a.y, -a.x}; }

// Streaming (nontemporal) global load and store of one complex entry, for
// fields read or written once per pass: a 512^3 slot is 1 GB, four times the
// last-level cache, so nothing a pass streams is found again by a later one.
// Marking those accesses streaming took the 512^3 FP64 step from 6.44 to
// 5.7 ms.  From 64^3 up every grid streams (the 64^3 LES step ran 0.271 ->
// 0.26 ms with it); 32^3 keeps cached accesses (streaming cost its step 4 %:
// 66 -> 69 us).
template <typename T>
__device__ __forceinline__ cpx<T> ld_nt(const cpx<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cpx<T>{x.x, x.y};
}
template <typename T>
__device__ __forceinline__ void st_nt(cpx<T>* p, cpx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p));
}

template <int N>
inline constexpr bool kStream = N >= 64;
__host__ __device__ inline bool streams(int n) { return n >= 64; }
template <bool NT, typename T>
__device__ __forceinline__ cpx<T> ld_g(const cpx<T>* p) {
    if constexpr (NT) return ld_nt(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_g(cpx<T>* p, cpx<T> v) {
    if constexpr (NT) st_nt(p, v);
    else *p = v;
}

// Pitch (complex entries per row) for grid size n.
__host__ __device__ inline int64_t row_pitch(int64_t n) { return ((n / 2 + 1 + 15) / 16) * 16; }

// Signed wavenumber of index j along y or z: x_fftw.f90:280-296
// (j > n/2 wraps to j-n; j == n/2 stays +n/2).
__host__ __device__ __forceinline__ int wavenum(int j, int n) { return (2 * j > n) ? j - n : j; }

// ialias(i,j,k), x_fftw.f90:298-310: how many of |kx|,|ky|,|kz| exceed n/3
// (integer division, as real(nx/3,8)).
__host__ __device__ __forceinline__ int alias_count(int kx, int ky, int kz, int n) {
    int t = n / 3;
    return (kx > t) + (abs(ky) > t) + (abs(kz) > t);
}

// nint(sqrt(real(k2,4))) for integer k2 >= 0, computed exactly in integers:
// nint(sqrt(k2)) = s + 1 when k2 > s*s + s (s = floor(sqrt(k2))), else s.
// For k2 < 2^24 this equals the float32 expression used by m_stats.f90:218,
// m_force.f90:139 and init_velocity.f90:228 (checked exhaustively in tests).
__host__ __device__ __forceinline__ int shell_of(int k2) {
    int s = (int)sqrt((double)k2);
    while ((int64_t)s * s > k2) --s;
    while ((int64_t)(s + 1) * (s + 1) <= k2) ++s;
    return (k2 > s * s + s) ? s + 1 : s;
}

}  // namespace hit3d
