// Register/LDS FFT engine for CDNA4 (64-wide wavefronts).
//
// A length-N complex DFT is computed cooperatively by TPF = N/E threads.
// Thread t owns E values in "stride order": slot s holds element t + s*TPF.
// The transform is a radix-r Stockham sequence (r = min(E, N/p) at span p),
// whose butterflies read a thread's own slots; outputs of every stage but
// the last go through LDS and are read back in stride order.  The output of
// the last stage is again in stride order, in the same slots, so a caller
// loads and stores global memory with consecutive lanes on consecutive
// columns (coalesced) and never needs a separate reordering pass.
//
// DIR = -1 is the forward transform (exp(-2 pi i jk/N), FFTW_FORWARD),
// DIR = +1 the unnormalised inverse.  Twiddles come from a table
// tw[j] = exp(-2 pi i j / N), j < N, built on the host in long double.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace hit3d {

// number of radix-E stages before span Pspan (Pspan a power of E)
__host__ __device__ constexpr int stage_index(int Pspan, int E) {
    int s = 0;
    while (Pspan > 1) {
        Pspan /= E;
        ++s;
    }
    return s;
}

template <typename T>
struct Consts;
template <>
struct Consts<double> {
    static constexpr double rsqrt2 = 0.70710678118654752440084436210484903928;
};
template <>
struct Consts<float> {
    static constexpr float rsqrt2 = 0.70710678118654752440084436210484903928f;
};

// multiply by -i (DIR=-1) or +i (DIR=+1)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot90(cpx<T> a) {
    if constexpr (DIR < 0) return mul_mi(a);
    else return mul_i(a);
}

// multiply by exp(DIR * i pi/4)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot45(cpx<T> a) {
    const T h = Consts<T>::rsqrt2;
    if constexpr (DIR < 0) return {(a.x + a.y) * h, (a.y - a.x) * h};
    else return {(a.x - a.y) * h, (a.y + a.x) * h};
}

// multiply by exp(DIR * i 3pi/4)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot135(cpx<T> a) {
    const T h = Consts<T>::rsqrt2;
    if constexpr (DIR < 0) return {(a.y - a.x) * h, -(a.x + a.y) * h};
    else return {-(a.x + a.y) * h, (a.x - a.y) * h};
}

// In-register DFT of R values v[o + m*S], m = 0..R-1 (natural order in/out).
template <int R, int DIR, typename T, int S>
struct DFT;

template <int DIR, typename T, int S>
struct DFT<1, DIR, T, S> {
    __device__ static __forceinline__ void run(cpx<T>*) {}
};

template <int DIR, typename T, int S>
struct DFT<2, DIR, T, S> {
    __device__ static __forceinline__ void run(cpx<T>* v) {
        cpx<T> a = v[0], b = v[S];
        v[0] = a + b;
        v[S] = a - b;
    }
};

template <int DIR, typename T, int S>
struct DFT<4, DIR, T, S> {
    __device__ static __forceinline__ void run(cpx<T>* v) {
        cpx<T> a0 = v[0], a1 = v[S], a2 = v[2 * S], a3 = v[3 * S];
        cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot90<DIR>(a1 - a3);
        v[0] = t0 + t2;
        v[2 * S] = t0 - t2;
        v[S] = t1 + t3;
        v[3 * S] = t1 - t3;
    }
};

template <int DIR, typename T, int S>
struct DFT<8, DIR, T, S> {
    __device__ static __forceinline__ void run(cpx<T>* v) {
        // even/odd split into two DFT4 then combine with W8^k
        cpx<T> e[4] = {v[0], v[2 * S], v[4 * S], v[6 * S]};
        cpx<T> o[4] = {v[S], v[3 * S], v[5 * S], v[7 * S]};
        DFT<4, DIR, T, 1>::run(e);
        DFT<4, DIR, T, 1>::run(o);
        cpx<T> w1 = rot45<DIR>(o[1]);
        cpx<T> w2 = rot90<DIR>(o[2]);
        cpx<T> w3 = rot135<DIR>(o[3]);
        v[0] = e[0] + o[0];
        v[4 * S] = e[0] - o[0];
        v[S] = e[1] + w1;
        v[5 * S] = e[1] - w1;
        v[2 * S] = e[2] + w2;
        v[6 * S] = e[2] - w2;
        v[3 * S] = e[3] + w3;
        v[7 * S] = e[3] - w3;
    }
};

template <int DIR, typename T, int S>
struct DFT<16, DIR, T, S> {
    __device__ static __forceinline__ void run(cpx<T>* v) {
        // 16 = 4 x 4: column DFT4s, twiddle W16^(a*b), row DFT4s
        cpx<T> a[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] = v[m * S];
        // DFT4 over index m = a + 4b along b for each a (stride 4)
#pragma unroll
        for (int i = 0; i < 4; ++i) DFT<4, DIR, T, 4>::run(a + i);
        // twiddles exp(DIR 2 pi i a*k1/16), a = 0..3, k1 = 0..3
        constexpr T c1 = (T)0.92387953251128675612818318939678828682;
        constexpr T s1 = (T)0.38268343236508977172845998403039886676;
        const T sg = (T)DIR;
        // k1 = 1: a=1 -> pi/8, a=2 -> pi/4, a=3 -> 3pi/8
        a[1 + 4] = cmul(a[1 + 4], mk(c1, sg * s1));
        a[2 + 4] = rot45<DIR>(a[2 + 4]);
        a[3 + 4] = cmul(a[3 + 4], mk(s1, sg * c1));
        // k1 = 2: a=1 -> pi/4, a=2 -> pi/2, a=3 -> 3pi/4
        a[1 + 8] = rot45<DIR>(a[1 + 8]);
        a[2 + 8] = rot90<DIR>(a[2 + 8]);
        a[3 + 8] = rot135<DIR>(a[3 + 8]);
        // k1 = 3: a=1 -> 3pi/8, a=2 -> 3pi/4, a=3 -> 9pi/8
        a[1 + 12] = cmul(a[1 + 12], mk(s1, sg * c1));
        a[2 + 12] = rot135<DIR>(a[2 + 12]);
        a[3 + 12] = cmul(a[3 + 12], mk(-c1, -sg * s1));
        // DFT4 along a for each k1 (contiguous groups of 4)
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) DFT<4, DIR, T, 1>::run(a + 4 * k1);
        // output index k = k1 + 4*k2 sits at a[4*k1 + k2]
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[(k1 + 4 * k2) * S] = a[4 * k1 + k2];
    }
};

template <typename T>
__device__ __forceinline__ cpx<T> ldtw(const cpx<T>* tw, int j) {
    return tw[j];
}

// LDS entry map of the exchanges.  PADSH > 0: one spare entry every
// 2^PADSH entries; PADSH = 0: none; PADSH = kSwizzle: entry i goes to
// i ^ ((i >> 3) & 15), a permutation inside each aligned group of 16 (the
// identity below 8).  A thread's radix-8 outputs are 8 consecutive entries,
// so in a transform whose threads are consecutive lanes (stride 1 between
// entries) the exchange writes land 8 entries apart: 128 bytes in FP64, the
// width of the 32 banks, so unpadded every lane of a group hit the same
// banks.  The swizzle spreads those writes and keeps the unit-stride reads
// conflict-free for 8- and 16-byte entries.  The row transforms lost their
// write conflicts (4.5-way unpadded; one pad per 16 or 32 entries still left
// 1.5- to 2-way, measured as 35 % of LDS cycles in the x pass).
inline constexpr int kSwizzle = -1;
template <int PADSH>
__device__ __forceinline__ int padi(int i) {
    if constexpr (PADSH == kSwizzle) return i ^ ((i >> 3) & 15);
    else if constexpr (PADSH > 0) return i + (i >> PADSH);
    else return i;
}
template <int PADSH>
__host__ __device__ constexpr int padded_size(int n) {
    return PADSH > 0 ? n + (n >> PADSH) + 1 : n;
}
// PADSH that gives one pad per 256 bytes of entries
template <typename T>
struct PadFor {
    static constexpr int value = sizeof(cpx<T>) == 16 ? 4 : 5;
};

// Synchronisation between LDS exchange phases: the whole block, or only the
// wavefront when every transform lives inside one wavefront (LDS operations
// of one wavefront complete in issue order; the fences and the wave barrier
// keep the compiler from moving accesses across the phase boundary).
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// One Stockham stage at span Pspan for NF independent transforms that share
// the thread mapping (and so the twiddles), then recursion to the next.
// v[f][s]: transform f, slot s.  lds: exchange buffer; element j of
// transform f sits at lds[f*FS + padi(j*LS)].  t: thread index in the transform.
// TWREC: the R-1 twiddles of a radix-8/4/2 butterfly come from one table load
// w and its powers w^2..w^7 by complex products (a few ulp of extra rounding,
// seven loads fewer); used where twiddle loads dominate the wait time.
// TW: precision of the twiddle table and of that recurrence (FP64 for FP32
// data keeps the twiddles within an ulp of FP32).
// TC: a TwCache of this thread's twiddles (see below), used instead of
// computing them when transforms with the same thread mapping repeat.
template <int N, int E, int DIR, typename T, int LS, int NF, int FS, int PADSH, bool WAVE,
          int Pspan, bool TWREC = false, typename TW = T, typename TC = void>
struct Stages {
    static constexpr int R = (E < N / Pspan) ? E : N / Pspan;
    static constexpr int TPF = N / E;
    static constexpr int NB = E / R;  // butterflies per thread
    __device__ static __forceinline__ void run(cpx<T> (&v)[NF][E], int t, cpx<T>* lds,
                                               const cpx<TW>* __restrict__ tw,
                                               const TC* tc = nullptr) {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int b = t + q * TPF;
            const int k = b & (Pspan - 1);
            if constexpr (Pspan > 1 && !std::is_void_v<TC>) {
                static_assert(NB == 1, "cached twiddles need one butterfly per thread per stage");
                constexpr int si = stage_index(Pspan, E) - 1;
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    // the cache holds the forward twiddles; the inverse uses their
                    // conjugates (exactly the TWREC recurrence of the conjugate)
                    const cpx<T> wc = tc->get(si, m);
                    const cpx<T> wm = DIR > 0 ? conj(wc) : wc;
#pragma unroll
                    for (int f = 0; f < NF; ++f) v[f][q + m * NB] = cmul(v[f][q + m * NB], wm);
                }
            } else if constexpr (Pspan > 1 && TWREC && (R == 8 || R == 4 || R == 2)) {
                cpx<TW> w[R];
                w[1] = ldtw(tw, k * (N / (Pspan * R)));
                if constexpr (DIR > 0) w[1].y = -w[1].y;
                if constexpr (R >= 4) {
                    w[2] = cmul(w[1], w[1]);
                    w[3] = cmul(w[2], w[1]);
                }
                if constexpr (R == 8) {
                    w[4] = cmul(w[2], w[2]);
                    w[5] = cmul(w[4], w[1]);
                    w[6] = cmul(w[3], w[3]);
                    w[7] = cmul(w[6], w[1]);
                }
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    const cpx<T> wm{(T)w[m].x, (T)w[m].y};
#pragma unroll
                    for (int f = 0; f < NF; ++f) v[f][q + m * NB] = cmul(v[f][q + m * NB], wm);
                }
            } else if constexpr (Pspan > 1) {
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    const cpx<TW> wt = ldtw(tw, m * k * (N / (Pspan * R)));
                    cpx<T> w{(T)wt.x, (T)wt.y};
                    if constexpr (DIR > 0) w.y = -w.y;
#pragma unroll
                    for (int f = 0; f < NF; ++f) v[f][q + m * NB] = cmul(v[f][q + m * NB], w);
                }
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) DFT<R, DIR, T, NB>::run(v[f] + q);
        }
        if constexpr (Pspan * R < N) {
            lds_sync<WAVE>();  // previous readers of lds are done
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int b = t + q * TPF;
                const int k = b & (Pspan - 1);
                const int base = (b - k) * R + k;
#pragma unroll
                for (int m = 0; m < R; ++m)
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        lds[f * FS + padi<PADSH>((base + m * Pspan) * LS)] = v[f][q + m * NB];
            }
            lds_sync<WAVE>();
#pragma unroll
            for (int s = 0; s < E; ++s)
#pragma unroll
                for (int f = 0; f < NF; ++f) v[f][s] = lds[f * FS + padi<PADSH>((t + s * TPF) * LS)];
            Stages<N, E, DIR, T, LS, NF, FS, PADSH, WAVE, Pspan * R, TWREC, TW, TC>::run(v, t, lds, tw,
                                                                                       tc);
        }
    }
};

// NF transforms at once.  Every thread taking part in the exchange must call
// this (it synchronises), also threads whose data is masked out.
template <int N, int E, int DIR, typename T, int LS, int NF, int FS, int PADSH, bool WAVE,
          bool TWREC = false, typename TW = T, typename TC = void>
__device__ __forceinline__ void multi_fft(cpx<T> (&v)[NF][E], int t, cpx<T>* lds,
                                          const cpx<TW>* __restrict__ tw,
                                          const TC* tc = nullptr) {
    Stages<N, E, DIR, T, LS, NF, FS, PADSH, WAVE, 1, TWREC, TW, TC>::run(v, t, lds, tw, tc);
}

// Forward twiddles of one thread for every twiddled stage of a length-N
// transform with E values per thread and radix E throughout (N a power of
// E): computed once, by the same recurrence as TWREC in TW and rounded to T,
// then shared by every transform with that thread mapping, either direction
// (the same values as TWREC up to the compiler's choice of fused
// multiply-adds).
template <typename T, int N, int E>
struct TwCache {
    static constexpr int NS = stage_index(N, E) - 1;  // twiddled stages
    cpx<T> w[NS > 0 ? NS : 1][E];
    __device__ __forceinline__ cpx<T> get(int si, int m) const { return w[si][m]; }
};

// The same twiddles for all N/E thread positions of a block, in LDS: entry
// (si, m) of position t at base[(si * E + m) * (N / E) + t].
template <typename T, int N, int E>
struct TwLds {
    static constexpr int NS = TwCache<T, N, E>::NS;
    static constexpr int SIZE = (NS > 0 ? NS : 1) * E * (N / E);
    const cpx<T>* base;
    int t;
    __device__ __forceinline__ cpx<T> get(int si, int m) const {
        return base[(si * E + m) * (N / E) + t];
    }
};

template <int N, int E, typename T, typename TW>
__device__ __forceinline__ void fill_twiddles(TwCache<T, N, E>& c, int t,
                                              const cpx<TW>* __restrict__ tw) {
    static_assert(E == 8, "the cache follows the radix-8 recurrence");
#pragma unroll
    for (int si = 0; si < TwCache<T, N, E>::NS; ++si) {
        int Pspan = E;
        for (int i = 0; i < si; ++i) Pspan *= E;
        const int k = t & (Pspan - 1);
        cpx<TW> w[8];
        w[1] = ldtw(tw, k * (N / (Pspan * E)));
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[2], w[1]);
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[3], w[3]);
        w[7] = cmul(w[6], w[1]);
#pragma unroll
        for (int m = 1; m < 8; ++m) c.w[si][m] = cpx<T>{(T)w[m].x, (T)w[m].y};
    }
}

// One transform with block-wide synchronisation; element j at lds[padi(j*LS)].
template <int N, int E, int DIR, typename T, int LS, int PADSH = 0, bool TWREC = false,
          typename TW = T, typename TC = void>
__device__ __forceinline__ void block_fft(cpx<T> (&v)[E], int t, cpx<T>* lds,
                                          const cpx<TW>* __restrict__ tw, const TC* tc = nullptr) {
    cpx<T>(&vv)[1][E] = *reinterpret_cast<cpx<T>(*)[1][E]>(&v);
    Stages<N, E, DIR, T, LS, 1, 0, PADSH, false, 1, TWREC, TW, TC>::run(vv, t, lds, tw, tc);
}

}  // namespace hit3d
