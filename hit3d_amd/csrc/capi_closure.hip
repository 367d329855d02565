// C ABI of the learned subgrid closure (les_model 11): parameters, scaling,
// training data, training and inference, and the filtering that makes its
// training set.  No reference counterpart.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "guard.hpp"
#include "hit3d.h"
#include "host.hpp"

using namespace hit3d;

static void mlp_alloc_params() {
    auto& M = lib().mlp;
    if (M.p) return;
    const size_t b = sizeof(float) * mlp::NPARAM;
    M.p = reinterpret_cast<float*>(dalloc(b));
    M.m = reinterpret_cast<float*>(dalloc(b));
    M.v = reinterpret_cast<float*>(dalloc(b));
    M.g = reinterpret_cast<float*>(dalloc(b));
    M.wb = reinterpret_cast<uint16_t*>(dalloc(sizeof(uint16_t) * mlp::NPARAM));
    M.bad = reinterpret_cast<unsigned int*>(dalloc(sizeof(unsigned int) * 4));
}

static void mlp_reserve(int64_t tiles, int64_t steps) {
    auto& M = lib().mlp;
    if (tiles > M.cap_tiles) {
        if (M.partial) check(hipFree(M.partial), "hipFree");
        if (M.lossp) check(hipFree(M.lossp), "hipFree");
        M.partial = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NPARAM * tiles));
        M.lossp = reinterpret_cast<float*>(dalloc(sizeof(float) * tiles));
        M.cap_tiles = tiles;
    }
    if (!M.red)
        M.red = reinterpret_cast<float*>(
            dalloc(sizeof(float) * kReduceGroups * (mlp::NPARAM + 1)));
    if (steps > M.cap_steps) {
        if (M.loss) check(hipFree(M.loss), "hipFree");
        M.loss = reinterpret_cast<float*>(dalloc(sizeof(float) * steps));
        M.cap_steps = steps;
    }
}

// Scratch of hit3d_mlp_grads_static: per-call staging of the samples, the
// per-tile gradient partials, the two-level reduction and the results, in
// one file-scope device array (no allocation per call).  Offsets in floats,
// each segment 256-byte aligned.
namespace {
constexpr int64_t kStaticCount = 8192;  // samples per call at most
constexpr int64_t kStaticTiles = kStaticCount / 32;
constexpr int64_t seg(int64_t floats) { return (floats + 63) / 64 * 64; }
constexpr int64_t kOffX = 0;
constexpr int64_t kOffT = kOffX + seg(kStaticCount * mlp::NIN);
constexpr int64_t kOffPartial = kOffT + seg(kStaticCount * mlp::NOUT);
constexpr int64_t kOffLossp = kOffPartial + seg(kStaticTiles * mlp::NPARAM);
constexpr int64_t kOffRed = kOffLossp + seg(kStaticTiles);
constexpr int64_t kOffGrad = kOffRed + seg((int64_t)kReduceGroups * (mlp::NPARAM + 1));
constexpr int64_t kOffLoss = kOffGrad + seg(mlp::NPARAM);
constexpr int64_t kScratchFloats = kOffLoss + seg(1);  // 2,153,984 floats = 8.2 MiB
constexpr int kMaxDevices = 64;
// device address of `scratch` on each device (hipGetSymbolAddress); the
// array's name in host code is only a stand-in for the symbol
float* g_scratch_addr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace hit3d {
// Resolves this device's copy of `scratch` once per device, with the device
// current; hit3d_init calls it, and the static entry point on its first call
// on a device that was not initialised through hit3d_init.
float* static_scratch_setup() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("static scratch: device index");
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("static scratch: no device address");
        g_scratch_addr[dev] = static_cast<float*>(p);
    }
    return g_scratch_addr[dev];
}
}  // namespace hit3d

extern "C" {

int hit3d_mlp_set_params(const float* params) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        mlp_alloc_params();
        auto& M = L.mlp;
        const size_t b = sizeof(float) * mlp::NPARAM;
        check(hipMemcpyAsync(M.p, params, b, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        check(hipMemsetAsync(M.m, 0, b, L.st), "hipMemsetAsync");
        check(hipMemsetAsync(M.v, 0, b, L.st), "hipMemsetAsync");
        launch_mlp_to_bf16(M.p, M.wb, L.st);
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        M.step = 0;
        M.have_params = true;
    });
}

int hit3d_mlp_get_params(float* params) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (!L.mlp.have_params) throw std::runtime_error("mlp: no parameters set");
        check(hipMemcpyAsync(params, L.mlp.p, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost,
                             L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_mlp_set_norm(const double* xmean, const double* xstd, const double* ymean,
                       const double* ystd) {
    return guarded([&] {
        auto& N = lib().mlp.norm;
        for (int i = 0; i < mlp::NIN; ++i) {
            if (!(xstd[i] > 0.0)) throw std::runtime_error("mlp: input scales must be positive");
            N.xmean[i] = (float)xmean[i];
            N.xstd[i] = (float)xstd[i];
        }
        for (int i = 0; i < mlp::NOUT; ++i) {
            if (!(ystd[i] > 0.0)) throw std::runtime_error("mlp: output scales must be positive");
            N.ymean[i] = (float)ymean[i];
            N.ystd[i] = (float)ystd[i];
        }
        lib().mlp.have_norm = true;
    }, false);
}

int hit3d_mlp_set_data(const int64_t* count, const float* X, const float* T) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        auto& M = L.mlp;
        if (*count < 1) throw std::runtime_error("mlp: empty training set");
        if (M.X) check(hipFree(M.X), "hipFree");
        if (M.T) check(hipFree(M.T), "hipFree");
        M.X = M.T = nullptr;
        M.X = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NIN * *count));
        M.T = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NOUT * *count));
        check(hipMemcpyAsync(M.X, X, sizeof(float) * mlp::NIN * *count, hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(M.T, T, sizeof(float) * mlp::NOUT * *count, hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        M.count = *count;
    });
}

int hit3d_mlp_train(const int64_t* steps, const int64_t* batch, const double* lr, double* losses) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        auto& M = L.mlp;
        if (!M.have_params || !M.X) throw std::runtime_error("mlp: parameters and data first");
        if (*batch < 32 || *batch % 32 != 0) throw std::runtime_error("mlp: batch must be a multiple of 32");
        if (*steps < 1) return;
        const int64_t tiles = *batch / 32;
        mlp_reserve(tiles, *steps);
        for (int64_t i = 0; i < *steps; ++i) {
            const int64_t start = ((int64_t)M.step * *batch) % M.count;
            launch_mlp_train_tiles(M.X, M.T, M.count, start, (int)tiles, *batch, M.wb, M.p,
                                   M.partial, M.lossp, L.st);
            launch_mlp_reduce(M.partial, M.lossp, (int)tiles, M.red, M.g, M.loss + i, L.st);
            ++M.step;
            launch_mlp_adam(M.p, M.wb, M.g, M.m, M.v, M.step, (float)*lr, L.st);
        }
        std::vector<float> ls(*steps);
        check(hipMemcpyAsync(ls.data(), M.loss, sizeof(float) * *steps, hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        if (losses)
            for (int64_t i = 0; i < *steps; ++i) losses[i] = (double)ls[i] / (double)(*batch * mlp::NOUT);
    });
}

int hit3d_mlp_grads(const int64_t* count, const float* X, const float* T, float* grads,
                    double* loss) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        auto& M = L.mlp;
        if (!M.have_params) throw std::runtime_error("mlp: parameters first");
        const int64_t n = *count;
        if (n < 32 || n % 32 != 0) throw std::runtime_error("mlp: count must be a multiple of 32");
        DevBuf bX(sizeof(float) * mlp::NIN * n), bT(sizeof(float) * mlp::NOUT * n);
        float* dX = bX.as<float>();
        float* dT = bT.as<float>();
        check(hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        check(hipMemcpyAsync(dT, T, sizeof(float) * mlp::NOUT * n, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        const int64_t tiles = n / 32;
        mlp_reserve(tiles, 1);
        launch_mlp_train_tiles(dX, dT, n, 0, (int)tiles, n, M.wb, M.p, M.partial, M.lossp, L.st);
        launch_mlp_reduce(M.partial, M.lossp, (int)tiles, M.red, M.g, M.loss, L.st);
        float ls = 0.f;
        check(hipMemcpyAsync(grads, M.g, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipMemcpyAsync(&ls, M.loss, sizeof(float), hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        if (loss) *loss = (double)ls / (double)(n * mlp::NOUT);
    });
}

int hit3d_mlp_grads_static(const int64_t* count, const float* X, const float* T, float* grads,
                           double* loss) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        auto& M = L.mlp;
        if (!M.have_params) throw std::runtime_error("mlp: parameters first");
        const int64_t n = *count;
        if (n < 32 || n % 32 != 0) throw std::runtime_error("mlp: count must be a multiple of 32");
        if (n > kStaticCount)
            throw std::runtime_error("mlp_grads_static: count " + std::to_string(n) +
                                     " exceeds the static scratch limit of " +
                                     std::to_string(kStaticCount) + " samples");
        float* base = static_scratch_setup();  // this device's address
        float* dX = base + kOffX;
        float* dT = base + kOffT;
        float* partial = base + kOffPartial;
        float* lossp = base + kOffLossp;
        float* red = base + kOffRed;
        float* g = base + kOffGrad;
        float* dloss = base + kOffLoss;
        // every scratch value read below is first written by this call (the
        // samples, one full partial row and loss per tile, every group sum,
        // the gradients and the loss), so nothing needs clearing
        check(hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(dT, T, sizeof(float) * mlp::NOUT * n, hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        const int64_t tiles = n / 32;
        launch_mlp_train_tiles(dX, dT, n, 0, (int)tiles, n, M.wb, M.p, partial, lossp, L.st);
        launch_mlp_reduce(partial, lossp, (int)tiles, red, g, dloss, L.st);
        float ls = 0.f;
        check(hipMemcpyAsync(grads, g, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(&ls, dloss, sizeof(float), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        if (loss) *loss = (double)ls / (double)(n * mlp::NOUT);
    });
}

int hit3d_mlp_forward(const int64_t* count, const float* X, float* Y) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        auto& M = L.mlp;
        if (!M.have_params) throw std::runtime_error("mlp: parameters first");
        const int64_t n = *count;
        if (n < 1) return;
        DevBuf bX(sizeof(float) * mlp::NIN * n), bY(sizeof(float) * mlp::NOUT * n);
        float* dX = bX.as<float>();
        float* dY = bY.as<float>();
        check(hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        launch_mlp_forward(dX, n, M.wb, M.p, dY, L.st);
        check(hipMemcpyAsync(Y, dY, sizeof(float) * mlp::NOUT * n, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_box_filter(const int64_t* ratio, double* ubar, double* tau) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int r = (int)*ratio;
        if (r < 1 || L.n % r != 0) throw std::runtime_error("box filter: ratio must divide n");
        if (9 * L.n * (int64_t)sizeof(double) > 65536)
            throw std::runtime_error("box filter: grids up to 512 points per direction");
        const int m = L.n / r;
        const size_t m3 = (size_t)m * m * m;
        DevBuf buf(sizeof(double) * 9 * m3);
        double* d = buf.as<double>();
        {
            Timed t(TC_POINTWISE);
            launch_box_filter(L.g, L.W[1], L.W[2], L.W[3], r, d, d + 3 * m3, L.st);
        }
        check(hipMemcpyAsync(ubar, d, sizeof(double) * 3 * m3, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipMemcpyAsync(tau, d + 3 * m3, sizeof(double) * 6 * m3, hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_sgs_features(const double* delta, float* X) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        mlp_gradients();
        const double* gr[9];
        for (int i = 0; i < 9; ++i) gr[i] = L.mlp.buf[i];
        const size_t n3 = (size_t)L.n * L.n * L.n;
        DevBuf buf(sizeof(float) * mlp::NIN * n3);
        float* d = buf.as<float>();
        {
            Timed t(TC_POINTWISE);
            launch_sgs_features(L.g, gr, *delta, d, L.st);
        }
        check(hipMemcpyAsync(X, d, sizeof(float) * mlp::NIN * n3, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_mlp_fallbacks(int64_t* count) {
    return guarded([&] { *count = lib().mlp.fallbacks; }, false);
}

int hit3d_mlp_set_mix(const double* weight) {
    return guarded([&] {
        if (!(*weight >= 0.0 && *weight <= 1.0))
            throw std::runtime_error("mlp: the Smagorinsky weight must lie in [0, 1]");
        clear_graphs();
        lib().mlp.mix = *weight;
    }, false);
}

}  // extern "C"
