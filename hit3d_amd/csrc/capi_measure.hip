// Measurement entry points (include/hit3d_measure.h): no reference
// counterpart.
#include <stdexcept>

#include "guard.hpp"
#include "hit3d_measure.h"
#include "host.hpp"

using namespace hit3d;

extern "C" {

int hit3d_stream_copy(const int64_t* reps, const int64_t* blocks, double* ms, double* bytes) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*reps < 1) throw std::runtime_error("stream_copy: reps must be >= 1");
        if (*blocks < 1 || *blocks > (1 << 20))
            throw std::runtime_error("stream_copy: blocks must be in [1, 2^20]");
        const int64_t b = (int64_t)slot_bytes();  // a multiple of 256
        launch_stream_copy(L.W[1], L.W[2], b, (int)*blocks, L.st);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        float t = 0.f;
        hipError_t err = hipEventRecord(e0, L.st);
        for (int64_t r = 0; r < *reps && err == hipSuccess; ++r) {
            launch_stream_copy(L.W[1], L.W[2], b, (int)*blocks, L.st);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, L.st);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        check(err, "stream_copy timing");
        *ms = (double)t / (double)*reps;
        *bytes = 2.0 * (double)b;
    });
}

}  // extern "C"
