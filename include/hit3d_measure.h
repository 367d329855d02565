/*
 * hit3d_measure.h -- measurement entry points of libhit3d_hip.so.  None has a
 * counterpart in the reference; they give the yardsticks the time step's
 * kernels are measured against.  Conventions as in hit3d.h (int64_t and
 * double by reference, 0 on success, hit3d_last_error for the message).
 */
#ifndef HIT3D_MEASURE_H
#define HIT3D_MEASURE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A plain streaming copy of one slot, wrk(2) = wrk(1) (every byte read and
 * written once, nontemporal 16-byte loads and stores, *blocks workgroups of
 * 256 threads), run *reps times after one untimed launch and timed with HIP
 * events on the library's stream: *ms = mean milliseconds per copy, *bytes =
 * bytes read + written per copy.  Overwrites wrk(2).  Needs hit3d_init. */
int hit3d_stream_copy(const int64_t *reps, const int64_t *blocks, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
