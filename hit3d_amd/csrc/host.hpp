// Host-side operations shared by the C ABI files: slot bookkeeping, the
// transform sequences, the reference subroutines as launch sequences, the
// statistics' host sides and the step orchestration.  All act on lib().
#pragma once
#include <cstdint>
#include <vector>

#include "state.hpp"

namespace hit3d {

// timers, state, slots (host_state.hip)
extern const char* const kTimerNames;  // comma-separated TimerClass names
void collect_timers();
void require_ready();
size_t slot_bytes();
double*& slot_ref(int64_t array_id, int64_t s);
void copy_slot(double* dst, const double* src);
void zero_slot(double* dst);
// rhs_old(1:3) to FP32 (the mixed step's format) or back to FP64; no-op when
// already so.  Every path that reads rhs_old(1:3), other than the mixed
// step, asks for FP64 first.
void rhs_old_format(bool fp32);
void pmul(const double* a, const double* b, double* out);
void clear_graphs();
void free_all();
double* dalloc(size_t bytes);

// A temporary device buffer of one call, freed when the call's scope ends,
// also when a later check throws.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) : p(dalloc(bytes)) {}
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// the device address of the allocation-free entry point's scratch array on
// the current device (capi_closure.hip)
float* static_scratch_setup();

// transforms (host_transforms.hip); fft_forward / fft_inverse are in state.hpp
void pass_y(double* slot, int dir);
void fft_forward_yz(double*& slot, int ph);
void fft_forward_product(const double* a, const double* b, double*& dst, int ph);
void fft_forward_product_acc(const double* a, const double* b, double* work, double* r0, int ax0,
                             bool mask0, double* r1 = nullptr, int ax1 = 0, bool mask1 = false);
void fft_inverse_from(const double* src, double*& dst, int ph);
// up to three slots per launch (slots[i] points at the slot pointer, which the
// out-of-place pass swaps like the single-slot forms do)
bool batch_spares();
void fft_forward_batch(double** const* slots, int nb);
void fft_inverse_batch(double** const* slots, int nb);
void fft_inverse_from_batch(const double* const* src, double** const* dst, int nb);

// reference subroutines (host_rhs.hip)
void les_turb_visc();
void les_velocity();
void les_scalar(int n);
void mlp_alloc_scratch();
void mlp_gradients();
void mlp_velocity();
// keep_shifted (phase shift with scalars): leave the shifted physical
// velocity in lib().keep for do_rhs_velocity(true) instead of discarding it
void do_rhs_scalars(bool int_scalars, bool keep_shifted = false);
double read_courant_max();
void do_rhs_velocity(bool use_kept = false);
void keep_slots();  // allocates lib().keep
void force_enqueue();
void do_force(double* energy);
void do_advance(int which, bool first, double dt, const double* dtp = nullptr);
void do_pressure();

// statistics (host_stats.hip)
double key_to_double(uint64_t k);
void minmax_to_host(const double* a, const double* b, const double* c, double* sum, double* mn,
                    double* mx);
void sums_to_host(const double* a, const double* b, const double* c, double out[2]);
void do_stat_scalars(double* diss, double* mean, double* var, double* mn, double* mx, double* es);
void do_divergence(double* dmin, double* dmax);

// fused velocity RHS + update of the phase-shift step (step_ps.hip)
bool fused_ps_supported();
void fused_ps_velocity(double* const shifted[3], bool first, const double* dtp);
bool fused_ps_scalars_supported();
void fused_ps_scalars(bool first, const double* dtp);

// one step (host_step.hip)
void step_body(bool fused, bool mixed, bool scal, bool first_v, bool first_s);
std::vector<void*> slot_state();
void set_slot_state(const std::vector<void*>& s);
void restore_work_slots(const std::vector<void*>& s);
void step_graph(bool fused, bool mixed, bool scal, bool first_v, bool first_s);

}  // namespace hit3d
