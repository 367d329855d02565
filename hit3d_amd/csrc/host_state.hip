// Library state: the singleton, error checks, timers, slot bookkeeping and
// allocation.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace hit3d {

static Lib g_lib;
Lib& lib() { return g_lib; }

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

const char* const kTimerNames =
    "fft_x,fft_y,fft_z,pointwise,reduce,step_inv_z,step_inv_y,step_x,step_fwd_y,step_fwd_z,"
    "step_update,particles";

Timed::Timed(int c) : cls(c) {
    Lib& L = lib();
    if (!L.timing) return;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    check(hipEventRecord(a, L.st), "hipEventRecord");
}
Timed::~Timed() {
    Lib& L = lib();
    if (!a) return;
    if (hipEventRecord(b, L.st) == hipSuccess) L.evs.push_back({cls, a, b});
}

void collect_timers() {
    Lib& L = lib();
    if (L.evs.empty()) return;
    check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    for (auto& e : L.evs) {
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, e.a, e.b), "hipEventElapsedTime");
        L.tms[e.cls] += ms;
        L.tcnt[e.cls] += 1;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    L.evs.clear();
}


// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
void require_ready() {
    if (!lib().ready) throw std::runtime_error("hit3d_init has not been called");
}

size_t slot_bytes() { return (size_t)lib().g.nc * sizeof(cpx<double>); }

double*& slot_ref(int64_t array_id, int64_t s) {
    Lib& L = lib();
    if (array_id == 1) {
        if (s < 1 || s > L.nf) throw std::runtime_error("fields slot out of range");
        return L.F[s];
    }
    if (array_id == 2) {
        if (s < 0 || s > L.nw) throw std::runtime_error("wrk slot out of range");
        return L.W[s];
    }
    if (array_id == 3) {
        if (s < 1 || s > L.nf) throw std::runtime_error("rhs_old slot out of range");
        if (s <= 3) rhs_old_format(false);
        return L.R[s];
    }
    throw std::runtime_error("unknown array id");
}

void copy_slot(double* dst, const double* src) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    check(hipMemcpyAsync(dst, src, slot_bytes(), hipMemcpyDeviceToDevice, L.st), "hipMemcpyAsync");
}

void zero_slot(double* dst) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_fill_zero(dst, (int64_t)(slot_bytes() / sizeof(double)), L.st);
}

void rhs_old_format(bool fp32) {
    Lib& L = lib();
    if (L.ro32 == fp32) return;
    Timed t(TC_POINTWISE);
    const int64_t count = (int64_t)(slot_bytes() / sizeof(double));
    // each slot goes through the spare: the two layouts overlap in place
    for (int c = 1; c <= 3; ++c) {
        if (fp32) {
            check(hipMemcpyAsync(L.spare, L.R[c], slot_bytes(), hipMemcpyDeviceToDevice, L.st),
                  "hipMemcpyAsync");
            launch_narrow(L.spare, reinterpret_cast<float*>(L.R[c]), count, L.st);
        } else {
            check(hipMemcpyAsync(L.spare, L.R[c], slot_bytes() / 2, hipMemcpyDeviceToDevice,
                                 L.st),
                  "hipMemcpyAsync");
            launch_widen(reinterpret_cast<const float*>(L.spare), L.R[c], count, L.st);
        }
    }
    L.ro32 = fp32;
}

void pmul(const double* a, const double* b, double* out) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_phys_mul(L.g, a, b, out, L.st);
}

void clear_graphs() {
    Lib& L = lib();
    for (auto& g : L.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    L.graphs.clear();
    L.graph_seen.clear();
}

void free_all() {
    Lib& L = lib();
    (void)hipDeviceSynchronize();
    clear_graphs();
    for (auto* p : L.F) if (p) (void)hipFree(p);
    for (auto* p : L.W) if (p) (void)hipFree(p);
    for (auto* p : L.R) if (p) (void)hipFree(p);
    for (auto* p : L.W32) if (p) (void)hipFree(p);
    if (L.spare) (void)hipFree(L.spare);
    for (auto* p : L.spares_b) if (p) (void)hipFree(p);
    for (auto* p : L.keep) if (p) (void)hipFree(p);
    if (L.tw64) (void)hipFree(L.tw64);
    if (L.tw32) (void)hipFree(L.tw32);
    if (L.ps_tab) (void)hipFree(L.ps_tab);
    if (L.dshell) (void)hipFree(L.dshell);
    if (L.dpart) (void)hipFree(L.dpart);
    if (L.dmax) (void)hipFree(L.dmax);
    if (L.tv) (void)hipFree(L.tv);
    for (void* q : {(void*)L.mlp.p, (void*)L.mlp.wb, (void*)L.mlp.m, (void*)L.mlp.v,
                    (void*)L.mlp.g, (void*)L.mlp.partial, (void*)L.mlp.lossp, (void*)L.mlp.loss,
                    (void*)L.mlp.red,
                    (void*)L.mlp.X, (void*)L.mlp.T, (void*)L.mlp.bad})
        if (q) (void)hipFree(q);
    for (auto* q : L.mlp.buf) if (q) (void)hipFree(q);
    if (L.xyzp) (void)hipFree(L.xyzp);
    if (L.uvwp) (void)hipFree(L.uvwp);
    if (L.dscal) (void)hipFree(L.dscal);
    if (L.hpin) (void)hipHostFree(L.hpin);
    if (L.st) (void)hipStreamDestroy(L.st);
    for (auto& e : L.evs) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    std::string keep = L.err;
    L = Lib();
    L.err = keep;
}

double* dalloc(size_t bytes) {
    void* p = nullptr;
    check(hipMalloc(&p, bytes), "hipMalloc");
    check(hipMemset(p, 0, bytes), "hipMemset");
    // hipMemset runs on the null stream, which the library's non-blocking
    // stream does not wait for: without this, a buffer allocated on first
    // use could be cleared after the kernels that write it
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return static_cast<double*>(p);
}

}  // namespace hit3d
