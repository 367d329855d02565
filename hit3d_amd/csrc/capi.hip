// C ABI of the library (include/hit3d.h): device-resident state and the
// entry points that replace the reference's x_fftw / rhs_* / pressure /
// m_force / dealias_all subroutines.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hit3d.h"
#include "init_host.hpp"
#include "passes.hpp"
#include "state.hpp"

namespace hit3d {

static Lib g_lib;
Lib& lib() { return g_lib; }

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

static const char* kTimerNames =
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

static void collect_timers() {
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
// transforms
// ---------------------------------------------------------------------------
static void pass_y(double* slot, int dir) {
    Lib& L = lib();
    auto* c = reinterpret_cast<cpx<double>*>(slot);
    const int64_t n = L.n, P = L.g.P;
    Timed t(TC_FFT_Y);
    launch_cols<double>(L.n, dir, c, c, n * P, n * P, P, P, L.tw64, L.st);
}

// y and z passes of a forward transform whose x pass is done; ph = -1
// multiplies the spectral result by exp(-i theta) (the phase shift back).
static void fft_forward_yz(double*& slot, int ph) {
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
static void fft_forward_product(const double* a, const double* b, double*& dst, int ph) {
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
static void fft_forward_product_acc(const double* a, const double* b, double* work, double* r0,
                                    int ax0, bool mask0, double* r1 = nullptr, int ax1 = 0,
                                    bool mask1 = false) {
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
static void fft_inverse_from(const double* src, double*& dst, int ph) {
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

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static void require_ready() {
    if (!lib().ready) throw std::runtime_error("hit3d_init has not been called");
}

static size_t slot_bytes() { return (size_t)lib().g.nc * sizeof(cpx<double>); }

static double*& slot_ref(int64_t array_id, int64_t s) {
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
        return L.R[s];
    }
    throw std::runtime_error("unknown array id");
}

static void copy_slot(double* dst, const double* src) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    check(hipMemcpyAsync(dst, src, slot_bytes(), hipMemcpyDeviceToDevice, L.st), "hipMemcpyAsync");
}

static void zero_slot(double* dst) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_fill_zero(dst, (int64_t)(slot_bytes() / sizeof(double)), L.st);
}

static void pmul(const double* a, const double* b, double* out) {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_phys_mul(L.g, a, b, out, L.st);
}

static void clear_graphs() {
    Lib& L = lib();
    for (auto& g : L.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    L.graphs.clear();
    L.graph_seen.clear();
}

static void free_all() {
    Lib& L = lib();
    (void)hipDeviceSynchronize();
    clear_graphs();
    for (auto* p : L.F) if (p) (void)hipFree(p);
    for (auto* p : L.W) if (p) (void)hipFree(p);
    for (auto* p : L.R) if (p) (void)hipFree(p);
    for (auto* p : L.W32) if (p) (void)hipFree(p);
    if (L.spare) (void)hipFree(L.spare);
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

static double* dalloc(size_t bytes) {
    void* p = nullptr;
    check(hipMalloc(&p, bytes), "hipMalloc");
    check(hipMemset(p, 0, bytes), "hipMemset");
    return static_cast<double*>(p);
}

// ---------------------------------------------------------------------------
// reference subroutines
// ---------------------------------------------------------------------------
// les_get_turb_visc_smag (m_les.f90:1142-1215): turb_visc from fields(1:3),
// using wrk(4:6) as work space like the reference.
static void les_turb_visc() {
    Lib& L = lib();
    auto& W = L.W;
    const double* U[3] = {L.F[1], L.F[2], L.F[3]};
    const double coef = (L.c_smag * L.dx) * (L.c_smag * L.dx);
    for (int group = 1; group <= 2; ++group) {
        {
            double* out[3] = {W[4], W[5], W[6]};
            Timed t(TC_POINTWISE);
            launch_strain(L.g, U, out, group, L.st);
        }
        for (int c = 4; c <= 6; ++c) fft_inverse(W[c]);
        Timed t(TC_POINTWISE);
        launch_tv_accum(L.g, W[4], W[5], W[6], L.tv, group, coef, L.st);
    }
}

static void mlp_velocity();

// les_rhs_velocity for model 1 (m_les.f90:636-690 with les_rhsv_turb_visc
// :902-1018): rhs += -d_j tau_ij, tau_ij = -2 nu_t S_ij, then the ialias mask.
static void les_velocity() {
    if (lib().les_model == 11) {
        mlp_velocity();
        return;
    }
    Lib& L = lib();
    auto& W = L.W;
    const double* U[3] = {L.F[1], L.F[2], L.F[3]};
    for (int group = 1; group <= 2; ++group) {
        {
            double* out[3] = {W[4], W[5], W[6]};
            Timed t(TC_POINTWISE);
            launch_strain(L.g, U, out, group, L.st);
        }
        for (int c = 4; c <= 6; ++c) {
            fft_inverse(W[c]);
            {
                Timed t(TC_POINTWISE);
                launch_tau(L.g, W[c], L.tv, L.st);
            }
            fft_forward(W[c]);
        }
        const double* tau[3] = {W[4], W[5], W[6]};
        double* rhs[3] = {W[1], W[2], W[3]};
        Timed t(TC_POINTWISE);
        launch_les_vel_accum(L.g, tau, rhs, group, L.st);
    }
    Timed t(TC_POINTWISE);
    for (int c = 1; c <= 3; ++c) launch_mask_alias(L.g, W[c], 0, L.st);
}

// m_les_rhss_turb_visc1 (m_les.f90:1041-1091) for field slot n, then the
// ialias mask of les_rhs_scalars (:863-871).
static void les_scalar(int n) {
    Lib& L = lib();
    auto& W = L.W;
    const int t1 = 3 + L.ns + L.nles + 1, t2 = t1 + 1;
    copy_slot(W[t1], L.F[n]);
    zero_slot(W[0]);
    for (int axis = 1; axis <= 3; ++axis) {
        {
            Timed t(TC_POINTWISE);
            launch_derivative(L.g, W[t1], W[t2], axis, L.st);
        }
        fft_inverse(W[t2]);
        pmul(W[t2], L.tv, W[t2]);
        fft_forward(W[t2]);
        Timed t(TC_POINTWISE);
        launch_derivative(L.g, W[t2], W[t2], axis, L.st);
        launch_add(L.g, W[0], W[t2], false, L.st);
    }
    Timed t(TC_POINTWISE);
    launch_add(L.g, W[n], W[0], true, L.st);
    launch_mask_alias(L.g, W[n], 0, L.st);
}

// nine physical velocity gradients du_i/dx_j (i-major) of fields(1:3) into
// the closure's buffers 0..8
static void mlp_alloc_scratch() {
    auto& B = lib().mlp.buf;
    if (B.empty()) {
        B.assign(15, nullptr);
        for (auto& q : B) q = dalloc(slot_bytes());
    }
}

static void mlp_gradients() {
    Lib& L = lib();
    auto& B = L.mlp.buf;
    mlp_alloc_scratch();
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            {
                Timed t(TC_POINTWISE);
                launch_derivative(L.g, L.F[c + 1], B[3 * c + a], a + 1, L.st);
            }
            fft_inverse(B[3 * c + a]);
        }
}

// The learned closure's stress divergence: tau_ij from the MLP at every point,
// rhs_i += -i k_j tau_ij on unaliased modes as for model 1.  A step whose
// network output is not finite everywhere uses the Smagorinsky stress
// instead (the fallback) and is counted.
static void mlp_velocity() {
    Lib& L = lib();
    auto& M = L.mlp;
    mlp_gradients();
    const double* gr[9];
    for (int i = 0; i < 9; ++i) gr[i] = M.buf[i];
    double* tau[6];
    for (int i = 0; i < 6; ++i) tau[i] = M.buf[9 + i];
    check(hipMemsetAsync(M.bad, 0, sizeof(unsigned int), L.st), "hipMemsetAsync");
    {
        Timed t(TC_POINTWISE);
        launch_mlp_les(L.g, gr, M.delta, M.norm, M.wb, M.p, tau, M.bad, L.st);
    }
    unsigned int bad = 0;
    check(hipMemcpyAsync(&bad, M.bad, sizeof(bad), hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
    check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    if (bad != 0) {
        ++M.fallbacks;
        if (!L.tv) L.tv = dalloc(slot_bytes());
        const int keep = L.les_model;
        L.les_model = 1;
        les_turb_visc();
        les_velocity();
        L.les_model = keep;
        return;
    }
    for (int i = 9; i < 15; ++i) fft_forward(M.buf[i]);
    double* rhs[3] = {L.W[1], L.W[2], L.W[3]};
    const double* t1[3] = {M.buf[9], M.buf[10], M.buf[11]};   // tau11, tau12, tau13
    const double* t2[3] = {M.buf[12], M.buf[13], M.buf[14]};  // tau22, tau23, tau33
    Timed t(TC_POINTWISE);
    launch_les_vel_accum(L.g, t1, rhs, 1, L.st);
    launch_les_vel_accum(L.g, t2, rhs, 2, L.st);
    for (int c = 1; c <= 3; ++c) launch_mask_alias(L.g, L.W[c], 0, L.st);
}

static void do_rhs_scalars(bool int_scalars) {
    Lib& L = lib();
    auto& F = L.F;
    auto& W = L.W;
    const Geo& g = L.g;
    if (L.nles > 0) throw std::runtime_error("LES scalars are not supported");
    if (L.les_model == 1) les_turb_visc();  // rhs_scalars.f90:18
    // scalar s of type 0 has the mean-gradient source -u (rhs_scalars.f90:387-389),
    // subtracted inside the kernel that completes its RHS
    auto source = [&](int s) -> const double* { return L.stype[s - 1] == 0 ? F[1] : nullptr; };
    if (!int_scalars || L.ns == 0) {  // rhs_scalars.f90:32-39
        for (int c = 1; c <= 3; ++c) fft_inverse_from(F[c], W[c], 0);
        return;
    }
    if (L.dealias == 0) {  // rhs_scalars.f90:48-134
        for (int c = 1; c <= 3; ++c) fft_inverse_from(F[c], W[c], 0);
        for (int s = 1; s <= L.ns; ++s) {
            fft_inverse_from(F[3 + s], W[0], 0);
            // products phi u_i straight into the forward x pass
            for (int i = 1; i <= 3; ++i) fft_forward_product(W[0], W[i], W[s + 2 + i], 0);
            const double* Q[3] = {W[s + 3], W[s + 4], W[s + 5]};
            {
                Timed t(TC_POINTWISE);
                launch_scalar_rhs(g, Q, F[3 + s], W[3 + s], L.pe[s - 1], source(s), L.st);
            }
            if (L.stype[s - 1] >= 100) {  // add_reaction + dealias_rhs, :124-130
                {
                    Timed t(TC_POINTWISE);
                    launch_reaction(g, W[0], W[0], (int)(L.stype[s - 1] / 100), L.reac[s - 1],
                                    F[3 + s], L.st);
                }
                fft_forward(W[0]);
                Timed t(TC_POINTWISE);
                launch_add(g, W[3 + s], W[0], false, L.st);
                launch_truncate(g, W[3 + s], L.kmax, L.st);
            }
        }
    } else {  // rhs_scalars.f90:144-271
        // the shift exp(i theta) on the first inverse pass, the products in
        // the forward x pass, the shift back on the last forward pass
        for (int c = 1; c <= 3 + L.ns; ++c) fft_inverse_from(F[c], W[c], +1);
        const int n1 = L.ns + 4, n2 = L.ns + 5;
        for (int c = 4; c <= 3 + L.ns; ++c) {
            fft_forward_product(W[c], W[1], W[n1], -1);
            fft_forward_product(W[c], W[2], W[n2], -1);
            fft_forward_product(W[c], W[3], W[c], -1);
            const double* Q[3] = {W[n1], W[n2], W[c]};
            Timed t(TC_POINTWISE);
            launch_scalar_rhs_ps(g, Q, F[c], W[c], 0.0, false, nullptr, L.st);
        }
        for (int c = 1; c <= 3; ++c) fft_inverse_from(F[c], W[c], 0);
        for (int c = 4; c <= 3 + L.ns; ++c) {
            fft_inverse_from(F[c], W[0], 0);
            fft_forward_product(W[0], W[1], W[n1], 0);
            fft_forward_product(W[0], W[2], W[n2], 0);
            fft_forward_product(W[0], W[3], W[0], 0);
            const double* Q[3] = {W[n1], W[n2], W[0]};
            Timed t(TC_POINTWISE);
            launch_scalar_rhs_ps(g, Q, F[c], W[c], L.pe[c - 4], true, source(c - 3), L.st);
        }
        for (int c = 4; c <= 3 + L.ns; ++c) {  // reaction_rates, rhs_scalars.f90:280-380
            if (L.stype[c - 4] <= 300) continue;
            copy_slot(W[n1], F[c]);
            {
                Timed t(TC_POINTWISE);
                launch_truncate(g, W[n1], L.n / 4, L.st);
            }
            fft_inverse(W[n1]);
            {
                Timed t(TC_POINTWISE);
                launch_reaction(g, W[n1], W[n2], 3, L.reac[c - 4], F[c], L.st);
            }
            fft_forward(W[n2]);
            Timed t(TC_POINTWISE);
            launch_add(g, W[c], W[n2], false, L.st);
        }
    }
    if (L.les_model == 1)  // rhs_scalars.f90:396-398
        for (int s = 1; s <= L.ns; ++s) les_scalar(3 + s);
}

static double read_courant_max() {
    Lib& L = lib();
    check(hipMemcpyAsync(L.hpin, L.dscal, sizeof(double), hipMemcpyDeviceToHost, L.st),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    uint64_t bits;
    std::memcpy(&bits, L.hpin, sizeof(bits));
    double m;
    std::memcpy(&m, &bits, sizeof(m));
    return m;
}

static void do_rhs_velocity() {
    Lib& L = lib();
    auto& F = L.F;
    auto& W = L.W;
    const Geo& g = L.g;
    {
        double* out[6] = {W[1], W[2], W[3], W[4], W[5], W[6]};
        Timed t(TC_POINTWISE);
        launch_products(g, W[1], W[2], W[3], out, L.dmax, L.dscal, L.st);
    }
    for (int c = 1; c <= 6; ++c) fft_forward(W[c]);
    {
        const double* P[6] = {W[1], W[2], W[3], W[4], W[5], W[6]};
        const double* U[3] = {F[1], F[2], F[3]};
        double* out[3] = {W[1], W[2], W[3]};
        Timed t(TC_POINTWISE);
        launch_rhs_assemble(g, P, U, out, L.nu, L.dealias == 1, L.st);
    }
    if (L.dealias == 1) {  // rhs_velocity.f90:242-408
        // shift on the first inverse pass, products in the forward x pass,
        // shift back on the last forward pass
        // and the half-RHS accumulation in the last forward pass's epilogue
        fft_inverse_from(F[1], W[4], +1);
        fft_inverse_from(F[2], W[5], +1);
        // (the last accumulation into each component applies the mask of
        // rhs_velocity.f90:400-404)
        fft_forward_product_acc(W[4], W[4], W[6], W[1], 1, false);
        fft_forward_product_acc(W[4], W[5], W[6], W[1], 2, false, W[2], 1, false);
        fft_forward_product_acc(W[5], W[5], W[6], W[2], 2, false);
        fft_inverse_from(F[3], W[6], +1);
        fft_forward_product_acc(W[4], W[6], W[4], W[1], 3, true, W[3], 1, false);
        fft_forward_product_acc(W[5], W[6], W[5], W[2], 3, true, W[3], 2, false);
        fft_forward_product_acc(W[6], W[6], W[6], W[3], 3, true);
    }
    if (L.les_model == 1 || L.les_model == 11) les_velocity();  // rhs_velocity.f90:414-416
}

// band energy into dscal[1] and the forcing added to the RHS (no host sync)
static void force_enqueue() {
    Lib& L = lib();
    if (L.kfmax < 1) throw std::runtime_error("force_velocity needs kfmax >= 1");
    Timed t(TC_REDUCE);
    launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kfmax, L.dshell, L.st);
    const double* f[3] = {L.F[1], L.F[2], L.F[3]};
    double* r[3] = {L.W[1], L.W[2], L.W[3]};
    launch_force_apply(L.g, f, r, L.dshell, L.kfmax, L.famp, L.dscal + 1, L.st);
}

static void do_force(double* energy) {
    Lib& L = lib();
    force_enqueue();
    if (energy) {
        check(hipMemcpyAsync(L.hpin + 1, L.dscal + 1, sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        *energy = L.hpin[1];
    }
}

static void do_advance(int which, bool first, double dt, const double* dtp = nullptr) {
    Lib& L = lib();
    int lo = 1, hi = 3;
    if (which == 2) {
        lo = 4;
        hi = 3 + L.ns + L.nles;
    } else if (which != 1) {
        throw std::runtime_error("advance: which must be 1 (velocity) or 2 (scalars)");
    }
    Timed t(TC_POINTWISE);
    for (int c = lo; c <= hi; ++c)
        launch_advance(L.g, L.F[c], L.W[c], L.R[c], dt, dtp, first, L.st);
}

static void do_pressure() {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_pressure(L.g, L.F[1], L.F[2], L.F[3], L.st);
}

static double key_to_double(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// min and max of (a [+ b + c]) in physical space, stored to sum if given.
static void minmax_to_host(const double* a, const double* b, const double* c, double* sum,
                           double* mn, double* mx) {
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

static void sums_to_host(const double* a, const double* b, const double* c, double out[2]) {
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
static void do_stat_scalars(double* diss, double* mean, double* var, double* mn, double* mx,
                            double* es) {
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
static void do_divergence(double* dmin, double* dmax) {
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

using namespace hit3d;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
template <typename Fn>
static int guarded(Fn&& fn, bool device = true) {
    try {
        fn();
        const hipError_t e = device ? hipGetLastError() : hipSuccess;
        if (e != hipSuccess) throw HipError{std::string("kernel launch: ") + hipGetErrorString(e)};
        return 0;
    } catch (const HipError& e) {
        lib().err = e.msg;
        return 2;
    } catch (const std::exception& e) {
        lib().err = e.what();
        return 1;
    } catch (...) {
        lib().err = "unknown error";
        return 3;
    }
}

extern "C" {

int hit3d_init(const int64_t* nx, const int64_t* ny, const int64_t* nz_all,
               const int64_t* n_scalars, const int64_t* n_les, const int64_t* dealias) {
    return guarded([&] {
        Lib& L = lib();
        if (L.ready) free_all();
        if (*nx != *ny || *nx != *nz_all)
            throw std::runtime_error("only cubic grids nx = ny = nz_all are supported");
        if (!supported_n((int)*nx))
            throw std::runtime_error("grid size must be a power of two in [8, 1024]");
        if (*n_scalars < 0 || *n_les < 0) throw std::runtime_error("negative field count");
        if (*dealias != 0 && *dealias != 1) throw std::runtime_error("dealias must be 0 or 1");
        L.n = (int)*nx;
        L.g.n = L.n;
        L.g.P = row_pitch(L.n);
        L.g.nc = (int64_t)L.n * L.n * L.g.P;
        L.ns = (int)*n_scalars;
        L.nles = (int)*n_les;
        L.nf = 3 + L.ns + L.nles;
        L.nw = std::max(6, 3 + L.ns + L.nles + 2);
        L.dealias = (int)*dealias;
        L.kmax = L.dealias == 0 ? L.n / 3 : (int)std::floor((double)L.n / 3.0 * std::sqrt(2.0));
        const double pi = 4.0 * std::atan(1.0);
        L.dx = 2.0 * pi / (double)L.n;
        check(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking), "hipStreamCreate");
        const size_t sb = slot_bytes();
        L.F.assign(L.nf + 1, nullptr);
        L.R.assign(L.nf + 1, nullptr);
        L.W.assign(L.nw + 1, nullptr);
        for (int i = 1; i <= L.nf; ++i) {
            L.F[i] = dalloc(sb);
            L.R[i] = dalloc(sb);
        }
        for (int i = 0; i <= L.nw; ++i) L.W[i] = dalloc(sb);
        L.spare = dalloc(sb);
        // twiddles exp(-2 pi i j / len) for every power of two len <= kMaxLen
        std::vector<cpx<double>> t64(2 * kMaxLen);
        std::vector<cpx<float>> t32(2 * kMaxLen);
        for (int len = 1; len <= kMaxLen; len *= 2)
            for (int j = 0; j < len; ++j) {
                const long double a = -2.0L * 3.14159265358979323846264338327950288L * j / len;
                t64[len + j] = {(double)cosl(a), (double)sinl(a)};
                t32[len + j] = {(float)cosl(a), (float)sinl(a)};
            }
        L.tw64 = reinterpret_cast<cpx<double>*>(dalloc(t64.size() * sizeof(cpx<double>)));
        L.tw32 = reinterpret_cast<cpx<float>*>(dalloc(t32.size() * sizeof(cpx<float>)));
        check(hipMemcpy(L.tw64, t64.data(), t64.size() * sizeof(cpx<double>), hipMemcpyHostToDevice),
              "hipMemcpy");
        check(hipMemcpy(L.tw32, t32.data(), t32.size() * sizeof(cpx<float>), hipMemcpyHostToDevice),
              "hipMemcpy");
        // phase-shift factors cos/sin(half*(kx+ky+kz)*dx), rhs_velocity.f90:248-249
        const int ntab = 3 * L.n;
        std::vector<double> tab(2 * ntab);
        for (int t = 0; t < ntab; ++t) {
            const double ks = (double)(t - L.n);
            const double a = 0.5 * ks * L.dx;
            tab[2 * t] = std::cos(a);
            tab[2 * t + 1] = std::sin(a);
        }
        L.ps_tab = dalloc(tab.size() * sizeof(double));
        check(hipMemcpy(L.ps_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy");
        L.dshell = dalloc((size_t)(L.n + 2) * sizeof(double) * 2);
        L.dpart = dalloc((size_t)L.n * 2 * sizeof(double));
        L.dmax = dalloc(std::max<size_t>(65536, (size_t)L.n * L.n) * sizeof(double));
        L.dscal = dalloc(16 * sizeof(double));
        check(hipHostMalloc((void**)&L.hpin, 16 * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        L.nu = 0.0;
        L.ready = true;
    });
}

int hit3d_set_physics(const double* nu, const int64_t* flow_type, const int64_t* kfmax,
                      const double* famp, const double* pe, const int64_t* scalar_type) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        clear_graphs();  // captured steps hold the old parameters as kernel arguments
        L.nu = *nu;
        L.flow_type = (int)*flow_type;
        L.kfmax = (int)*kfmax;
        L.famp = *famp;
        if (L.kfmax > L.kmax) throw std::runtime_error("kfmax exceeds kmax");
        L.pe.assign(L.ns, 0.0);
        L.stype.assign(L.ns, 0);
        L.reac.assign(L.ns, 0.0);
        for (int s = 0; s < L.ns; ++s) {
            if (!pe || !scalar_type) throw std::runtime_error("scalar parameters missing");
            L.pe[s] = pe[s];
            L.stype[s] = scalar_type[s];
            const int64_t rtype = L.stype[s] / 100;
            if (L.stype[s] < 0 || rtype > 3)
                throw std::runtime_error("scalar type: reaction digit must be 0..3");
            if (L.dealias == 1 && L.stype[s] > 300 && L.stype[s] != 311)
                throw std::runtime_error(
                    "The scalar type has a reaction rate that is not supported yet (phase "
                    "shifting supports 311 only)");
        }
    });
}

int hit3d_set_reaction(const double* reac_sc) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if ((int)L.reac.size() != L.ns) throw std::runtime_error("set_reaction: set_physics first");
        clear_graphs();
        for (int s = 0; s < L.ns; ++s) L.reac[s] = reac_sc[s];
    }, false);
}

int hit3d_finalize(void) {
    return guarded([&] { free_all(); });
}

int hit3d_put(const int64_t* array_id, const int64_t* slot, const double* host) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double* d = slot_ref(*array_id, *slot);
        const size_t w = (size_t)(L.n + 2) * sizeof(double);
        check(hipMemcpy2DAsync(d, (size_t)L.g.P * 2 * sizeof(double), host, w, w,
                               (size_t)L.n * L.n, hipMemcpyHostToDevice, L.st),
              "hipMemcpy2DAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_get(const int64_t* array_id, const int64_t* slot, double* host) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double* d = slot_ref(*array_id, *slot);
        const size_t w = (size_t)(L.n + 2) * sizeof(double);
        check(hipMemcpy2DAsync(host, w, d, (size_t)L.g.P * 2 * sizeof(double), w,
                               (size_t)L.n * L.n, hipMemcpyDeviceToHost, L.st),
              "hipMemcpy2DAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_copy(const int64_t* dst_array, const int64_t* dst_slot, const int64_t* src_array,
               const int64_t* src_slot) {
    return guarded([&] {
        require_ready();
        double* d = slot_ref(*dst_array, *dst_slot);
        double* s = slot_ref(*src_array, *src_slot);
        if (d != s) copy_slot(d, s);
    });
}

int hit3d_xfft3d(const int64_t* flag, const int64_t* n) {
    return guarded([&] {
        require_ready();
        double*& s = slot_ref(2, *n);
        if (*flag == 1) fft_forward(s);
        else if (*flag == -1) fft_inverse(s);
        else throw std::runtime_error("xFFT3d: flag must be +1 or -1");
    });
}

int hit3d_xfft3d_fields(const int64_t* flag, const int64_t* n) {
    return guarded([&] {
        require_ready();
        double*& s = slot_ref(1, *n);
        if (*flag == 1) fft_forward(s);
        else if (*flag == -1) fft_inverse(s);
        else throw std::runtime_error("xFFT3d_fields: flag must be +1 or -1");
    });
}

int hit3d_x_derivative(const int64_t* n, const int64_t* axis, const int64_t* nto) {
    return guarded([&] {
        require_ready();
        if (*axis < 1 || *axis > 3) throw std::runtime_error("x_derivative: wrong value of axis");
        Lib& L = lib();
        double* a = slot_ref(2, *n);
        double* b = slot_ref(2, *nto);
        Timed t(TC_POINTWISE);
        launch_derivative(L.g, a, b, (int)*axis, L.st);
    });
}

int hit3d_dealias_all(void) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        Timed t(TC_POINTWISE);
        for (int c = 1; c <= 3 + L.ns; ++c) launch_truncate(L.g, L.F[c], L.kmax, L.st);
    });
}

int hit3d_rhs_scalars(const int64_t* int_scalars) {
    return guarded([&] {
        require_ready();
        do_rhs_scalars(*int_scalars != 0);
    });
}

int hit3d_rhs_velocity(const double* dt, double* courant) {
    return guarded([&] {
        require_ready();
        do_rhs_velocity();
        const double m = read_courant_max();
        if (courant) *courant = m * (*dt) / lib().dx;
    });
}

int hit3d_force_velocity(double* energy) {
    return guarded([&] {
        require_ready();
        do_force(energy);
    });
}

int hit3d_advance(const int64_t* which, int64_t* first_order, const double* dt) {
    return guarded([&] {
        require_ready();
        do_advance((int)*which, *first_order != 0, *dt);
        *first_order = 0;
    });
}

int hit3d_pressure(void) {
    return guarded([&] {
        require_ready();
        do_pressure();
    });
}

// The device work of one driver step (main.f90:124-167), enqueued without any
// host synchronisation; the step size is read from dscal[12].
static void step_body(bool fused, bool mixed, bool scal, bool first_v, bool first_s) {
    Lib& L = lib();
    const double* dtp = L.dscal + 12;
    if (fused) {
        // rhs_scalars + rhs_velocity + force + update + pressure in one pipeline
        fused_step(dtp, first_v, mixed);
        return;
    }
    do_rhs_scalars(scal);
    if (scal) do_advance(2, first_s, 0.0, dtp);
    do_rhs_velocity();
    if (L.flow_type == 1) force_enqueue();
    {  // main.f90:156-164 and pressure.f90:13-47 in one pass
        Timed t(TC_POINTWISE);
        double* f[3] = {L.F[1], L.F[2], L.F[3]};
        const double* r[3] = {L.W[1], L.W[2], L.W[3]};
        double* ro[3] = {L.R[1], L.R[2], L.R[3]};
        launch_advance_project(L.g, f, r, ro, 0.0, dtp, first_v, L.st);
    }
}

static std::vector<void*> slot_state() {
    Lib& L = lib();
    std::vector<void*> s;
    for (auto* p : L.F) s.push_back(p);
    for (auto* p : L.R) s.push_back(p);
    for (auto* p : L.W) s.push_back(p);
    for (auto* p : L.W32) s.push_back(p);
    for (auto* p : L.mlp.buf) s.push_back(p);  // closure scratch, transformed in place
    s.push_back(L.spare);
    return s;
}

static void set_slot_state(const std::vector<void*>& s) {
    Lib& L = lib();
    size_t i = 0;
    for (auto*& p : L.F) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.R) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.W) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.W32) p = static_cast<float*>(s[i++]);
    for (auto*& p : L.mlp.buf) p = static_cast<double*>(s[i++]);
    L.spare = static_cast<double*>(s[i++]);
}

// Puts the work-slot, closure-scratch and spare pointers back to the state s
// (from slot_state); fields and rhs_old must not have moved.
static void restore_work_slots(const std::vector<void*>& s) {
    Lib& L = lib();
    if (s.size() != slot_state().size())
        throw std::runtime_error("internal: scratch slots were allocated during the step");
    size_t i = 0;
    for (auto* p : L.F)
        if (p != s[i++]) throw std::runtime_error("internal: a field slot moved during the step");
    for (auto* p : L.R)
        if (p != s[i++]) throw std::runtime_error("internal: an rhs_old slot moved during the step");
    for (auto*& p : L.W) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.W32) p = static_cast<float*>(s[i++]);
    for (auto*& p : L.mlp.buf) p = static_cast<double*>(s[i++]);
    L.spare = static_cast<double*>(s[i++]);
}

// step_body from a captured graph: replayed when one was captured from the
// same slot pointers and flags, captured when this combination was already
// run once eagerly (which also makes every lazy allocation happen outside a
// capture), run eagerly otherwise.
static void step_graph(bool fused, bool mixed, bool scal, bool first_v, bool first_s) {
    Lib& L = lib();
    const unsigned flags = (fused ? 1u : 0u) | (mixed ? 2u : 0u) | (scal ? 4u : 0u) |
                           (first_v ? 8u : 0u) | (first_s ? 16u : 0u);
    const std::vector<void*> before = slot_state();
    for (auto& g : L.graphs)
        if (g.flags == flags && g.before == before) {
            check(hipGraphLaunch(g.exec, L.st), "hipGraphLaunch");
            set_slot_state(g.after);
            ++L.graph_replays;
            return;
        }
    bool seen = false;
    for (auto& k : L.graph_seen) seen = seen || (k.second == flags && k.first == before);
    if (!seen || L.graphs.size() >= 16) {
        if (!seen && L.graph_seen.size() < 64) L.graph_seen.push_back({before, flags});
        step_body(fused, mixed, scal, first_v, first_s);
        return;
    }
    hipGraph_t graph = nullptr;
    check(hipStreamBeginCapture(L.st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
        step_body(fused, mixed, scal, first_v, first_s);
    } catch (...) {
        (void)hipStreamEndCapture(L.st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        set_slot_state(before);
        throw;
    }
    check(hipStreamEndCapture(L.st, &graph), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) set_slot_state(before);
    check(e, "hipGraphInstantiate");
    L.graphs.push_back({before, slot_state(), flags, exec});
    check(hipGraphLaunch(exec, L.st), "hipGraphLaunch");
    ++L.graph_replays;
}

int hit3d_step(const double* dt, int64_t* fov, int64_t* fos, const int64_t* int_scalars,
               const int64_t* precision, double* courant, double* energy) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const bool mixed = *precision == 1;
        if (*precision != 0 && *precision != 1)
            throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        const bool scal = (*int_scalars != 0) && L.ns > 0;
        const bool fused = !scal && L.fuse && L.les_model == 0 && fused_step_supported(mixed);
        if (!fused && mixed) throw std::runtime_error("mixed precision needs the fused path");
        const bool first_v = *fov != 0, first_s = *fos != 0;
        // the step size goes to the device (dscal[12]) ahead of the step's
        // work, so a captured step replays with the current dt
        L.hpin[12] = *dt;
        check(hipMemcpyAsync(L.dscal + 12, L.hpin + 12, sizeof(double), hipMemcpyHostToDevice,
                             L.st),
              "hipMemcpyAsync");
        // The out-of-place passes swap work slots with the spare; fields and
        // rhs_old never move.  wrk is scratch of the step (undefined after
        // it), so the work-slot labels are put back afterwards: every step
        // then starts from the same pointers and one captured graph serves
        // them all.
        if (L.les_model == 11) mlp_alloc_scratch();  // before the pointer snapshot
        const std::vector<void*> s0 = slot_state();
        // the learned closure checks its output on the host mid-step
        if (L.graph_mode && !L.timing && L.les_model != 11)
            step_graph(fused, mixed, scal, first_v, first_s);
        else
            step_body(fused, mixed, scal, first_v, first_s);
        restore_work_slots(s0);
        *fov = 0;
        if (scal) *fos = 0;
        // Courant max (dscal[0], order-free bits of a non-negative double)
        // and band energy (dscal[1]) back in one copy
        check(hipMemcpyAsync(L.hpin, L.dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        double cmax;
        std::memcpy(&cmax, L.hpin, sizeof(cmax));
        if (energy && L.flow_type == 1) *energy = L.hpin[1];
        if (courant) *courant = cmax * (*dt) / L.dx;
    });
}

int hit3d_stats(double* e_spec, double* enstrophy) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        {
            Timed t(TC_REDUCE);
            launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kmax, L.dshell, L.st);
            launch_enstrophy(L.g, L.F[1], L.F[2], L.F[3], L.dpart, L.dscal + 2, L.st);
        }
        if (e_spec)
            check(hipMemcpyAsync(e_spec, L.dshell, (size_t)L.kmax * sizeof(double),
                                 hipMemcpyDeviceToHost, L.st),
                  "hipMemcpyAsync");
        check(hipMemcpyAsync(L.hpin + 2, L.dscal + 2, sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        if (enstrophy) *enstrophy = L.hpin[2];
    });
}

int hit3d_particles_set(const int64_t* np, const double* xyzp, const double* uvwp) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int64_t cnt = *np;
        if (cnt < 0) throw std::runtime_error("particles_set: negative particle count");
        const double hi = (double)(L.n + 1);
        for (int64_t i = 0; i < 3 * cnt; ++i)
            if (!(xyzp[i] >= 1.0 && xyzp[i] < hi))
                throw std::runtime_error("particles_set: position outside [1, n+1) grid units");
        if (cnt > L.np_cap) {
            if (L.xyzp) check(hipFree(L.xyzp), "hipFree");
            if (L.uvwp) check(hipFree(L.uvwp), "hipFree");
            L.xyzp = L.uvwp = nullptr;
            L.np_cap = 0;
            L.xyzp = static_cast<double*>(dalloc((size_t)cnt * 3 * sizeof(double)));
            L.uvwp = static_cast<double*>(dalloc((size_t)cnt * 3 * sizeof(double)));
            L.np_cap = cnt;
        }
        L.np = cnt;
        if (cnt == 0) return;
        const size_t b = (size_t)cnt * 3 * sizeof(double);
        check(hipMemcpyAsync(L.xyzp, xyzp, b, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        if (uvwp) {
            check(hipMemcpyAsync(L.uvwp, uvwp, b, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        } else {  // particles_init: uvwp = -999 (m_particles.f90:327)
            std::vector<double> fill((size_t)cnt * 3, -999.0);
            check(hipMemcpyAsync(L.uvwp, fill.data(), b, hipMemcpyHostToDevice, L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        }
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_particles_get(double* xyzp, double* uvwp) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const size_t b = (size_t)L.np * 3 * sizeof(double);
        if (b == 0) return;
        if (xyzp) check(hipMemcpyAsync(xyzp, L.xyzp, b, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        if (uvwp) check(hipMemcpyAsync(uvwp, L.uvwp, b, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_fields_to_parts(void) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        for (int c = 1; c <= 3; ++c) {
            copy_slot(L.W[c], L.F[c]);
            fft_inverse(L.W[c]);
        }
    });
}

int hit3d_particles_move(const double* dt, const int64_t* scheme) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*scheme != 0 && *scheme != 1)
            throw std::runtime_error("particles_tracking_scheme must be 0 (trilinear) or 1 (CINT)");
        if (*scheme == 1 && L.n < 4) throw std::runtime_error("CINT needs n >= 4");
        Timed t(TC_PARTICLES);
        launch_particles_move(L.g, L.W[1], L.W[2], L.W[3], L.xyzp, L.uvwp, L.np, *dt, L.dx,
                              (int)*scheme, L.st);
    });
}

int hit3d_stat_scalars(double* sc_diss, double* mean, double* variance, double* sc_min,
                       double* sc_max, double* es_spec) {
    return guarded([&] {
        require_ready();
        do_stat_scalars(sc_diss, mean, variance, sc_min, sc_max, es_spec);
    });
}

int hit3d_skewness(double* skew) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        double tot = 0.0;
        for (int c = 1; c <= 3; ++c) {
            {
                Timed t(TC_POINTWISE);
                launch_derivative(L.g, L.F[c], L.W[0], c, L.st);
            }
            fft_inverse(L.W[0]);
            {
                Timed t(TC_REDUCE);
                launch_phys_moments23(L.g, L.W[0], L.dpart, L.dscal + 4, L.st);
            }
            check(hipMemcpyAsync(L.hpin + 4, L.dscal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                 L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            const double n3 = (double)L.n * L.n * L.n;
            const double m2 = L.hpin[4] / n3, m3 = L.hpin[5] / n3;
            tot += m3 / std::pow(m2, 1.5);
        }
        *skew = tot / 3.0;
    });
}

int hit3d_divergence(double* dmin, double* dmax) {
    return guarded([&] {
        require_ready();
        do_divergence(dmin, dmax);
    });
}

int hit3d_init_velocity(const double* rn1, const int64_t* isp_type, const int64_t* ir_exp,
                        const double* peak) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int n = L.n;
        const size_t hsz = (size_t)(n + 2) * n * n;
        std::vector<double> host(hsz);
        if (*isp_type == -1) {  // init_velocity.f90:252-293
            for (int c = 0; c < 3; ++c) {
                taylor_green_slot(n, c, host.data());
                const int64_t a = 2, s = c + 1;
                if (hit3d_put(&a, &s, host.data()) != 0) throw std::runtime_error(L.err);
                fft_forward(L.W[c + 1]);
                copy_slot(L.F[c + 1], L.W[c + 1]);
            }
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            return;
        }
        // init_velocity.f90:67-120: uniforms, Box-Muller, forward transforms
        for (int c = 0; c < 3; ++c) {
            gaussian_slot(n, *rn1, c, host.data());
            const int64_t a = 2, s = c + 1;
            if (hit3d_put(&a, &s, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.W[c + 1]);
        }
        L.rng_iy = velocity_rng_iy(n, *rn1);
        // :123-149 curl on shells 1..kmax
        launch_init_curl(L.g, L.W[1], L.W[2], L.W[3], L.F[1], L.F[2], L.F[3], L.kmax, L.st);
        // :156-186 shell spectrum
        launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kmax, L.dshell, L.st);
        std::vector<double> e(L.kmax);
        check(hipMemcpyAsync(e.data(), L.dshell, L.kmax * sizeof(double), hipMemcpyDeviceToHost,
                             L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        // :192-237 target spectrum and per-shell ratio
        std::vector<double> tgt = target_spectrum(L.kmax, (int)*isp_type, *peak, (double)*ir_exp);
        std::vector<double> ratio(L.kmax + 1, -1.0);
        for (int k = 1; k <= L.kmax; ++k)
            if (e[k - 1] > 0.0) ratio[k] = std::sqrt(tgt[k - 1] / e[k - 1]);
        check(hipMemcpyAsync(L.dshell, ratio.data(), ratio.size() * sizeof(double),
                             hipMemcpyHostToDevice, L.st),
              "hipMemcpyAsync");
        for (int c = 1; c <= 3; ++c) launch_init_rescale(L.g, L.F[c], L.dshell, L.kmax, L.st);
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_init_scalar(const int64_t* n_scalar, const double* rn2, const double* peak) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int s = (int)*n_scalar;
        if ((int)L.stype.size() != L.ns) throw std::runtime_error("init_scalar: set_physics first");
        if (s < 1 || s > L.ns) throw std::runtime_error("init_scalar: scalar number out of range");
        const int n = L.n;
        std::vector<double> host((size_t)(n + 2) * n * n);
        const int64_t st = L.stype[s - 1];
        const int ic = (int)(st - (st / 100) * 100);
        double* F = L.F[3 + s];
        if (ic == 0) {
            zero_slot(F);
        } else if (ic < 10) {  // init_scalar_spectrum, :84-270
            if (st < 1 || st > 3)
                throw std::runtime_error("INIT_SCALARS: WRONG INITIAL SPECTRUM TYPE");
            L.rng_iy = scalar_gaussian_slot(n, *rn2, L.rng_iy, host.data());
            const int64_t a = 2, w3 = 3;
            if (hit3d_put(&a, &w3, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.W[3]);
            {
                Timed t(TC_REDUCE);
                launch_shell_sums1(L.g, L.W[3], 1, L.kmax, L.dshell, L.st);
            }
            std::vector<double> e(L.kmax);
            check(hipMemcpyAsync(e.data(), L.dshell, L.kmax * sizeof(double),
                                 hipMemcpyDeviceToHost, L.st),
                  "hipMemcpyAsync");
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            const std::vector<double> tgt = scalar_target_spectrum(L.kmax, st, *peak);
            std::vector<double> ratio(L.kmax + 1, -1.0);
            for (int k = 1; k <= L.kmax; ++k)
                if (e[k - 1] > 0.0) ratio[k] = std::sqrt(tgt[k - 1] / e[k - 1]);
            check(hipMemcpyAsync(L.dshell, ratio.data(), ratio.size() * sizeof(double),
                                 hipMemcpyHostToDevice, L.st),
                  "hipMemcpyAsync");
            copy_slot(F, L.W[3]);
            {
                Timed t(TC_POINTWISE);
                launch_init_rescale(L.g, F, L.dshell, L.kmax, L.st);
            }
            check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
            if (st == 3) {  // double-delta PDF, :239-262
                copy_slot(L.W[0], F);
                fft_inverse(L.W[0]);
                {
                    Timed t(TC_POINTWISE);
                    launch_double_delta(L.g, L.W[0], L.st);
                }
                fft_forward(L.W[0]);
                {
                    Timed t(TC_POINTWISE);
                    launch_shell_cut(L.g, L.W[0], L.kmax * 2 / 3, L.st);
                }
                copy_slot(F, L.W[0]);
            }
        } else {  // init_scalar_space, :278-383
            scalar_space_slot(n, ic, *peak, host.data());
            const int64_t a = 1, slot = 3 + s;
            if (hit3d_put(&a, &slot, host.data()) != 0) throw std::runtime_error(L.err);
            fft_forward(L.F[3 + s]);  // may exchange the slot with the spare buffer
            if (ic == 11 || ic == 12)  // fields(1,1,1,3+n) = 0
                check(hipMemsetAsync(L.F[3 + s], 0, sizeof(double), L.st), "hipMemsetAsync");
        }
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_init_scalars(const double* rn2, const double* peak_sc) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        for (int64_t s = 1; s <= L.ns; ++s)  // init_scalar, init_scalars.f90:20-22
            if (hit3d_init_scalar(&s, rn2, peak_sc + (s - 1)) != 0) throw std::runtime_error(L.err);
        Timed t(TC_POINTWISE);  // :27-36 zero ialias > 0 on every scalar
        for (int s = 1; s <= L.ns; ++s) launch_mask_alias(L.g, L.F[3 + s], 0, L.st);
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_snapshot(const int64_t* n_field, float* out) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const int64_t f = *n_field;
        if (f < 1 || f > L.nf) throw std::runtime_error("snapshot: field number out of range");
        copy_slot(L.W[0], L.F[f]);
        {
            Timed t(TC_POINTWISE);
            launch_sphere_cut(L.g, L.W[0], L.kmax, L.st);
        }
        fft_inverse(L.W[0]);
        float* stage = reinterpret_cast<float*>(L.spare);  // free between transforms
        {
            Timed t(TC_POINTWISE);
            launch_pack_f32(L.g, L.W[0], stage, L.st);
        }
        check(hipMemcpyAsync(out, stage, sizeof(float) * (size_t)L.n * L.n * L.n,
                             hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
    });
}

int hit3d_rng_state(int64_t* iy) {
    return guarded([&] { *iy = lib().rng_iy; }, false);
}

int hit3d_scalar_gaussian(const int64_t* nx, const double* rn2, const int64_t* iy_prev,
                          double* out, int64_t* iy_next) {
    return guarded([&] {
        if (*nx < 2) throw std::runtime_error("bad grid size");
        const int64_t iy = scalar_gaussian_slot((int)*nx, *rn2, *iy_prev, out);
        if (iy_next) *iy_next = iy;
    }, false);
}

int hit3d_velocity_rng_state(const int64_t* nx, const double* rn1, int64_t* iy) {
    return guarded([&] {
        if (*nx < 2) throw std::runtime_error("bad grid size");
        *iy = velocity_rng_iy((int)*nx, *rn1);
    }, false);
}

int hit3d_random_draws(const double* seed, const int64_t* skip, const int64_t* count,
                       double* out) {
    return guarded([&] {
        if (*skip < 0 || *count < 0) throw std::runtime_error("negative draw count");
        random_draws(*seed, *skip, *count, out);
    }, false);
}

int hit3d_init_gaussian(const int64_t* nx, const double* rn1, const int64_t* c, double* out) {
    return guarded([&] {
        if (*c < 1 || *c > 3) throw std::runtime_error("component must be 1..3");
        if (*nx < 2) throw std::runtime_error("bad grid size");
        gaussian_slot((int)*nx, *rn1, (int)*c - 1, out);
    }, false);
}

int hit3d_get_ialias(int8_t* out) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        const size_t bytes = (size_t)L.n * L.n * (L.n + 2);
        void* d = nullptr;
        check(hipMalloc(&d, bytes), "hipMalloc");
        launch_ialias(L.g, static_cast<int8_t*>(d), L.st);
        check(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        check(hipFree(d), "hipFree");
    });
}

int hit3d_set_les(const int64_t* les_model, const double* c_smag) {
    return guarded([&] {
        require_ready();
        Lib& L = lib();
        if (*les_model != 0 && *les_model != 1 && *les_model != 11)
            throw std::runtime_error(
                "only les_model 0 (none), 1 (Smagorinsky) and 11 (learned closure) are supported");
        if (*les_model == 11 && !(L.mlp.have_params && L.mlp.have_norm))
            throw std::runtime_error("les_model 11: set the closure's parameters and scaling first");
        if (*les_model == 11 && L.ns > 0)
            throw std::runtime_error("les_model 11 closes the velocity only (no scalars)");
        if (*les_model != 0 && L.dealias != 0)
            throw std::runtime_error("LES needs the 2/3 rule (m_parameters.f90:415-420)");
        clear_graphs();
        L.les_model = (int)*les_model;
        L.c_smag = *c_smag;
        if (L.les_model != 0 && !L.tv) L.tv = dalloc(slot_bytes());
        if (L.les_model == 11) L.mlp.delta = L.dx;  // filter width of the LES grid
    });
}

// ------------------------------------------------------------ learned closure
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
        float* dX = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NIN * n));
        float* dT = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NOUT * n));
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
        check(hipFree(dX), "hipFree");
        check(hipFree(dT), "hipFree");
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
        float* dX = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NIN * n));
        float* dY = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NOUT * n));
        check(hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, L.st), "hipMemcpyAsync");
        launch_mlp_forward(dX, n, M.wb, M.p, dY, L.st);
        check(hipMemcpyAsync(Y, dY, sizeof(float) * mlp::NOUT * n, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        check(hipFree(dX), "hipFree");
        check(hipFree(dY), "hipFree");
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
        double* d = dalloc(sizeof(double) * 9 * m3);
        {
            Timed t(TC_POINTWISE);
            launch_box_filter(L.g, L.W[1], L.W[2], L.W[3], r, d, d + 3 * m3, L.st);
        }
        check(hipMemcpyAsync(ubar, d, sizeof(double) * 3 * m3, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipMemcpyAsync(tau, d + 3 * m3, sizeof(double) * 6 * m3, hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        check(hipFree(d), "hipFree");
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
        float* d = reinterpret_cast<float*>(dalloc(sizeof(float) * mlp::NIN * n3));
        {
            Timed t(TC_POINTWISE);
            launch_sgs_features(L.g, gr, *delta, d, L.st);
        }
        check(hipMemcpyAsync(X, d, sizeof(float) * mlp::NIN * n3, hipMemcpyDeviceToHost, L.st), "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        check(hipFree(d), "hipFree");
    });
}

int hit3d_mlp_fallbacks(int64_t* count) {
    return guarded([&] { *count = lib().mlp.fallbacks; }, false);
}

int hit3d_set_fusion(const int64_t* on) {
    return guarded([&] {
        require_ready();
        clear_graphs();
        lib().fuse = *on != 0;
    });
}

int hit3d_set_graphs(const int64_t* on) {
    return guarded([&] {
        require_ready();
        clear_graphs();
        lib().graph_mode = *on != 0;
    });
}

int hit3d_graph_stats(int64_t* captured, int64_t* replays) {
    return guarded([&] {
        require_ready();
        if (captured) *captured = (int64_t)lib().graphs.size();
        if (replays) *replays = lib().graph_replays;
    });
}

int hit3d_timing_enable(const int64_t* on) {
    return guarded([&] {
        Lib& L = lib();
        collect_timers();
        L.timing = *on != 0;
        for (int i = 0; i < TC_COUNT; ++i) {
            L.tms[i] = 0.0;
            L.tcnt[i] = 0;
        }
    });
}

int hit3d_timing_read(const int64_t* cap, double* ms, int64_t* count) {
    return guarded([&] {
        Lib& L = lib();
        collect_timers();
        for (int i = 0; i < TC_COUNT && i < *cap; ++i) {
            ms[i] = L.tms[i];
            count[i] = L.tcnt[i];
        }
    });
}

int hit3d_timing_names(char* buf, const int64_t* len) {
    return guarded([&] {
        std::strncpy(buf, kTimerNames, (size_t)*len);
        if (*len > 0) buf[*len - 1] = '\0';
    }, false);
}

int hit3d_synchronize(void) {
    return guarded([&] {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipGetLastError(), "kernel launch");
    });
}

int hit3d_last_error(char* buf, const int64_t* len) {
    if (*len <= 0) return 0;
    std::strncpy(buf, lib().err.c_str(), (size_t)*len);
    buf[*len - 1] = '\0';
    return 0;
}

}  // extern "C"
