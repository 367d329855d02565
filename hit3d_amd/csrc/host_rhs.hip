// The reference subroutines of the time step as launch sequences:
// rhs_scalars, rhs_velocity, force_velocity, the AB2 update, pressure and
// the LES closures.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace hit3d {

// ---------------------------------------------------------------------------
// reference subroutines
// ---------------------------------------------------------------------------
// les_get_turb_visc_smag (m_les.f90:1142-1215): turb_visc from fields(1:3),
// using wrk(4:6) as work space like the reference.
void les_turb_visc() {
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
        double** b[3] = {&W[4], &W[5], &W[6]};
        fft_inverse_batch(b, 3);
        Timed t(TC_POINTWISE);
        launch_tv_accum(L.g, W[4], W[5], W[6], L.tv, group, coef, L.st);
    }
}


// les_rhs_velocity for model 1 (m_les.f90:636-690 with les_rhsv_turb_visc
// :902-1018): rhs += -d_j tau_ij, tau_ij = -2 nu_t S_ij, then the ialias mask.
void les_velocity() {
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
        double** b[3] = {&W[4], &W[5], &W[6]};
        fft_inverse_batch(b, 3);
        {
            Timed t(TC_POINTWISE);
            for (int c = 4; c <= 6; ++c) launch_tau(L.g, W[c], L.tv, L.st);
        }
        fft_forward_batch(b, 3);
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
void les_scalar(int n) {
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
void mlp_alloc_scratch() {
    auto& B = lib().mlp.buf;
    if (B.empty()) {
        B.assign(15, nullptr);
        for (auto& q : B) q = dalloc(slot_bytes());
    }
}

void mlp_gradients() {
    Lib& L = lib();
    auto& B = L.mlp.buf;
    mlp_alloc_scratch();
    for (int c = 0; c < 3; ++c) {
        for (int a = 0; a < 3; ++a) {
            Timed t(TC_POINTWISE);
            launch_derivative(L.g, L.F[c + 1], B[3 * c + a], a + 1, L.st);
        }
        double** b[3] = {&B[3 * c], &B[3 * c + 1], &B[3 * c + 2]};
        fft_inverse_batch(b, 3);
    }
}

// The learned closure's stress divergence: tau_ij from the MLP at every point,
// rhs_i += -i k_j tau_ij on unaliased modes as for model 1.  A step whose
// network output is not finite everywhere uses the Smagorinsky stress
// instead (the fallback) and is counted.
void mlp_velocity() {
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
        const double cd = L.c_smag * L.dx;
        launch_mlp_les(L.g, gr, M.delta, M.norm, M.wb, M.p, tau, M.bad, M.mix, cd * cd, L.st);
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
    for (int i = 9; i < 15; i += 3) {
        double** b[3] = {&M.buf[i], &M.buf[i + 1], &M.buf[i + 2]};
        fft_forward_batch(b, 3);
    }
    double* rhs[3] = {L.W[1], L.W[2], L.W[3]};
    const double* t1[3] = {M.buf[9], M.buf[10], M.buf[11]};   // tau11, tau12, tau13
    const double* t2[3] = {M.buf[12], M.buf[13], M.buf[14]};  // tau22, tau23, tau33
    Timed t(TC_POINTWISE);
    launch_les_vel_accum(L.g, t1, rhs, 1, L.st);
    launch_les_vel_accum(L.g, t2, rhs, 2, L.st);
    for (int c = 1; c <= 3; ++c) launch_mask_alias(L.g, L.W[c], 0, L.st);
}

// physical velocity in wrk(1:3) from fields(1:3), one batch
static void inverse_velocity(std::vector<double*>& F, std::vector<double*>& W) {
    const double* src[3] = {F[1], F[2], F[3]};
    double** dst[3] = {&W[1], &W[2], &W[3]};
    fft_inverse_from_batch(src, dst, 3);
}

void keep_slots() {
    Lib& L = lib();
    if (L.keep.empty()) {
        L.keep.assign(3, nullptr);
        for (auto& q : L.keep) q = dalloc(slot_bytes());
    }
}

void do_rhs_scalars(bool int_scalars, bool keep_shifted) {
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
        inverse_velocity(F, W);
        return;
    }
    if (L.dealias == 0) {  // rhs_scalars.f90:48-134
        inverse_velocity(F, W);
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
        if (keep_shifted)  // the shifted velocity moves to lib().keep, free
            for (int c = 0; c < 3; ++c) std::swap(W[1 + c], L.keep[c]);
        inverse_velocity(F, W);
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

double read_courant_max() {
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

void do_rhs_velocity(bool use_kept) {
    Lib& L = lib();
    auto& F = L.F;
    auto& W = L.W;
    const Geo& g = L.g;
    {
        double* out[6] = {W[1], W[2], W[3], W[4], W[5], W[6]};
        Timed t(TC_POINTWISE);
        launch_products(g, W[1], W[2], W[3], out, L.dmax, L.dscal, L.st);
    }
    for (int c = 1; c <= 6; c += 3) {
        double** b[3] = {&W[c], &W[c + 1], &W[c + 2]};
        fft_forward_batch(b, 3);
    }
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
        // (or the shifted velocity rhs_scalars kept, swapped in)
        if (use_kept) {
            std::swap(W[4], L.keep[0]);
            std::swap(W[5], L.keep[1]);
        } else {
            fft_inverse_from(F[1], W[4], +1);
            fft_inverse_from(F[2], W[5], +1);
        }
        // (the last accumulation into each component applies the mask of
        // rhs_velocity.f90:400-404)
        fft_forward_product_acc(W[4], W[4], W[6], W[1], 1, false);
        fft_forward_product_acc(W[4], W[5], W[6], W[1], 2, false, W[2], 1, false);
        fft_forward_product_acc(W[5], W[5], W[6], W[2], 2, false);
        if (use_kept)
            std::swap(W[6], L.keep[2]);
        else
            fft_inverse_from(F[3], W[6], +1);
        fft_forward_product_acc(W[4], W[6], W[4], W[1], 3, true, W[3], 1, false);
        fft_forward_product_acc(W[5], W[6], W[5], W[2], 3, true, W[3], 2, false);
        fft_forward_product_acc(W[6], W[6], W[6], W[3], 3, true);
    }
    if (L.les_model == 1 || L.les_model == 11) les_velocity();  // rhs_velocity.f90:414-416
}

// band energy into dscal[1] and the forcing added to the RHS (no host sync)
void force_enqueue() {
    Lib& L = lib();
    if (L.kfmax < 1) throw std::runtime_error("force_velocity needs kfmax >= 1");
    if (L.kfmax > L.kmax) throw std::runtime_error("force_velocity: kfmax exceeds kmax");
    Timed t(TC_REDUCE);
    launch_shell_sums(L.g, L.F[1], L.F[2], L.F[3], 1, L.kfmax, L.dshell, L.st);
    const double* f[3] = {L.F[1], L.F[2], L.F[3]};
    double* r[3] = {L.W[1], L.W[2], L.W[3]};
    launch_force_apply(L.g, f, r, L.dshell, L.kfmax, L.famp, L.dscal + 1, L.st);
}

void do_force(double* energy) {
    Lib& L = lib();
    force_enqueue();
    if (energy) {
        check(hipMemcpyAsync(L.hpin + 1, L.dscal + 1, sizeof(double), hipMemcpyDeviceToHost, L.st),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(L.st), "hipStreamSynchronize");
        *energy = L.hpin[1];
    }
}

void do_advance(int which, bool first, double dt, const double* dtp) {
    Lib& L = lib();
    int lo = 1, hi = 3;
    if (which == 2) {
        lo = 4;
        hi = 3 + L.ns + L.nles;
    } else if (which != 1) {
        throw std::runtime_error("advance: which must be 1 (velocity) or 2 (scalars)");
    }
    if (which == 1) rhs_old_format(false);
    Timed t(TC_POINTWISE);
    for (int c = lo; c <= hi; ++c)
        launch_advance(L.g, L.F[c], L.W[c], L.R[c], dt, dtp, first, L.st);
}

void do_pressure() {
    Lib& L = lib();
    Timed t(TC_POINTWISE);
    launch_pressure(L.g, L.F[1], L.F[2], L.F[3], L.st);
}

}  // namespace hit3d
