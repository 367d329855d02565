// One driver step (main.f90:124-167) on the device, launched directly or
// replayed from a captured HIP graph.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace hit3d {

// The device work of one driver step (main.f90:124-167), enqueued without any
// host synchronisation; the step size is read from dscal[12].
void step_body(bool fused, bool mixed, bool scal, bool first_v, bool first_s) {
    Lib& L = lib();
    const double* dtp = L.dscal + 12;
    if (fused) {
        // rhs_scalars + rhs_velocity + force + update + pressure in one pipeline
        fused_step(dtp, first_v, mixed);
        return;
    }
    if (L.fuse && scal && L.keep.size() == 3 && fused_ps_scalars_supported()) {
        // phase shift: scalar RHS, scalar update, velocity RHS, forcing,
        // update and projection in the fused passes of step_ps.hip
        fused_ps_scalars(first_s, dtp);
        double* sh[3] = {L.keep[0], L.keep[1], L.keep[2]};
        fused_ps_velocity(sh, first_v, dtp);
        return;
    }
    // with the phase shift and scalars, the shifted velocity rhs_scalars
    // transforms is kept for rhs_velocity (the fields do not change between)
    const bool keep = scal && L.dealias == 1 && L.keep.size() == 3;
    do_rhs_scalars(scal, keep);
    if (scal) do_advance(2, first_s, 0.0, dtp);
    if (L.fuse && fused_ps_supported() && L.keep.size() == 3) {
        // the velocity RHS, forcing, update and projection in the fused
        // passes of step_ps.hip; they need the shifted physical velocity
        // before the fields change
        if (!keep)
            for (int c = 0; c < 3; ++c) fft_inverse_from(L.F[1 + c], L.keep[c], +1);
        double* sh[3] = {L.keep[0], L.keep[1], L.keep[2]};
        fused_ps_velocity(sh, first_v, dtp);
        return;
    }
    do_rhs_velocity(keep);
    if (L.flow_type == 1) force_enqueue();
    {  // main.f90:156-164 and pressure.f90:13-47 in one pass
        Timed t(TC_POINTWISE);
        double* f[3] = {L.F[1], L.F[2], L.F[3]};
        const double* r[3] = {L.W[1], L.W[2], L.W[3]};
        double* ro[3] = {L.R[1], L.R[2], L.R[3]};
        launch_advance_project(L.g, f, r, ro, 0.0, dtp, first_v, L.st);
    }
}

std::vector<void*> slot_state() {
    Lib& L = lib();
    std::vector<void*> s;
    for (auto* p : L.F) s.push_back(p);
    for (auto* p : L.R) s.push_back(p);
    for (auto* p : L.W) s.push_back(p);
    for (auto* p : L.W32) s.push_back(p);
    for (auto* p : L.mlp.buf) s.push_back(p);  // closure scratch, transformed in place
    for (auto* p : L.spares_b) s.push_back(p);
    for (auto* p : L.keep) s.push_back(p);
    s.push_back(L.spare);
    return s;
}

void set_slot_state(const std::vector<void*>& s) {
    Lib& L = lib();
    size_t i = 0;
    for (auto*& p : L.F) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.R) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.W) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.W32) p = static_cast<float*>(s[i++]);
    for (auto*& p : L.mlp.buf) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.spares_b) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.keep) p = static_cast<double*>(s[i++]);
    L.spare = static_cast<double*>(s[i++]);
}

// Puts the work-slot, closure-scratch and spare pointers back to the state s
// (from slot_state); fields and rhs_old must not have moved.
void restore_work_slots(const std::vector<void*>& s) {
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
    for (auto*& p : L.spares_b) p = static_cast<double*>(s[i++]);
    for (auto*& p : L.keep) p = static_cast<double*>(s[i++]);
    L.spare = static_cast<double*>(s[i++]);
}

// step_body from a captured graph: replayed when one was captured from the
// same slot pointers and flags, captured when this combination was already
// run once eagerly (which also makes every lazy allocation happen outside a
// capture), run eagerly otherwise.
void step_graph(bool fused, bool mixed, bool scal, bool first_v, bool first_s) {
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

}  // namespace hit3d
