// Library state: the device-resident arrays of the reference's modules and
// the tables the kernels share.
#pragma once
#include <string>
#include <vector>

#include "common.hpp"
#include "particles.hpp"
#include "sgs_mlp.hpp"
#include "spectral.hpp"

namespace hit3d {

enum TimerClass {
    TC_FFT_X = 0,
    TC_FFT_Y,
    TC_FFT_Z,
    TC_POINTWISE,
    TC_REDUCE,
    TC_STEP_INV_Z,
    TC_STEP_INV_Y,
    TC_STEP_X,
    TC_STEP_FWD_Y,
    TC_STEP_FWD_Z,
    TC_STEP_UPDATE,
    TC_PARTICLES,
    TC_COUNT
};

struct HipError {
    std::string msg;
};

void check(hipError_t e, const char* what);

struct Lib {
    bool ready = false;
    int n = 0;
    Geo g{};
    int ns = 0, nles = 0, nf = 0, nw = 0;
    int dealias = 0, kmax = 0;
    double nu = 0.0, famp = 0.0, dx = 0.0;
    int flow_type = 0, kfmax = 0;
    std::vector<double> pe;
    std::vector<int64_t> stype;
    std::vector<double> reac;  // reaction rates reac_sc (m_parameters.f90:445)
    int64_t rng_iy = 0;        // the generator's saved iy (RANDu.f:22-24) across inits

    std::vector<double*> F;  // fields, index 1..nf (0 unused)
    std::vector<double*> W;  // wrk, index 0..nw
    std::vector<double*> R;  // rhs_old, index 1..nf
    // rhs_old(1:3) holds FP32 values (cpx<float> at the FP64 entry index, in
    // the first half of each slot): the mixed step's format, converted back
    // before anything else reads rhs_old
    bool ro32 = false;
    double* spare = nullptr;
    std::vector<double*> spares_b;  // two more spares for batched transforms (N <= 128)
    std::vector<double*> keep;      // shifted physical velocity kept between the two RHS (3 slots)

    cpx<double>* tw64 = nullptr;
    cpx<float>* tw32 = nullptr;
    double* ps_tab = nullptr;        // cos/sin(0.5*(kx+ky+kz)*dx), index (kx+ky+kz+n)
    double* dshell = nullptr;        // shell sums (kmax+2 doubles)
    double* dpart = nullptr;         // per-plane partial sums (2n doubles)
    double* dmax = nullptr;          // per-block maxima of |u|+|v|+|w| (max(65536, n*n) doubles)
    double* dscal = nullptr;         // [0] courant max bits, [1] band energy, [2] enstrophy,
                                     // [4..5] physical sums, [8..9] min/max keys
    double* hpin = nullptr;          // pinned host mirror of dscal
    hipStream_t st = nullptr;
    int ncu = 0;  // compute units of the device (persistent launches)

    // mixed-precision work buffers (allocated on first use): nine FP32 slots
    std::vector<float*> W32;

    bool fuse = true;  // hit3d_step uses the fused pipeline where it applies
    // form of the fused step's nonlinear term: 0 divergence (the reference's),
    // 1 rotational (u x omega), 2 divergence of the trace-free tensor
    int nlform = 0;

    // hit3d_step replayed from captured HIP graphs (graph mode).  A capture is
    // valid for the slot pointers it started from (the out-of-place passes
    // swap them) and the step's flags; it records the pointers it leaves.
    struct StepGraph {
        std::vector<void*> before, after;
        unsigned flags = 0;
        hipGraphExec_t exec = nullptr;
    };
    bool graph_mode = true;
    int64_t graph_replays = 0;  // steps run from a captured graph
    std::vector<StepGraph> graphs;
    std::vector<std::pair<std::vector<void*>, unsigned>> graph_seen;  // run once eagerly first
    int les_model = 0;       // 0 none, 1 Smagorinsky (m_les.f90:94-104)
    double c_smag = 0.18;    // m_les.f90:1148
    double* tv = nullptr;    // turb_visc, physical slot layout
    // Lagrangian particles (m_particles.f90:20-31): positions in grid units
    // and the previous interpolated velocity, (3, np) order, device-resident
    int64_t np = 0, np_cap = 0;
    double* xyzp = nullptr;
    double* uvwp = nullptr;

    // learned closure (les_model 11): parameters, Adam state, training data
    struct Mlp {
        float* p = nullptr;       // FP32 master parameters (mlp::NPARAM)
        uint16_t* wb = nullptr;   // bfloat16 copy
        float *m = nullptr, *v = nullptr, *g = nullptr;
        float* partial = nullptr; // per-tile gradient partials
        float* lossp = nullptr;
        float* red = nullptr;     // group sums of the two-level gradient reduction
        float* loss = nullptr;    // per-step losses of the last train call
        int64_t cap_tiles = 0, cap_steps = 0;
        int step = 0;
        float *X = nullptr, *T = nullptr;
        int64_t count = 0;
        MlpNorm norm{};
        bool have_params = false, have_norm = false;
        std::vector<double*> buf;  // 9 gradient + 6 stress slots
        unsigned int* bad = nullptr;
        int64_t fallbacks = 0;     // steps that used Smagorinsky instead
        double delta = 0.0;
        double mix = 0.0;          // weight of the Smagorinsky stress in the LES stress
    } mlp;

    bool timing = false;
    struct Ev {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Ev> evs;
    double tms[TC_COUNT] = {0};
    int64_t tcnt[TC_COUNT] = {0};

    std::string err;
};

Lib& lib();

// RAII event pair around launches of one class when timing is enabled.
struct Timed {
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    explicit Timed(int c);
    ~Timed();
};

// internal transforms on a slot pointer (the pointer may be swapped with
// the spare buffer, which is why it is passed by reference)
// ph = -1: the spectral result is multiplied by exp(-i theta), theta = (kx+ky+kz) dx/2
void fft_forward(double*& slot, int ph = 0);
void fft_inverse(double*& slot);

// fused FP64 / mixed step (step.hip)
// Enqueues one fused step (2/3 rule): RHS, forcing, update and projection;
// the step size is read from device memory at dtp; the Courant max and the
// band energy land in dscal[0], dscal[1].  No host synchronisation.
void fused_step(const double* dtp, bool first, bool mixed);
bool fused_step_supported(bool mixed);

}  // namespace hit3d
