"""Benchmark of the HIT3DP time step on one MI355X.

Workload: 512^3 forced turbulence, RE=800, 2/3-rule dealiasing, band forcing
KFMAX=2 FAMP=0.5, variable DT, from the seeded k^4 exp(-2(k/3)^2) field
(RN1=67748937).  A step is one pass of the driver loop body (main.f90:124-176):
RHS, forcing, AB2 update, projection and the dt controller, with all state
resident in HBM.  Both precision forms run from the same initial field: FP64
throughout, and mixed (FP32 transforms and nonlinear term, FP64 state,
update and reductions); the E(k) difference between them after the timed
steps is reported.  The nonlinear term is formed (--form) as the divergence
of the trace-free product tensor by default: five forward transforms instead
of the reference's six, and the same fields to rounding, since the two
differ by a gradient the projection removes; the reference's own form and
the rotational form are timed beside it.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
FORM_NAMES = {0: "divergence -i k_j (u_i u_j)^ (the reference's)",
              1: "rotational (u x omega)^",
              2: "divergence of the trace-free tensor u_i u_j - delta_ij |u|^2/3 (differs "
                 "from the reference's by a gradient the projection removes)"}
PS_FIVE_PRODUCTS = ("divergence of u_i u_j - delta_ij w w in both phase-shift halves, five "
                    "products (differs from the reference's by a gradient the projection "
                    "removes)")
METRIC = "s/step at 512^3 FP64 vs mixed (E(k) err stated) + % of HBM roofline"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def params(n):
    from hit3d_amd.driver import Params
    return Params(nx=n, ny=n, nz_all=n, flow_type=1, re=800.0, dt=0.001, variable_dt=True,
                  isp_type=2, ir_exp=4, peak_wavenum=3.0, kfmax=2, famp=0.5, dealias=0)


def slot_bytes(n, word=8):
    return (n + 2) * n * n * word


# the other configurations of the measurement plan, run with --workload
WORKLOADS = {
    "c1": dict(desc="32^3 decaying, RE=100, fixed DT=0.01, 2/3 rule", n=32, n_fft=9, ns=0,
               kw=dict(flow_type=0, re=100.0, dt=0.01, variable_dt=False, dealias=0)),
    "c2": dict(desc="256^3 forced, RE=400, phase shift, KFMAX=2, FAMP=0.5, 2 scalars "
                    "(Sc 0.7, 1.0, mean gradient)", n=256, n_fft=37, ns=2,
               kw=dict(flow_type=1, re=400.0, dt=0.001, variable_dt=True, kfmax=2, famp=0.5,
                       dealias=1, n_scalars=2, scalar_type=[0, 0], sc=[0.7, 1.0],
                       ir_exp_sc=[0, 0], peak_wavenum_sc=[1.0, 1.0], reac_sc=[0.0, 0.0],
                       tscalar=-1.0)),
    "c5": dict(desc="1024^3 forced, RE=1600, phase shift, 3 scalars (Sc 0.5, 1, 2, mean "
                    "gradient), 1,048,576 particles (CINT)", n=1024, n_fft=45, ns=3,
               kw=dict(flow_type=1, re=1600.0, dt=5e-4, variable_dt=True, kfmax=2, famp=0.5,
                       dealias=1, n_scalars=3, scalar_type=[0, 0, 0], sc=[0.5, 1.0, 2.0],
                       ir_exp_sc=[0, 0, 0], peak_wavenum_sc=[1.0] * 3, reac_sc=[0.0] * 3,
                       tscalar=-1.0, nptot=1048576, particles_tracking_scheme=1)),
}


# configuration C4: 64^3 LES
C4_DESC = ("64^3 forced LES, RE=800, 2/3 rule, KFMAX=2, FAMP=0.5: Smagorinsky (C_s 0.18) and the "
           "learned closure (les_model 11, random-init weights of the 10-64-64-6 MLP); closure "
           "training step at batch 65,536 (Adam, lr 1e-3)")
# multiply-adds per sample of one training step: forward 16x64 + 64x64 + 64x16,
# the same again for the weight gradients, 64x64 + 64x16 for the input gradients
MLP_TRAIN_FLOP_PER_SAMPLE = 2 * (2 * 6144 + 5120)
PEAK_BF16_TFLOPS = 2500.0


def run_c4(steps, warmup, train_steps=50, batch=65536):
    """C4 timing: LES steps with each closure, then closure training steps on
    synthetic standardised samples.  Prints its own JSON line."""
    from hit3d_amd.driver import Params, Run
    from hit3d_amd.sgs import init_params
    n = 64
    out = {"workload": "c4", "config": C4_DESC, "steps": steps, "warmup": warmup}
    for model in ("smagorinsky", "mlp"):
        p = Params(nx=n, ny=n, nz_all=n, flow_type=1, re=800.0, dt=0.002, variable_dt=True,
                   isp_type=2, ir_exp=4, peak_wavenum=3.0, kfmax=2, famp=0.5, dealias=0,
                   les_model=1)
        r = Run(p)
        r.begin_new()
        if model == "mlp":
            r.h.mlp_set_params(init_params(0))
            # inputs as they come (unit scale), outputs scaled to a small stress
            r.h.mlp_set_norm(np.zeros(10), np.ones(10), np.zeros(6), np.full(6, 1e-3))
            r.h.set_les(11, 0.18)
        for _ in range(warmup):
            r.step()
        r.h.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.step()
        r.h.synchronize()
        t = (time.perf_counter() - t0) / steps
        out[f"{model}_s_per_step"] = t
        if model == "mlp":
            out["mlp_fallback_steps"] = r.h.mlp_fallbacks()
        r.h.close()
        log(f"[bench] c4 {model} {t * 1e3:.3f} ms/step")
    from hit3d_amd import Hit3D
    rng = np.random.default_rng(12345)
    X = rng.standard_normal((4 * batch, 10)).astype(np.float32)
    T = (0.1 * rng.standard_normal((4 * batch, 6))).astype(np.float32)
    h = Hit3D(n)
    try:
        h.mlp_set_params(init_params(0))
        h.mlp_set_norm(np.zeros(10), np.ones(10), np.zeros(6), np.ones(6))
        h.mlp_set_data(X, T)
        h.mlp_train(5, batch, 1e-3)
        h.synchronize()
        t0 = time.perf_counter()
        losses = h.mlp_train(train_steps, batch, 1e-3)
        h.synchronize()
        t = (time.perf_counter() - t0) / train_steps
    finally:
        h.close()
    flop = MLP_TRAIN_FLOP_PER_SAMPLE * batch
    out["train"] = {"batch": batch, "steps": train_steps, "s_per_step": t,
                    "samples_per_s": batch / t, "flop_per_step": flop,
                    "achieved_tflops": flop / t / 1e12, "peak_bf16_tflops": PEAK_BF16_TFLOPS,
                    "frac": flop / t / 1e12 / PEAK_BF16_TFLOPS,
                    "loss_first_last": [float(losses[0]), float(losses[-1])]}
    print(json.dumps(out), flush=True)


def run_fft_micro(sizes=(256, 512, 1024), reps=20, warmup=3):
    """Transform microbenchmark of the measurement plan: a uniform(-1, 1) FP64
    field from a fixed seed (12345) in one wrk slot of a library holding all
    its slots, forward then inverse xFFT3d, median of `reps` after `warmup`.
    A transform moves 3 passes x (read + write) of one slot.  Prints its own
    JSON line."""
    from hit3d_amd import WRK, Hit3D
    out = {"workload": "fft", "config": "xFFT3d(+1) and xFFT3d(-1) on one FP64 slot, "
                                        "uniform(-1, 1) field, seed 12345", "reps": reps,
           "warmup": warmup, "sizes": {}}
    for n in sizes:
        h = Hit3D(n)
        try:
            a = np.zeros((n, n, n + 2))
            a[:, :, :n] = np.random.default_rng(12345).uniform(-1.0, 1.0, (n, n, n))
            h.put(WRK, 1, a)
            del a
            fwd, inv = [], []
            for i in range(warmup + reps):
                h.synchronize()
                t0 = time.perf_counter()
                h.xFFT3d(1, 1)
                h.synchronize()
                t1 = time.perf_counter()
                h.xFFT3d(-1, 1)
                h.synchronize()
                t2 = time.perf_counter()
                if i >= warmup:
                    fwd.append(t1 - t0)
                    inv.append(t2 - t1)
        finally:
            h.close()
        byt = 6 * slot_bytes(n)
        tf, ti = float(np.median(fwd)), float(np.median(inv))
        out["sizes"][str(n)] = {"forward_ms": tf * 1e3, "inverse_ms": ti * 1e3,
                                "bytes_per_transform": byt,
                                "forward_GBs": byt / tf / 1e9, "inverse_GBs": byt / ti / 1e9,
                                "forward_frac_of_8TBs": byt / tf / (PEAK_HBM_GBS * 1e9),
                                "inverse_frac_of_8TBs": byt / ti / (PEAK_HBM_GBS * 1e9)}
        log(f"[bench] fft n={n}: forward {tf * 1e3:.3f} ms, inverse {ti * 1e3:.3f} ms")
    print(json.dumps(out), flush=True)


def run_workload(name, steps, warmup, form=0):
    """One of the other configurations: FP64 driver steps, with particles
    moved every step when the configuration has them; form 2 runs the
    phase-shift configurations with five products instead of six (the same
    fields to rounding).  Prints its own JSON line."""
    from hit3d_amd.driver import Params, Run
    w = WORKLOADS[name]
    n = w["n"]
    kw = dict(w["kw"])
    nptot = kw.pop("nptot", 0)
    scheme = kw.pop("particles_tracking_scheme", 0)
    p = Params(nx=n, ny=n, nz_all=n, isp_type=2, ir_exp=4, peak_wavenum=3.0, **kw)
    r = Run(p)
    if form and p.dealias == 1:
        r.h.set_nonlinear(form)
    t0 = time.time()
    r.begin_new()
    if nptot:
        rng = np.random.default_rng(12345)
        r.ipart = np.arange(1, nptot + 1, dtype=np.int32)
        r.h.particles_set(rng.uniform(1.0, n + 1.0, (nptot, 3)))
        p.nptot, p.particles_tracking_scheme = nptot, scheme
    r.h.synchronize()
    log(f"[bench] {name} init {time.time() - t0:.1f}s")
    for _ in range(warmup):
        r.step()
    r.h.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        r.step()
        log(f"[bench] {name} step {i + 1}/{steps}")
    r.h.synchronize()
    t1 = time.perf_counter()
    graphs, replays = r.h.graph_stats()
    # per-kernel-class times: as many further steps with HIP events around
    # the launches (direct launches, no graph replay)
    r.h.timing_enable(True)
    for i in range(steps):
        r.step()
    r.h.synchronize()
    timers = r.h.timing_read()
    r.h.close()
    t = (t1 - t0) / steps
    s = slot_bytes(n)
    b_alg = 6 * w["n_fft"] * s + 4 * (3 + w["ns"]) * s
    out = {"workload": name, "config": w["desc"], "s_per_step": t, "ns_per_point_step": t / n ** 3 * 1e9,
           "nonlinear_form": (PS_FIVE_PRODUCTS if (form == 2 and p.dealias == 1) else FORM_NAMES[0]),
           "plan_alg_bytes": b_alg,
           "plan_bytes_over_step_time_vs_peak": b_alg / t / (PEAK_HBM_GBS * 1e9),
           "particles": nptot, "steps": steps, "warmup": warmup,
           "graph_replays": replays,
           "kernel_ms_per_step": {k: v[0] / steps for k, v in timers.items() if v[1]}}
    print(json.dumps(out), flush=True)


def run_form(n, precision, steps, warmup, form=0):
    from hit3d_amd.driver import Run
    p = params(n)
    r = Run(p, precision=precision)
    if form:
        r.h.set_nonlinear(form)
    t0 = time.time()
    r.begin_new()
    r.h.synchronize()
    log(f"[bench] n={n} precision={precision} init {time.time() - t0:.1f}s")
    for _ in range(warmup):
        r.step()
    r.h.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:  # torch is plumbing only; the library's own sync is authoritative
        pass
    # the timed steps as the library runs them (captured graphs); the
    # per-kernel HIP-event timers need direct launches, so they time the same
    # number of further steps of their own
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    r.h.synchronize()
    t1 = time.perf_counter()
    st = r.stats()
    r.h.timing_enable(True)
    for _ in range(steps):
        r.step()
    r.h.synchronize()
    timers = r.h.timing_read()
    r.h.timing_enable(False)
    log(f"[bench] n={n} precision={precision} {1e3 * (t1 - t0) / steps:.3f} ms/step "
        f"courant={r.courant:.4f} dt={r.dt:.3e} energy={st.energy:.6e}")
    r.h.close()
    return (t1 - t0) / steps, timers, st, r


# algorithmic HBM bytes per launch of each timed kernel class (FP64 slot S):
# every axis pass reads and writes one slot's valid entries; the fused step
# kernels are charged the slots they must read and write.
def kernel_bytes(n, cls, precision, form=0):
    """Bytes each kernel must move: the unfused axis passes read and write one
    slot's valid entries; the fused step kernels move only the 2/3-rule cube
    (kmax+1 kx columns, 2kmax+1 ky/kz rows) they read or produce.  form: the
    nonlinear term's form (0 divergence, six product buffers; 1 rotational,
    six inverse and three forward buffers; 2 deviatoric, five products)."""
    s64 = slot_bytes(n)
    k = n // 3
    nk, nky = k + 1, 2 * k + 1
    w = 16 if precision == 0 else 8          # work complex word
    cube = nk * nky * nky * 16               # one FP64 field inside the cube
    cube_ro = nk * nky * nky * w             # rhs_old inside the cube (FP32 in the mixed form)
    zy = n * nky * nk * w                     # one work buffer with kept ky rows, all z
    full = n * n * nk * w                     # one work buffer, all y and z
    ni = 6 if form == 1 else 3                # buffers through the inverse passes
    nf = {0: 6, 1: 3, 2: 5}[form]             # buffers through the forward passes
    table = {
        "fft_x": 2 * s64, "fft_y": 2 * s64, "fft_z": 2 * s64,
        "step_inv_z": 3 * cube + ni * zy,
        "step_inv_y": ni * zy + ni * full,
        "step_x": ni * full + nf * full,
        "step_fwd_y": nf * full + nf * zy,
        "step_fwd_z": nf * zy + 6 * cube + 6 * cube_ro,   # products in; fields, rhs_old in/out
    }
    return table.get(cls)


def accuracy_run(n, steps, every, form=0):
    """Untimed: FP64 and mixed runs from the same field, one after the other
    (the library holds one grid per process), for `steps` driver steps;
    E(k), dissipation and velocity-derivative skewness sampled every `every`
    steps.  Returns the errors of the mixed form's time averages against
    FP64's, and how the instantaneous E(k) difference grows (the two
    trajectories separate as any two chaotic trajectories do; the time
    averages are what the mixed form must reproduce)."""
    from hit3d_amd.driver import Run
    samples = {0: [], 1: []}
    for prec in (0, 1):
        r = Run(params(n), precision=prec)
        try:
            if form:
                r.h.set_nonlinear(form)
            r.begin_new()
            for i in range(steps):
                r.step()
                if (i + 1) % 200 == 0:
                    log(f"[bench] accuracy run ({'mixed' if prec else 'FP64'}): step {i + 1}/{steps}")
                if (i + 1) % every == 0:
                    st = r.stats()
                    samples[prec].append((st.time, st.e_spec, st.eps_v, r.h.skewness()))
        finally:
            r.h.close()
    inst = []
    for j, (a64, a32) in enumerate(zip(samples[0], samples[1])):
        e64, e32 = a64[1], a32[1]
        sig = e64 > 1e-12 * e64.max()
        inst.append(((j + 1) * every, float(np.max(np.abs(e32[sig] - e64[sig]) / e64[sig]))))
    avg = {prec: (np.mean([x[1] for x in v], axis=0), float(np.mean([x[2] for x in v])),
                  float(np.mean([x[3] for x in v]))) for prec, v in samples.items()}
    e64, e32 = avg[0][0], avg[1][0]
    sig = e64 > 1e-12 * e64.max()
    marks = sorted({1, 10, 100, 1000, steps} & {k for k, _ in inst} | {inst[-1][0]})
    return {"steps": steps, "samples": len(samples[0]), "every": every,
            "sim_time_fp64": samples[0][-1][0], "sim_time_mixed": samples[1][-1][0],
            "ek_time_avg_max_rel_err_vs_fp64": float(np.max(np.abs(e32[sig] - e64[sig]) / e64[sig])),
            "eps_time_avg_rel_err_vs_fp64": abs(avg[1][1] - avg[0][1]) / avg[0][1],
            "skewness_time_avg_fp64": avg[0][2], "skewness_time_avg_mixed": avg[1][2],
            "skewness_time_avg_rel_err_vs_fp64": abs(avg[1][2] - avg[0][2]) / abs(avg[0][2]),
            "eps_rel_err_vs_fp64": abs(samples[1][-1][2] - samples[0][-1][2]) / samples[0][-1][2],
            "skewness_fp64": samples[0][-1][3], "skewness_mixed": samples[1][-1][3],
            "skewness_rel_err_vs_fp64": abs(samples[1][-1][3] - samples[0][-1][3])
            / abs(samples[0][-1][3]),
            "ek_instantaneous_max_rel_err_by_step": {str(k): v for k, v in inst if k in marks}}


def copy_bandwidth(n, reps=20):
    """Two copies of one slot (wrk(2) = wrk(1), read + write of every byte):
    the runtime's copy kernel (hipMemcpyAsync device to device), which also
    scales the FETCH_SIZE / WRITE_SIZE counters, and the library's own
    streaming copy (nontemporal dwordx4 loads and stores, several launch
    shapes, best kept): the achievable rate of a pass that reads and writes
    each byte once.  GB/s of bytes read + written."""
    import hit3d_amd
    from hit3d_amd import measure
    h = hit3d_amd.Hit3D(n)
    try:
        h.copy(hit3d_amd.WRK, 2, hit3d_amd.WRK, 1)
        h.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            h.copy(hit3d_amd.WRK, 2, hit3d_amd.WRK, 1)
        h.synchronize()
        t = (time.perf_counter() - t0) / reps
        best = None
        for blocks in (512, 1024, 2048, 4096, 8192, 16384, 65536):
            ms, byt = measure.stream_copy(h, reps, blocks)
            gbs = byt / (ms * 1e-3) / 1e9
            if best is None or gbs > best[0]:
                best = (gbs, blocks, ms, byt)
    finally:
        h.close()
    p = ((n // 2 + 1 + 15) // 16) * 16
    byt = n * n * p * 16
    log(f"[bench] runtime copy: {reps + 1} x {byt} bytes, {t * 1e3:.3f} ms each; streaming "
        f"copy {best[0]:.0f} GB/s at {best[1]} blocks")
    return {"runtime_copy_GBs": 2 * byt / t / 1e9, "stream_copy_GBs": best[0],
            "stream_copy_blocks": best[1], "stream_copy_ms": best[2],
            "stream_copy_bytes": best[3]}


def measured_traffic(n, cls, form=2):
    """HBM bytes per launch of a kernel class from the committed PMC summary
    (profiles/pmc_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by
    the same counters on the streaming copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    for run in d.get("runs", []):
        if run.get("n") == n and run.get("form") == form and cls in run.get("classes", {}):
            c = run["classes"][cls]
            return c["fetch_bytes"] + c["write_bytes"]
    return None


def host_cpu():
    """CPU model and logical core count of the host the baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(n):
    """The oracle (numpy/C restatement of the reference step: numpy's
    pocketfft and elementwise numpy, one thread) timed for one driver step at
    the benchmark's own grid, from the same seeded field: no scaling."""
    from oracle import hit3d_oracle as o
    import hit3d_amd
    h = hit3d_amd.Hit3D(n)
    h.init_velocity(67748937.0, 2, 4, 3.0)
    h.dealias_all()
    f0 = np.stack([h.get(hit3d_amd.FIELDS, c) for c in (1, 2, 3)])
    h.close()
    cfg = o.Config(n=n, nu=1 / 800.0, dt=0.001, dealias=0, flow_type=1, kfmax=2, famp=0.5,
                   variable_dt=True)
    st = o.State(fields=f0, rhs_old=np.zeros_like(f0), dt=cfg.dt)
    log(f"[bench] cpu baseline: one oracle step at {n}^3")
    t0 = time.perf_counter()
    st = o.step(cfg, st)
    dt_n = time.perf_counter() - t0
    model, ncpu = host_cpu()
    return {"value": dt_n, "unit": "s/step", "cores": 1, "kind": "port",
            "sample": f"one driver step of the oracle (numpy pocketfft + elementwise numpy, "
                      f"single thread) at {n}^3 from the seeded field, measured directly; "
                      f"host {model}, {ncpu} logical cores, 1 used"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", type=int, default=2, choices=[0, 1, 2],
                    help="nonlinear term of the timed step: 0 divergence (the reference's "
                         "products), 1 rotational, 2 divergence of the trace-free tensor "
                         "(default; five forward transforms, the same fields to rounding)")
    ap.add_argument("--no-rotational", action="store_true",
                    help="skip the timing of the rotational and deviatoric forms")
    ap.add_argument("--workload", default="c3", choices=["c3", "c4", "fft"] + sorted(WORKLOADS),
                    help="c3 (default) prints the benchmark line; c1, c2, c4, c5 time the other "
                         "configurations; fft is the transform microbenchmark")
    ap.add_argument("--accuracy-steps", type=int, default=100,
                    help="untimed FP64 vs mixed run for time-averaged E(k) and skewness errors "
                         "(0 skips it)")
    ap.add_argument("--accuracy-only", action="store_true",
                    help="run only the FP64 vs mixed accuracy comparison (--accuracy-steps "
                         "steps, sampled every 10) and print it as one JSON line")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs the whole box on one GPU; --gpus {a.gpus} is not "
              "supported (multi-GPU sharding is design only)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "c4":
        run_c4(a.steps, a.warmup)
        return
    if a.workload == "fft":
        run_fft_micro()
        return
    if a.workload != "c3":
        run_workload(a.workload, a.steps, a.warmup, a.form)
        return
    if a.accuracy_only:
        t0 = time.perf_counter()
        acc = accuracy_run(a.n, a.accuracy_steps, 10, a.form)
        acc.update({"n": a.n, "nonlinear_form": FORM_NAMES[a.form],
                    "wall_s": time.perf_counter() - t0})
        print(json.dumps(acc), flush=True)
        return
    n = a.n
    form = a.form
    t64, tim64, st64, _ = run_form(n, 0, a.steps, a.warmup, form=form)
    mixed = None
    if not a.no_mixed:
        try:
            t32, tim32, st32, _ = run_form(n, 1, a.steps, a.warmup, form=form)
            e64, e32 = st64.e_spec, st32.e_spec
            sig = e64 > 1e-12 * e64.max()
            mixed = {"s_per_step": t32, "speedup_vs_fp64": t64 / t32,
                     "ek_max_rel_err_vs_fp64": float(np.max(np.abs(e32[sig] - e64[sig]) / e64[sig])),
                     "energy_rel_err_vs_fp64": abs(st32.energy - st64.energy) / st64.energy,
                     "eps_rel_err_vs_fp64": abs(st32.eps_v - st64.eps_v) / st64.eps_v,
                     "after_steps": a.steps + a.warmup,
                     "kernel_ms_per_step": {k: v[0] / a.steps for k, v in tim32.items() if v[1]}}
            if a.accuracy_steps > 0:
                mixed["accuracy"] = accuracy_run(n, a.accuracy_steps, 10, form)
        except Exception as e:  # reported, never silently replaced
            mixed = {"error": str(e)}
    # the other forms of the nonlinear term, FP64: rotational (u x omega)
    # and the trace-free tensor's divergence; same fields to rounding
    forms = {}
    if not a.no_rotational:
        for name, fo in (("divergence_fp64", 0), ("rotational_fp64", 1), ("deviatoric_fp64", 2)):
            if fo == form:
                continue
            tr, timr, str_, _ = run_form(n, 0, a.steps, a.warmup, form=fo)
            forms[name] = {"s_per_step": tr, "vs_headline_form": tr / t64,
                           "energy_rel_diff_vs_headline_form":
                               abs(str_.energy - st64.energy) / st64.energy,
                           "kernel_ms_per_step": {k: v[0] / a.steps for k, v in timr.items()
                                                  if v[1]}}
    copy = copy_bandwidth(n)
    copy_gbs = copy["stream_copy_GBs"]
    # dominant kernel class of the FP64 step
    best, best_ms = None, -1.0
    for cls, (ms, cnt) in tim64.items():
        if cnt > 0 and ms > best_ms and kernel_bytes(n, cls, 0, form) is not None:
            best, best_ms = cls, ms
    roof = None
    if best is not None:
        ms, cnt = tim64[best]
        avg_s = ms / cnt / 1e3
        byt = kernel_bytes(n, best, 0, form)
        ach = byt / avg_s / 1e9
        roof = {"bound": "hbm", "kernel_class": best, "achieved": ach, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": measured_traffic(n, best, form),
                "alg_bytes_per_launch": byt, "avg_launch_ms": avg_s * 1e3,
                "launches_per_step": cnt / a.steps, "stream_copy_GBs": copy_gbs,
                "frac_of_stream_copy": ach / copy_gbs}
    # every kernel class of the FP64 step against the same peak (the dominant
    # one above is the x pass or the z/update kernel, which take about the
    # same time at 512^3 and trade places from one GPU to the next)
    by_kernel = {}
    for cls, (ms, cnt) in tim64.items():
        byt = kernel_bytes(n, cls, 0, form)
        if cnt > 0 and byt is not None:
            ach = byt / (ms / cnt / 1e3) / 1e9
            by_kernel[cls] = {"achieved_GBs": ach, "frac": ach / PEAK_HBM_GBS,
                              "frac_of_stream_copy": ach / copy_gbs,
                              "alg_bytes_per_launch": byt, "avg_launch_ms": ms / cnt,
                              "launches_per_step": cnt / a.steps}
    # the step's own algorithmic bytes: what its five kernels must move (the
    # pruned transforms), per step
    step_bytes = sum(v["alg_bytes_per_launch"] * v["launches_per_step"]
                     for v in by_kernel.values())
    s = slot_bytes(n)
    # the measurement plan's byte model of an unpruned step (9 full transforms
    # of 6 slot passes each + the state): bytes this step never moves
    model_bytes = 6 * 9 * s + 4 * 3 * s
    out = {
        "metric": METRIC,
        "value": t64,
        "unit": "s/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded RANDu field, spectrum k^4 exp(-2(k/3)^2), RN1=67748937",
        "config": {"workload": f"{n}^3 forced DNS, RE=800, 2/3-rule, KFMAX=2, FAMP=0.5, "
                               "variable DT, no scalars", "grid": n, "precision": "fp64",
                   "nonlinear_form": FORM_NAMES[form], "parallelism": "single GPU"},
        "ns_per_point_step": t64 / n ** 3 * 1e9,
        "step_alg_bytes": step_bytes,
        "step_hbm_roofline_frac": step_bytes / t64 / (PEAK_HBM_GBS * 1e9),
        "step_frac_of_stream_copy": step_bytes / t64 / (copy_gbs * 1e9),
        "model_alg_bytes_unpruned": model_bytes,
        "model_bytes_over_step_time_vs_peak": model_bytes / t64 / (PEAK_HBM_GBS * 1e9),
        "copy": copy,
        "mixed": mixed,
        "nonlinear_forms": forms,
        "kernel_ms_per_step": {k: v[0] / a.steps for k, v in tim64.items() if v[1]},
        "roofline": roof,
        "roofline_by_kernel": by_kernel,
        "cpu_baseline": None if a.no_cpu else cpu_baseline(n),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
