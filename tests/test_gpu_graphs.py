"""GPU checks of hit3d_step replayed from captured HIP graphs: every step of
a run with graphs must equal, bit for bit, the same step launched directly
(the kernels are deterministic), for the fused 2/3-rule step (FP64 and
mixed), the unfused phase-shift step with scalars, the unfused 2/3 step with
scalars and Smagorinsky LES.  Statistics, field reads and a changing dt run
between the replays, and the captured steps are replayed several times.
"""
import numpy as np
import pytest

from hit3d_amd import FIELDS, RHS_OLD
from hit3d_amd.driver import Params, Run

pytestmark = pytest.mark.gpu

N = 32
STEPS = 8


def note(msg):
    print(msg, flush=True)


def trajectory(graphs, precision=0, **kw):
    p = Params(nx=N, ny=N, nz_all=N, isp_type=2, ir_exp=4, peak_wavenum=3.0, **kw)
    r = Run(p, precision=precision)
    try:
        r.h.set_graphs(graphs)
        r.begin_new()
        nf = 3 + p.n_scalars
        out = []
        for _ in range(STEPS):
            r.step()
            e, ens = r.h.stats()  # direct work between replays
            out.append(dict(courant=r.courant, energy=r.energy, dt=r.dt, e=e, ens=ens,
                            f=np.stack([r.h.get(FIELDS, c) for c in range(1, nf + 1)]),
                            ro=np.stack([r.h.get(RHS_OLD, c) for c in range(1, nf + 1)])))
        captured, replays = r.h.graph_stats()
    finally:
        r.h.close()
    return out, captured, replays


CASES = {
    "fused_fp64": dict(kw=dict(flow_type=1, re=400.0, dt=0.01, variable_dt=True, kfmax=2,
                               famp=0.5, dealias=0)),
    "fused_mixed": dict(precision=1, kw=dict(flow_type=1, re=400.0, dt=0.01, variable_dt=True,
                                             kfmax=2, famp=0.5, dealias=0)),
    "phase_shift_scalars": dict(kw=dict(flow_type=1, re=400.0, dt=0.005, variable_dt=True,
                                        kfmax=2, famp=0.5, dealias=1, n_scalars=2,
                                        scalar_type=[0, 0], sc=[0.7, 1.0], ir_exp_sc=[0, 0],
                                        peak_wavenum_sc=[1.0, 1.0], reac_sc=[0.0, 0.0],
                                        tscalar=-1.0)),
    "two_thirds_scalars": dict(kw=dict(flow_type=1, re=400.0, dt=0.005, variable_dt=True,
                                       kfmax=2, famp=0.5, dealias=0, n_scalars=1,
                                       scalar_type=[0], sc=[1.0], ir_exp_sc=[0],
                                       peak_wavenum_sc=[1.0], reac_sc=[0.0], tscalar=-1.0)),
    "smagorinsky": dict(kw=dict(flow_type=1, re=800.0, dt=0.005, variable_dt=True, kfmax=2,
                                famp=0.5, dealias=0, les_model=1)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replay_equals_direct_launches(case):
    c = CASES[case]
    note(f"[graphs] {case}: direct launches")
    ref, cap0, rep0 = trajectory(False, c.get("precision", 0), **c["kw"])
    assert (cap0, rep0) == (0, 0)
    note(f"[graphs] {case}: captured graphs")
    got, cap, rep = trajectory(True, c.get("precision", 0), **c["kw"])
    note(f"[graphs] {case}: {cap} graph(s), {rep} of {STEPS} steps replayed")
    assert cap >= 1 and rep >= 3
    for i, (a, b) in enumerate(zip(ref, got)):
        assert a["dt"] == b["dt"], (case, i)
        assert a["courant"] == b["courant"], (case, i)
        assert a["energy"] == b["energy"], (case, i)
        assert np.array_equal(a["e"], b["e"]), (case, i)
        assert np.array_equal(a["f"], b["f"]), (case, i, np.abs(a["f"] - b["f"]).max())
        assert np.array_equal(a["ro"], b["ro"]), (case, i)
    # the run moved: dt changed and the fields evolved
    assert ref[0]["dt"] != ref[-1]["dt"] or not c["kw"].get("variable_dt")
    assert not np.array_equal(ref[0]["f"], ref[-1]["f"])
