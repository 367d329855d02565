"""The learned subgrid-stress closure's workflow on the GPU library: training
data from box-filtered DNS, standardisation, training, and coarse LES runs
with the closure or with Smagorinsky, measured against the filtered-DNS spectrum.

No counterpart in the reference: the closure stands in for the eddy
viscosity of m_les.f90:1142-1215 (kept as the fallback).  Every field
operation runs on the GPU through the C ABI; this module only moves
host-side arrays between the two grids (one library instance holds one grid
at a time).
"""
from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from .driver import Params, Run, my_dt
from .lib import FIELDS, WRK, Hit3D

# tau components in the closure's order and the velocity pairs they belong to
TAU_PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


@dataclass
class Snapshot:
    time: float
    ubar: np.ndarray   # [3, m, m, m] physical, box-filtered
    tau: np.ndarray    # [6, m, m, m] avg(u_i u_j) - ubar_i ubar_j


@dataclass
class Dataset:
    X: np.ndarray      # [P, 10] raw inputs
    T: np.ndarray      # [P, 6] deviatoric stress
    m: int
    delta: float
    ek_filtered: list = field(default_factory=list)  # E(k) of each filtered snapshot


def dns_params(n: int, re: float = 800.0) -> Params:
    """Forced DNS as in the 512^3 configuration: band forcing KFMAX=2,
    FAMP=0.5, 2/3 rule, variable dt, seeded k^4 exp(-2(k/3)^2) field."""
    return Params(nx=n, ny=n, nz_all=n, flow_type=1, re=re, dt=1e-3, variable_dt=True,
                  isp_type=2, ir_exp=4, peak_wavenum=3.0, kfmax=2, famp=0.5, dealias=0)


def run_dns(n: int, ratio: int, spinup: int, nsnap: int, every: int, re: float = 800.0,
            precision: int = 0) -> list:
    """Box-filtered snapshots of a forced DNS: spinup steps, then nsnap
    snapshots every `every` steps (the filter ratio sets the coarse grid)."""
    r = Run(dns_params(n, re), precision=precision)
    snaps = []
    try:
        r.begin_new()
        t0 = time.time()
        for i in range(spinup):
            r.step()
            if (i + 1) % max(1, spinup // 10) == 0:
                log(f"[sgs] dns n={n} spin-up {i + 1}/{spinup} t={r.time:.3f} "
                    f"courant={r.courant:.3f} ({time.time() - t0:.0f}s)")
        for s in range(nsnap):
            for _ in range(every):
                r.step()
            r.h.fields_to_parts()
            ubar, tau = r.h.box_filter(ratio)
            snaps.append(Snapshot(r.time, ubar, tau))
            log(f"[sgs] dns snapshot {s + 1}/{nsnap} t={r.time:.3f} energy={r.stats().energy:.4f}")
    finally:
        r.h.close()
    return snaps


def shell_spectrum_host(ubar: np.ndarray) -> np.ndarray:
    """E(k) of a physical field on the coarse grid, shells 1..m/2 (the
    reference's convention: 1/2 sum |u_k|^2 / m^6 over a sphere shell)."""
    m = ubar.shape[1]
    F = [np.fft.fftn(ubar[c]) for c in range(3)]
    k = np.fft.fftfreq(m, 1.0 / m)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    sh = np.floor(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) + 0.5).astype(int)
    e = 0.5 * sum(np.abs(f) ** 2 for f in F) / float(m) ** 6
    return np.bincount(sh.ravel(), weights=e.ravel(), minlength=m)[1:m // 2 + 1]


def _truncated(h: Hit3D, phys3: list, m: int) -> list:
    """The 2/3-rule truncation (dealias_all) of up to three physical coarse
    fields, on the GPU: transform, truncate, transform back."""
    for c, f in enumerate(phys3):
        a = np.zeros((m, m, m + 2))
        a[:, :, :m] = f
        h.put(WRK, 1, a)
        h.xFFT3d(1, 1)
        h.copy(FIELDS, c + 1, WRK, 1)
    h.dealias_all()
    out = []
    for c in range(len(phys3)):
        h.copy(WRK, 1, FIELDS, c + 1)
        h.xFFT3d(-1, 1)
        out.append(h.get(WRK, 1)[:, :, :m])
    return out


def features(snaps: list, m: int, delta: float) -> Dataset:
    """Inputs and targets on the coarse grid for the filter the LES actually
    applies: the box filter B followed by the 2/3-rule truncation T of the
    coarse grid.  Inputs: the nine gradients of T(B(u)) (spectral
    derivatives, as the LES closure computes them) and the filter width;
    targets: the deviatoric part of T(B(u_i u_j)) - T(B(u_i)) T(B(u_j))."""
    h = Hit3D(m)
    Xs, Ts, eks = [], [], []
    try:
        for sn in snaps:
            ut = _truncated(h, [sn.ubar[0], sn.ubar[1], sn.ubar[2]], m)
            bprod = [sn.tau[k] + sn.ubar[i] * sn.ubar[j] for k, (i, j) in enumerate(TAU_PAIRS)]
            tprod = _truncated(h, bprod[:3], m) + _truncated(h, bprod[3:], m)
            t = np.stack([tprod[k] - ut[i] * ut[j] for k, (i, j) in enumerate(TAU_PAIRS)]).reshape(6, -1)
            # the LES state: T(B(u)) as spectral fields
            for c in range(3):
                a = np.zeros((m, m, m + 2))
                a[:, :, :m] = ut[c]
                h.put(WRK, 1, a)
                h.xFFT3d(1, 1)
                h.copy(FIELDS, c + 1, WRK, 1)
            h.dealias_all()
            Xs.append(h.sgs_features(delta))
            tr = (t[0] + t[3] + t[5]) / 3.0
            dev = t.copy()
            dev[0] -= tr
            dev[3] -= tr
            dev[5] -= tr
            Ts.append(dev.T.astype(np.float32))
            eks.append(shell_spectrum_host(np.stack(ut)))
    finally:
        h.close()
    return Dataset(np.concatenate(Xs), np.concatenate(Ts), m, delta, eks)


def standardise(ds: Dataset):
    xm, xs = ds.X.mean(0).astype(np.float64), ds.X.std(0).astype(np.float64)
    xs = np.where(xs > 1e-12 * np.maximum(1.0, np.abs(xm)), xs, 1.0)
    ym, ys = ds.T.mean(0).astype(np.float64), ds.T.std(0).astype(np.float64)
    ys = np.where(ys > 0, ys, 1.0)
    return xm, xs, ym, ys


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights in the closure's parameter layout."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((16, 64))
    W1[:10] = rng.standard_normal((10, 64)) * math.sqrt(2.0 / 10)
    W2 = rng.standard_normal((64, 64)) * math.sqrt(2.0 / 64)
    W3 = np.zeros((64, 16))
    W3[:, :6] = rng.standard_normal((64, 6)) * math.sqrt(1.0 / 64)
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), W3.ravel(),
                           np.zeros(16)]).astype(np.float32)


@dataclass
class Trained:
    params: np.ndarray
    norm: tuple
    losses: np.ndarray
    val_loss: float
    smag_val_loss: float


def smagorinsky_tau(X: np.ndarray, delta: float, cs: float) -> np.ndarray:
    """Deviatoric Smagorinsky stress -2 (cs delta)^2 |S| S_ij from the raw
    gradient inputs (the a-priori comparison of the fallback model)."""
    g = X[:, :9].astype(np.float64).reshape(-1, 3, 3)   # g[p, i, j] = du_i/dx_j
    S = 0.5 * (g + g.transpose(0, 2, 1))
    smag = np.sqrt(2.0 * np.einsum("pij,pij->p", S, S))
    nut = (cs * delta) ** 2 * smag
    tr = (S[:, 0, 0] + S[:, 1, 1] + S[:, 2, 2]) / 3.0
    out = np.stack([S[:, i, j] for (i, j) in TAU_PAIRS], axis=1)
    out[:, 0] -= tr
    out[:, 3] -= tr
    out[:, 5] -= tr
    return -2.0 * nut[:, None] * out


def train(ds: Dataset, steps: int = 2000, batch: int = 65536, lr: float = 1e-3, seed: int = 0,
          val_frac: float = 0.1) -> Trained:
    """Adam on the GPU (the library's own forward/backward/Adam kernels) on a
    shuffled training split; returns the parameters, the scaling and the
    validation loss next to the Smagorinsky (Cs = 0.17) stress's."""
    norm = standardise(ds)
    xm, xs, ym, ys = norm
    rng = np.random.default_rng(seed)
    idx = rng.permutation(ds.X.shape[0])
    nval = int(len(idx) * val_frac)
    val, tr = idx[:nval], idx[nval:]
    Xn = ((ds.X - xm) / xs).astype(np.float32)
    Tn = ((ds.T - ym) / ys).astype(np.float32)
    h = Hit3D(16)   # a small grid: the closure's kernels do not use it
    try:
        h.mlp_set_params(init_params(seed))
        h.mlp_set_norm(xm, xs, ym, ys)
        h.mlp_set_data(Xn[tr], Tn[tr])
        losses = h.mlp_train(steps, min(batch, (len(tr) // 32) * 32), lr)
        p = h.mlp_get_params()
        Yv = h.mlp_forward(Xn[val])
    finally:
        h.close()
    val_loss = float(np.mean((Yv - Tn[val]) ** 2))
    sm = (smagorinsky_tau(ds.X[val], ds.delta, 0.17) - ym) / ys
    smag_val = float(np.mean((sm - Tn[val]) ** 2))
    return Trained(p, norm, losses, val_loss, smag_val)


def a_priori(tr: Trained, ds: Dataset, cs: float = 0.17) -> tuple:
    """Standardised MSE of the trained closure and of Smagorinsky on a whole
    dataset (for data the closure was not trained on), with the training
    set's scaling."""
    xm, xs, ym, ys = tr.norm
    Xn = ((ds.X - xm) / xs).astype(np.float32)
    Tn = ((ds.T - ym) / ys).astype(np.float32)
    h = Hit3D(16)
    try:
        h.mlp_set_params(tr.params)
        h.mlp_set_norm(xm, xs, ym, ys)
        Y = h.mlp_forward(Xn)
    finally:
        h.close()
    sm = (smagorinsky_tau(ds.X, ds.delta, cs) - ym) / ys
    return float(np.mean((Y - Tn) ** 2)), float(np.mean((sm - Tn) ** 2))


def dissipation_scale(tr: Trained, ds: Dataset) -> float:
    """Ratio of the data's mean subgrid dissipation <-tau:S> to that of the
    closure's stress as the LES applies it (after the backscatter clipping of
    k_les).  A regression fit shrinks its predictions toward the mean, and
    with them the energy it drains; scaling the stress by this ratio (a
    degree-one change: the clipping commutes with it) matches the mean
    drain of the data it was trained on."""
    xm, xs, ym, ys = tr.norm
    Xn = ((ds.X - xm) / xs).astype(np.float32)
    h = Hit3D(16)
    try:
        h.mlp_set_params(tr.params)
        h.mlp_set_norm(xm, xs, ym, ys)
        Y = h.mlp_forward(Xn)
    finally:
        h.close()
    tau = Y.astype(np.float64) * ys + ym
    g = ds.X[:, :9].astype(np.float64)
    S = np.stack([g[:, 0], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]), g[:, 4],
                  0.5 * (g[:, 5] + g[:, 7]), g[:, 8]], axis=1)
    w = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
    ts = (w * tau * S).sum(1)
    ss = (w * S * S).sum(1)
    back = (ts > 0) & (ss > 0)
    tau[back] += (-ts[back] / ss[back])[:, None] * S[back]
    d_pred = -(w * tau * S).sum(1).mean()
    d_true = -(w * ds.T.astype(np.float64) * S).sum(1).mean()
    return float(d_true / d_pred)


def les_run(ubar0: np.ndarray, model: str, steps: int, trained: Trained | None = None,
            cs: float = 0.17, re: float = 800.0, sample_every: int = 50, dt0: float = 2e-3,
            mix: float = 0.0):
    """Forced coarse LES from the filtered field ubar0 (physical, [3, m, m, m])
    with the learned closure (model 'mlp', its stress blended with weight mix
    of the Smagorinsky stress) or Smagorinsky ('smag').  Returns
    the time-averaged E(k) (shells 1..kmax), the energy history and the count
    of fallback steps."""
    m = ubar0.shape[1]
    p = dns_params(m, re)
    p.dt = dt0
    h = Hit3D(m)
    try:
        h.set_physics(p.nu, flow_type=1, kfmax=p.kfmax, famp=p.famp)
        for c in range(3):
            a = np.zeros((m, m, m + 2))
            a[:, :, :m] = ubar0[c]
            h.put(WRK, 1, a)
            h.xFFT3d(1, 1)
            h.copy(FIELDS, c + 1, WRK, 1)
        h.dealias_all()
        h.pressure()
        if model == "mlp":
            h.mlp_set_params(trained.params)
            h.mlp_set_norm(*trained.norm)
            h.mlp_set_mix(mix)
            h.set_les(11, cs)
        elif model == "smag":
            h.set_les(1, cs)
        else:
            raise ValueError("model must be 'mlp' or 'smag'")
        dt, fov, fos = p.dt, True, True
        eks, energy = [], []
        for i in range(steps):
            courant, en, fov, fos = h.step(dt, fov, fos)
            dt = my_dt(dt, courant)
            if not np.isfinite(courant) or courant > 10.0:
                raise FloatingPointError(f"LES ({model}) diverged at step {i}")
            if (i + 1) % sample_every == 0:
                e, _ = h.stats()
                eks.append(e)
                energy.append(float(e.sum()))
        fb = h.mlp_fallbacks() if model == "mlp" else 0
    finally:
        h.close()
    half = len(eks) // 2
    return np.mean(eks[half:], axis=0), np.array(energy), fb


def les_matched(ubar0: np.ndarray, model: str, times: list, trained: Trained | None = None,
                cs: float = 0.17, re: float = 800.0, dt0: float = 2e-3, mix: float = 0.0) -> list:
    """Forced coarse LES from the filtered field ubar0 for a given span of
    time: E(k) (shells 1..kmax) each time the LES clock first reaches one of
    `times` (offsets from the start, increasing), to set beside the filtered
    DNS at the same times."""
    m = ubar0.shape[1]
    p = dns_params(m, re)
    h = Hit3D(m)
    out = []
    try:
        h.set_physics(p.nu, flow_type=1, kfmax=p.kfmax, famp=p.famp)
        for c in range(3):
            a = np.zeros((m, m, m + 2))
            a[:, :, :m] = ubar0[c]
            h.put(WRK, 1, a)
            h.xFFT3d(1, 1)
            h.copy(FIELDS, c + 1, WRK, 1)
        h.dealias_all()
        h.pressure()
        if model == "mlp":
            h.mlp_set_params(trained.params)
            h.mlp_set_norm(*trained.norm)
            h.mlp_set_mix(mix)
            h.set_les(11, cs)
        else:
            h.set_les(1, cs)
        dt, fov, fos, t, i = dt0, True, True, 0.0, 0
        for target in times:
            while t < target:
                courant, _, fov, fos = h.step(dt, fov, fos)
                t += dt
                dt = my_dt(dt, courant)
                i += 1
                if not np.isfinite(courant) or courant > 10.0:
                    raise FloatingPointError(f"LES ({model}) diverged at step {i}")
            e, _ = h.stats()
            out.append(e)
    finally:
        h.close()
    return out


def spectrum_error(ek_les: np.ndarray, ek_ref: np.ndarray, kmin: int = 1) -> float:
    """Mean relative E(k) error over shells kmin..kmax of the LES."""
    k = min(len(ek_les), len(ek_ref))
    a, b = ek_les[kmin - 1:k], ek_ref[kmin - 1:k]
    return float(np.mean(np.abs(a - b) / b))


def compare_windows(a, snaps: list, m: int, delta: float, t_dns: float) -> dict:
    """The closure trained on the first window's snapshots, then for every
    window: a-priori errors on that window's data (held out in time for the
    later windows) and 64^3 LES runs with the closure and with Smagorinsky
    from the window's last filtered snapshot, their time-averaged E(k)
    against the window's time-averaged filtered DNS."""
    kmax = m // 3
    wins = [snaps[w * a.snaps:(w + 1) * a.snaps] for w in range(a.windows)]
    sets = [features(ws, m, delta) for ws in wins]
    t0 = time.time()
    tr = train(sets[0], a.train_steps, a.batch, a.lr)
    log(f"[sgs] trained on window 1: validation {tr.val_loss:.4f} vs Smagorinsky "
        f"{tr.smag_val_loss:.4f} ({time.time() - t0:.1f}s)")
    # the data's mean subgrid dissipation over the closure's (after the
    # backscatter clearing)
    cal = dissipation_scale(tr, sets[0])
    log(f"[sgs] mean dissipation, data over closure: {cal:.3f}")
    # the network alone, and blended half and half with Smagorinsky's stress;
    # the control smag_half is the blend's Smagorinsky part alone (constant
    # cs / sqrt 2: half the stress)
    models = {"smag": ("smag", None, 0.0, a.cs), "mlp": ("mlp", tr, 0.0, a.cs),
              "mlp_mix": ("mlp", tr, 0.5, a.cs), "smag_half": ("smag", None, 0.0, a.cs / math.sqrt(2))}
    out = []
    for w, (ws, ds) in enumerate(zip(wins, sets)):
        ek_ref = np.mean(ds.ek_filtered, axis=0)
        if w == 0:
            ap_mlp, ap_smag = tr.val_loss, tr.smag_val_loss   # the held-out 10 %
        else:
            ap_mlp, ap_smag = a_priori(tr, ds, a.cs)
        row = {"window": w + 1, "t_first": ws[0].time, "t_last": ws[-1].time,
               "a_priori_mlp": ap_mlp, "a_priori_smagorinsky": ap_smag}
        offsets = [sn.time - ws[0].time for sn in ws[1:]]
        for model, (kind, trm, mix, cs) in models.items():
            try:
                ek, energy, fb = les_run(ws[-1].ubar, kind, a.les_steps, trm, cs, a.re, mix=mix)
                row[f"ek_err_{model}"] = spectrum_error(ek[:kmax], ek_ref[:kmax])
                row[f"fallback_steps_{model}"] = int(fb)
            except FloatingPointError as e:
                row[f"ek_err_{model}"] = None
                row[f"error_{model}"] = str(e)
            # time-matched: from the window's first filtered snapshot, E(k) at
            # the times of the others against theirs
            try:
                eks = les_matched(ws[0].ubar, kind, offsets, trm, cs, a.re, mix=mix)
                row[f"ek_err_matched_{model}"] = float(np.mean(
                    [spectrum_error(e[:kmax], r[:kmax]) for e, r in zip(eks, ds.ek_filtered[1:])]))
                # E_LES / E_DNS per shell, averaged over the matched times
                row[f"ek_ratio_matched_{model}"] = [round(float(x), 3) for x in np.mean(
                    [e[:kmax] / r[:kmax] for e, r in zip(eks, ds.ek_filtered[1:])], axis=0)]
            except FloatingPointError as e:
                row[f"ek_err_matched_{model}"] = None
                row[f"error_matched_{model}"] = str(e)
        log(f"[sgs] window {w + 1} (t {row['t_first']:.2f}-{row['t_last']:.2f}): a priori "
            f"{ap_mlp:.3f} vs {ap_smag:.3f}; a posteriori {row.get('ek_err_mlp')} vs "
            f"{row.get('ek_err_smag')} (blend {row.get('ek_err_mlp_mix')}); time-matched "
            f"{row.get('ek_err_matched_mlp')} vs {row.get('ek_err_matched_smag')} (blend "
            f"{row.get('ek_err_matched_mlp_mix')}; Smagorinsky at cs/sqrt 2: "
            f"{row.get('ek_err_smag_half')}, {row.get('ek_err_matched_smag_half')})")
        out.append(row)
    wins_post = sum(1 for r in out if r.get("ek_err_mlp") is not None and r.get("ek_err_smag")
                    is not None and r["ek_err_mlp"] < r["ek_err_smag"])
    wins_prio = sum(1 for r in out if r["a_priori_mlp"] < r["a_priori_smagorinsky"])
    wins_matched = sum(1 for r in out if r.get("ek_err_matched_mlp") is not None
                       and r.get("ek_err_matched_smag") is not None
                       and r["ek_err_matched_mlp"] < r["ek_err_matched_smag"])
    wins_mix = sum(1 for r in out if r.get("ek_err_mlp_mix") is not None and r.get("ek_err_smag")
                   is not None and r["ek_err_mlp_mix"] < r["ek_err_smag"])
    wins_mix_matched = sum(1 for r in out if r.get("ek_err_matched_mlp_mix") is not None
                           and r.get("ek_err_matched_smag") is not None
                           and r["ek_err_matched_mlp_mix"] < r["ek_err_matched_smag"])
    wins_mix_half = sum(1 for r in out if r.get("ek_err_mlp_mix") is not None
                        and r.get("ek_err_smag_half") is not None
                        and r["ek_err_mlp_mix"] < r["ek_err_smag_half"]
                        and r["ek_err_matched_mlp_mix"] < r["ek_err_matched_smag_half"])
    return {"dns": a.dns, "ratio": a.ratio, "les_grid": m, "windows": out,
            "blend_better_than_its_smagorinsky_part_both_measures": wins_mix_half,
            "dissipation_data_over_closure": cal,
            "blend_better_a_posteriori": wins_mix,
            "blend_better_time_matched": wins_mix_matched,
            "closure_better_a_priori": wins_prio, "closure_better_a_posteriori": wins_post,
            "closure_better_time_matched": wins_matched,
            "train_steps": a.train_steps, "batch": a.batch, "snaps_per_window": a.snaps,
            "every": a.every, "spinup": a.spinup, "les_steps": a.les_steps,
            "dns_seconds": t_dns}


def main(argv=None):
    """Train the closure on box-filtered DNS and compare the coarse LES with
    it and with Smagorinsky against the filtered-DNS spectrum; prints one
    JSON line (and saves the parameters with --out)."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--dns", type=int, default=512)
    ap.add_argument("--ratio", type=int, default=8)
    ap.add_argument("--re", type=float, default=800.0)
    ap.add_argument("--spinup", type=int, default=6000)
    ap.add_argument("--snaps", type=int, default=8)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--les-steps", type=int, default=5000)
    ap.add_argument("--cs", type=float, default=0.17)
    ap.add_argument("--out", default="",
                    help="write the trained closure as a closure file (<run_name>.sgs) for "
                         "les_model 11, with the blend weight --mix")
    ap.add_argument("--mix", type=float, default=0.5,
                    help="weight of the Smagorinsky stress in the blended LES stress (the "
                         "model 'mlp_mix' of the comparison and the weight written by --out)")
    ap.add_argument("--windows", type=int, default=1,
                    help="consecutive DNS windows of --snaps snapshots each: the closure is "
                         "trained on the first, then compared a priori and a posteriori against "
                         "every window's own filtered DNS")
    a = ap.parse_args(argv)
    m = a.dns // a.ratio
    delta = 2.0 * math.pi / m
    t0 = time.time()
    snaps = run_dns(a.dns, a.ratio, a.spinup, a.snaps * a.windows, a.every, a.re)
    t_dns = time.time() - t0
    if a.windows > 1:
        print(json.dumps(compare_windows(a, snaps, m, delta, t_dns)), flush=True)
        return
    ds = features(snaps, m, delta)
    ek_ref = np.mean(ds.ek_filtered, axis=0)
    log(f"[sgs] dataset {ds.X.shape[0]} samples on {m}^3 ({t_dns:.0f}s of DNS)")
    t0 = time.time()
    tr = train(ds, a.train_steps, a.batch, a.lr)
    t_train = time.time() - t0
    log(f"[sgs] trained: loss {tr.losses[0]:.4f} -> {tr.losses[-20:].mean():.4f}, "
        f"validation {tr.val_loss:.4f} vs Smagorinsky a priori {tr.smag_val_loss:.4f} ({t_train:.1f}s)")
    res = {"dns": a.dns, "ratio": a.ratio, "les_grid": m, "samples": int(ds.X.shape[0]),
           "train_steps": a.train_steps, "batch": a.batch, "lr": a.lr,
           "train_loss_first": float(tr.losses[0]), "train_loss_last": float(tr.losses[-20:].mean()),
           "val_loss_mlp": tr.val_loss, "val_loss_smagorinsky": tr.smag_val_loss,
           "train_seconds": t_train, "dns_seconds": t_dns}
    u0 = snaps[-1].ubar
    kmax = m // 3
    for model, kind, mix in (("smag", "smag", 0.0), ("mlp", "mlp", 0.0), ("mlp_mix", "mlp", a.mix)):
        t0 = time.time()
        try:
            ek, energy, fb = les_run(u0, kind, a.les_steps, tr, a.cs, a.re, mix=mix)
            res[f"ek_err_{model}"] = spectrum_error(ek[:kmax], ek_ref[:kmax])
            res[f"energy_{model}"] = float(energy[len(energy) // 2:].mean())
            res[f"fallback_steps_{model}"] = int(fb)
            res[f"ek_{model}"] = [float(v) for v in ek[:kmax]]
        except FloatingPointError as e:
            res[f"ek_err_{model}"] = None
            res[f"error_{model}"] = str(e)
        res[f"les_seconds_{model}"] = time.time() - t0
        log(f"[sgs] LES {model}: E(k) error {res.get(f'ek_err_{model}')} ({res[f'les_seconds_{model}']:.1f}s)")
    res["ek_filtered_dns"] = [float(v) for v in ek_ref[:kmax]]
    res["energy_filtered_dns"] = float(np.mean([e.sum() for e in ds.ek_filtered]))
    if a.out:
        from .driver import write_closure_file
        write_closure_file(a.out, tr.params, tr.norm, mix=a.mix)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
