"""Host-side driver: the reference's input file, time-step controller and
main loop (main.f90:22-303) over the GPU library.

This mirrors the Fortran driver for Python callers (tests, benchmark); the
Fortran drop-in is hit3d_amd/fortran/.  All field arithmetic runs on the GPU.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from .lib import FIELDS, Hit3D


@dataclass
class Params:
    """read_input_file (m_parameters.f90:172-511)."""
    nx: int = 32
    ny: int = 32
    nz_all: int = 32
    itmin: int = 0
    itmax: int = 10
    iprint1: int = 1
    iprint2: int = 10 ** 9
    iwrite4: int = 10 ** 9
    tmax: float = 1e6
    trescale: float = 1e6
    nrescale: int = 0
    tscalar: float = 0.0
    flow_type: int = 0
    re: float = 100.0
    dt: float = 0.01
    variable_dt: bool = False
    isp_type: int = 2
    ir_exp: int = 4
    peak_wavenum: float = 3.0
    force_type: int = 1
    kfmax: int = 2
    famp: float = 0.5
    dealias: int = 0
    det_rand: int = 0
    rn1: float = 67748937.0
    rn2: float = 93712139.0
    rn3: float = 10765397.0
    nptot: int = 0
    particles_tracking_scheme: int = 0
    starttime_particles: float = 0.0
    particles_filter_size: float = 0.0
    les_model: int = 0
    n_scalars: int = 0
    scalar_type: list = field(default_factory=list)
    sc: list = field(default_factory=list)
    ir_exp_sc: list = field(default_factory=list)
    peak_wavenum_sc: list = field(default_factory=list)
    reac_sc: list = field(default_factory=list)

    @property
    def nu(self) -> float:
        return 1.0 / self.re

    @property
    def pe(self):
        return [self.nu / s for s in self.sc]

    @property
    def kmax(self) -> int:
        if self.dealias == 0:
            return self.nx // 3
        return int(math.floor(float(self.nx) / 3.0 * math.sqrt(2.0)))


def _first(line: str) -> str:
    return line.split()[0]


def read_input_file(path: str) -> Params:
    """Parses the positional input file of m_parameters.f90:196-451 (run name
    checks aside).  A negative DT means a fixed step (:277-284)."""
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)
    nxt = lambda: next(it)  # noqa: E731
    p = Params()
    for _ in range(3):
        nxt()
    p.nx, p.ny, p.nz_all = (int(v) for v in nxt().split()[:3])
    nxt()
    p.itmin = int(_first(nxt()))
    p.itmax = int(_first(nxt()))
    p.iprint1 = int(_first(nxt()))
    p.iprint2 = int(_first(nxt()))
    p.iwrite4 = int(_first(nxt()))
    nxt()
    p.tmax = float(_first(nxt()))
    parts = nxt().split()
    p.trescale = float(parts[0])
    try:
        p.nrescale = int(parts[1])
    except (IndexError, ValueError):  # m_parameters.f90:8000
        p.nrescale = 1 if p.trescale > 0 else 0
    p.tscalar = float(_first(nxt()))
    nxt()
    p.flow_type = int(_first(nxt()))
    nxt()
    p.re = float(_first(nxt()))
    dt = float(_first(nxt()))
    p.variable_dt = dt >= 0.0
    p.dt = abs(dt)
    nxt()
    p.isp_type = int(_first(nxt()))
    p.ir_exp = int(float(_first(nxt())))
    p.peak_wavenum = float(_first(nxt()))
    nxt()
    p.force_type = int(_first(nxt()))
    p.kfmax = int(_first(nxt()))
    p.famp = float(_first(nxt()))
    nxt()
    p.dealias = int(_first(nxt()))
    nxt()
    p.det_rand = int(_first(nxt()))
    p.rn1 = float(_first(nxt()))
    p.rn2 = float(_first(nxt()))
    p.rn3 = float(_first(nxt()))
    nxt()
    p.nptot = int(_first(nxt()))
    p.particles_tracking_scheme = int(_first(nxt()))
    p.starttime_particles = float(_first(nxt()))
    p.particles_filter_size = float(_first(nxt()))
    nxt()
    p.les_model = int(_first(nxt()))
    nxt()
    if p.les_model > 0 and p.dealias != 0:  # m_parameters.f90:415-420
        p.dealias = 0
    p.n_scalars = int(_first(nxt()))
    nxt()
    if p.n_scalars > 0:
        nxt()
        for _ in range(p.n_scalars):
            v = nxt().split()
            p.scalar_type.append(int(v[0]))
            p.sc.append(float(v[1]))
            p.ir_exp_sc.append(float(v[2]))
            p.peak_wavenum_sc.append(float(v[3]))
            p.reac_sc.append(float(v[4]) if len(v) > 4 else 0.0)
    return p


def my_dt(dt: float, courant: float, int_scalars: bool = False, sc=()) -> float:
    """my_dt.f90:1-34 (flow_type <= 1)."""
    courant_max = 0.1
    if courant < courant_max:
        dt = dt * 1.01
        dt = min(dt, 0.01)
    else:
        dt = dt / 1.05
    if int_scalars:
        for s in sc:
            dt = min(dt, dt * courant_max / courant / s)
    return dt


@dataclass
class StatLine:
    itime: int
    time: float
    energy: float
    eps_v: float
    eta: float
    enstrophy: float
    re_lambda: float
    x_length: float
    lambda_: float
    re_lambda1: float
    tau_e: float
    etakmax: float
    e_spec: np.ndarray
    diss_spec: np.ndarray


def derived_stats(e_spec: np.ndarray, enstrophy: float, nu: float, re: float, kmax: int,
                  itime: int = 0, time: float = 0.0) -> StatLine:
    """stat_velocity (m_stats.f90:115-148) from the device shell sums."""
    energy = float(np.sum(e_spec))
    k = np.arange(1, kmax + 1, dtype=np.float64)
    eps_v = float(np.sum(e_spec * k ** 2 * 2.0 * nu))
    eta = (nu ** 3 / eps_v) ** 0.25
    uvar = 2.0 / 3.0 * energy
    x_length = math.pi / 2.0 * float(np.sum(e_spec / k)) / uvar
    lam = math.sqrt(15.0 * uvar * nu / eps_v)
    return StatLine(itime=itime, time=time, energy=energy, eps_v=eps_v, eta=eta,
                    enstrophy=enstrophy, re_lambda=uvar * math.sqrt(15.0 / eps_v * re),
                    x_length=x_length, lambda_=lam, re_lambda1=math.sqrt(uvar) * lam / nu,
                    tau_e=x_length / math.sqrt(uvar), etakmax=eta * kmax, e_spec=e_spec,
                    diss_spec=e_spec * k ** 2 * 2.0 * nu)


def fortran_e(v: float, w: int = 15, d: int = 6) -> str:
    """Fortran Ew.d output editing: sign, 0., d digits of the correctly
    rounded mantissa in [0.1, 1), E and a signed two-digit exponent (the E is
    dropped for exponents beyond 99), right-justified in w columns."""
    v = float(v)
    if v != v or v in (float("inf"), float("-inf")):
        return f"{'NaN' if v != v else ('-Infinity' if v < 0 else 'Infinity'):>{w}}"
    if v == 0.0:
        digits, e = "0" * d, 0
    else:
        m, ex = f"{abs(v):.{d - 1}e}".split("e")
        digits, e = m.replace(".", ""), int(ex) + 1
    sign = "-" if (v < 0 or (v == 0.0 and str(v).startswith("-"))) else ""
    exp = f"E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) <= 99 else f"{'-' if e < 0 else '+'}{abs(e):03d}"
    return f"{sign}0.{digits}{exp}".rjust(w)


def _e15(v: float) -> str:
    return fortran_e(v, 15, 6)


def write_stats(s: StatLine, directory: str):
    """stat1.gp / stat2.gp / es.gp in the reference's formats (m_stats.f90:152-185)."""
    p1 = os.path.join(directory, "stat1.gp")
    new = not os.path.exists(p1)
    with open(p1, "a") as f:
        if new:
            f.write(" # 1.itime 2.time         3.energy       4.diss         5.eta          "
                    "6.enstrophy    7.R_lambda\n")
        f.write(f"{s.itime:8d}" + "".join(_e15(v) for v in (s.time, s.energy, s.eps_v, s.eta,
                                                               s.enstrophy, s.re_lambda)) + "\n")
    p2 = os.path.join(directory, "stat2.gp")
    new = not os.path.exists(p2)
    with open(p2, "a") as f:
        if new:
            f.write("# 1.itime  2.time         3.int LS       4. lambda      5.R_lambda1    "
                    "6.tau_e        7.etakmax\n")
        f.write(f"{s.itime:8d}" + "".join(_e15(v) for v in (s.time, s.x_length, s.lambda_,
                                                               s.re_lambda1, s.tau_e,
                                                               s.etakmax)) + "\n")
    with open(os.path.join(directory, "es.gp"), "a") as f:
        f.write("\n\n")
        f.write(f"# ITIME={s.itime:7d} TIME={fortran_e(s.time, 17, 8)}\n")
        for k, (e, d) in enumerate(zip(s.e_spec, s.diss_spec), start=1):
            f.write(f"{k:4d}{_e15(e)}{_e15(d)}\n")


def write_scalar_stats(st: dict, itime: int, time: float, directory: str):
    """scNN.gp (m_stats.f90:66-80) and es_NN.gp (:421-430) per scalar from
    Hit3D.stat_scalars()."""
    for i in range(len(st["sc_diss"])):
        p = os.path.join(directory, f"sc{i + 1:02d}.gp")
        new = not os.path.exists(p)
        with open(p, "a") as f:
            if new:
                f.write("# 1.itime 2.time          3.sc.diss       4. mean      5.variance"
                        "    6.min     7.max\n")
            vals = (time, st["sc_diss"][i], st["mean"][i], st["variance"][i], st["sc_min"][i],
                    st["sc_max"][i])
            f.write(f"{itime:7d}" + "".join(_e15(v) for v in vals) + "\n")
        with open(os.path.join(directory, f"es_{i + 1:02d}.gp"), "a") as f:
            f.write("\n\n")
            f.write(f"# ITIME={itime:7d} TIME={fortran_e(time, 17, 8)}\n")
            for k, e in enumerate(st["es_spec"][i], start=1):
                f.write(f"{k:4d}{_e15(e)}\n")


# ----------------------------------------------------------------------------
# restart files and float32 snapshots
# ----------------------------------------------------------------------------

def write_restart(path: str, fields, n: int, nums_out: int, n_les: int, time: float, dt: float):
    """restart_write_parallel (restart_io.f90:240-331): 36-byte header of int32
    nx, ny, nz, nums, nles and FP64 TIME, DT; then each field as n^3 FP64
    values in the spectral (i = 1..n, kz, ky) order, the Nyquist pair
    i = n+1, n+2 dropped.  fields: spectral slots [z(ky), y(kz), n+2]."""
    hdr = np.array([n, n, n, nums_out, n_les], "<i4").tobytes() + np.array([time, dt], "<f8").tobytes()
    with open(path, "wb") as f:
        f.write(hdr)
        for a in fields:
            f.write(np.ascontiguousarray(np.asarray(a)[:, :, :n], "<f8").tobytes())


def read_restart(path: str, n: int):
    """restart_read_parallel (restart_io.f90:337-506): returns (header dict,
    fields [count, n, n, n+2] with the Nyquist pair zero)."""
    with open(path, "rb") as f:
        hdr = f.read(36)
        if len(hdr) != 36:
            raise ValueError(f"*** error: Cannot read restart header of {path}")
        nx1, ny1, nz1, nums1, nles1 = (int(v) for v in np.frombuffer(hdr[:20], "<i4"))
        time, dt = (float(v) for v in np.frombuffer(hdr[20:36], "<f8"))
        if (nx1, ny1, nz1) != (n, n, n):
            raise ValueError(f"*** error: Dimensions are different: {(nx1, ny1, nz1)} vs {n}")
        cnt = 3 + nums1 + nles1
        data = np.frombuffer(f.read(cnt * n ** 3 * 8), "<f8")
        if data.size != cnt * n ** 3:
            raise ValueError(f"*** error: restart file {path} is truncated")
    out = np.zeros((cnt, n, n, n + 2))
    out[:, :, :, :n] = data.reshape(cnt, n, n, n)
    return dict(nx=nx1, ny=ny1, nz=nz1, nums=nums1, nles=nles1, time=time, dt=dt), out


def write_tmp4(path: str, a: np.ndarray):
    """write_tmp4 (write_tmp4.f90:14-30): int32 (nx, ny, nz) then the float32
    array, x fastest, as one stream."""
    nz, ny, nx = a.shape
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, nz], "<i4").tobytes())
        f.write(np.ascontiguousarray(a, "<f4").tobytes())


# ----------------------------------------------------------------------------
# particles: initial positions and the binary particle file
# ----------------------------------------------------------------------------

_PT_REC = np.dtype([("ipart", "<i4"), ("xyz", "<f8", (3,))])  # packed, 28 bytes


def _to_grid_units(xyz: np.ndarray, n: int) -> np.ndarray:
    """m_particles.f90:320-324: x / (2 pi / n) + 1 (one rank)."""
    return xyz / (2.0 * math.pi / float(n)) + 1.0


def particles_init(nptot: int, n: int, pt_text: str | None = None):
    """particles_init (m_particles.f90:157-337) on one rank: positions from
    the text file <run_name>.pt (3 numbers per particle, physical units) or,
    without it, the lattice of the nearest cube to nptot.  Positions are
    wrapped into [0, 2 pi) and converted to grid units.  Returns (ipart,
    xyzp[np, 3])."""
    two_pi = 2.0 * math.pi
    if pt_text is not None and os.path.exists(pt_text):
        with open(pt_text) as f:
            vals = np.array(f.read().replace(",", " ").split()[:3 * nptot], dtype=np.float64)
        if vals.size < 3 * nptot:
            raise ValueError(f"PARTICLES_INIT: {pt_text} holds fewer than {nptot} particles")
        xyz = vals.reshape(nptot, 3)
    else:
        npthird = int(float(nptot) ** (1.0 / 3.0))
        if nptot - npthird ** 3 >= (npthird + 1) ** 3 - nptot:
            npthird += 1
        nptot = min(npthird ** 3, 2 ** 30)
        sctmp = two_pi / float(npthird)
        i = np.arange(npthird, dtype=np.float64)
        kk, jj, ii = np.meshgrid(i, i, i, indexing="ij")
        xyz = np.stack([(ii.ravel() + 0.5) * sctmp, (jj.ravel() + 0.5) * sctmp,
                        (kk.ravel() + 0.5) * sctmp], axis=1)
    bad = (xyz >= two_pi) | (xyz < 0.0)
    xyz = np.where(bad, xyz - two_pi * np.floor(xyz / two_pi), xyz)
    return np.arange(1, nptot + 1, dtype=np.int32), _to_grid_units(xyz, n)


def write_particles_binary(path: str, ipart: np.ndarray, xyzp: np.ndarray, n: int, time: float):
    """particles_restart_write_binary (m_particles.f90:558-648): stream file
    of int32 nptot, then per particle int32 id and the FP64 position
    (x - 1) * dx, then the FP64 time."""
    dx = 2.0 * math.pi / float(n)
    rec = np.empty(len(ipart), _PT_REC)
    rec["ipart"] = ipart
    rec["xyz"] = (xyzp - 1.0) * dx
    with open(path, "wb") as f:
        f.write(np.array([len(ipart)], "<i4").tobytes())
        f.write(rec.tobytes())
        f.write(np.array([time], "<f8").tobytes())


def read_particles_binary(path: str, n: int):
    """particles_restart_read_binary (m_particles.f90:655-744) on one rank.
    Returns (ipart, xyzp[np, 3] in grid units, time)."""
    with open(path, "rb") as f:
        nptot = int(np.frombuffer(f.read(4), "<i4")[0])
        rec = np.frombuffer(f.read(_PT_REC.itemsize * nptot), _PT_REC)
        if len(rec) != nptot:
            raise ValueError(f"Error reading particle restart file {path}")
        tail = f.read(8)
    t = float(np.frombuffer(tail, "<f8")[0]) if len(tail) == 8 else float("nan")
    return rec["ipart"].astype(np.int32), _to_grid_units(rec["xyz"].astype(np.float64), n), t


def write_closure_file(path: str, params: np.ndarray, norm, mix: float = 0.0):
    """The learned closure for les_model 11 as the Fortran driver reads it
    (<run_name>.sgs): 6288 float32 parameters, then float64 xmean(10),
    xstd(10), ymean(6), ystd(6) and the weight of the Smagorinsky stress in
    the LES stress (hit3d_mlp_set_mix)."""
    xm, xs, ym, ys = (np.asarray(v, "<f8") for v in norm)
    with open(path, "wb") as f:
        f.write(np.asarray(params, "<f4").tobytes())
        for v in (xm, xs, ym, ys, np.array([mix], "<f8")):
            f.write(v.tobytes())


def read_closure_file(path: str):
    """Returns (params[6288] float32, (xmean, xstd, ymean, ystd), mix); a
    file without the trailing weight (32 doubles) has mix 0."""
    raw = open(path, "rb").read()
    if len(raw) not in (6288 * 4 + 32 * 8, 6288 * 4 + 33 * 8):
        raise ValueError(f"{path}: not a closure file ({len(raw)} bytes)")
    params = np.frombuffer(raw[:6288 * 4], "<f4").copy()
    d = np.frombuffer(raw[6288 * 4:], "<f8")
    mix = float(d[32]) if len(d) > 32 else 0.0
    return params, (d[0:10].copy(), d[10:20].copy(), d[20:26].copy(), d[26:32].copy()), mix


class Run:
    """A run of the driver loop (main.f90:56-288) on the GPU library."""

    def __init__(self, p: Params, precision: int = 0, closure: str | None = None):
        if p.nx != p.ny or p.nx != p.nz_all:
            raise ValueError("only cubic grids are supported")
        if p.les_model not in (0, 1, 11):
            raise ValueError("only les_model 0 (none), 1 (Smagorinsky) and 11 (the learned "
                             "closure) are built")
        if p.les_model == 11 and closure is None:
            raise ValueError("les_model 11 needs the closure file (<run_name>.sgs)")
        if p.nptot > 0 and p.particles_filter_size > 0.0:
            raise ValueError("particles on a locally averaged velocity are not built")
        self.p = p
        self.precision = precision
        self.h = Hit3D(p.nx, n_scalars=p.n_scalars, dealias=p.dealias)
        self.h.set_physics(p.nu, flow_type=p.flow_type, kfmax=p.kfmax, famp=p.famp, pe=p.pe,
                           scalar_type=p.scalar_type)
        if p.n_scalars > 0:
            self.h.set_reaction(p.reac_sc if len(p.reac_sc) else [0.0] * p.n_scalars)
        if p.les_model == 1:
            self.h.set_les(1, 0.18)   # m_les.f90:1148
        if p.les_model == 11:
            params, norm, mix = read_closure_file(closure)
            self.h.mlp_set_params(params)
            self.h.mlp_set_norm(*norm)
            self.h.mlp_set_mix(mix)
            self.h.set_les(11, 0.18)
        self.ipart = None
        self.int_particles = False
        self.time = 0.0
        self.dt = p.dt
        self.itime = p.itmin
        self.fov = True
        self.fos = True
        self.int_scalars = False
        self.courant = 0.0

    def begin_new(self):
        """begin_new.f90:10-31 + main.f90:71-76: velocity, scalars (integrated
        from the start when TSCALAR <= 0), dealias_all."""
        p = self.p
        self.time = 0.0
        self.int_scalars = p.tscalar <= 0.0 and p.n_scalars > 0
        self.h.init_velocity(p.rn1, p.isp_type, p.ir_exp, p.peak_wavenum)
        if p.n_scalars > 0:
            self.h.init_scalars(p.rn2, p.peak_wavenum_sc)
        self.fov = self.fos = True
        self.h.dealias_all()

    def restart_write(self, run_name: str, directory: str = "."):
        """restart_write_parallel (restart_io.f90:240-331) to <run>.64.<itime>."""
        p = self.p
        nums_out = p.n_scalars if self.int_scalars else 0
        fields = [self.h.get(FIELDS, c) for c in range(1, 4 + nums_out)]
        path = os.path.join(directory, f"{run_name}.64.{self.itime:06d}")
        write_restart(path, fields, p.nx, nums_out, 0, self.time, self.dt)
        return path

    def begin_restart(self, run_name: str, directory: str = "."):
        """begin_restart (begin_restart.f90:1-90) with restart_read_parallel:
        fields from <run>.64.<itmin>; scalars missing from the file are
        initialised (init_scalar, restart_io.f90:420-431), extra ones dropped;
        dt halved for variable dt; scalars integrated when TIME > TSCALAR."""
        p = self.p
        self.itime = p.itmin
        hdr, f = read_restart(os.path.join(directory, f"{run_name}.64.{p.itmin:06d}"), p.nx)
        self.time, self.dt = hdr["time"], hdr["dt"]
        nums_read = min(p.n_scalars, hdr["nums"])
        for c in range(3 + nums_read):
            self.h.put(FIELDS, c + 1, f[c])
        for s in range(nums_read + 1, p.n_scalars + 1):
            self.h.init_scalar(s, p.rn2, p.peak_wavenum_sc[s - 1])
        if p.variable_dt:
            self.dt = 0.5 * self.dt
        self.int_scalars = p.n_scalars > 0 and self.time > p.tscalar
        self.fov = self.fos = True
        # main.f90:73-76: dealias_all after begin_restart too -- truncates
        # what the file or init_scalar put outside the kmax cube
        self.h.dealias_all()

    def write_snapshot(self, directory: str = "."):
        """io_write_4 (io_write_4.f90:1-85): u., v., w.<itime> and scNN.<itime>
        as float32 with the k > kmax modes removed, and nu_t.<itime> (the
        turbulent viscosity of the last right-hand side, :73-79) when the
        Smagorinsky model is on.  The turbulent viscosity exists once a
        step has run; before that, as the reference's zero-initialised
        array, it is written as zeros."""
        ext = f"{self.itime:06d}"
        for c, name in ((1, "u"), (2, "v"), (3, "w")):
            write_tmp4(os.path.join(directory, f"{name}.{ext}"), self.h.snapshot(c))
        if self.int_scalars:
            for s in range(1, self.p.n_scalars + 1):
                write_tmp4(os.path.join(directory, f"sc{s:02d}.{ext}"), self.h.snapshot(3 + s))
        if self.p.les_model == 1:
            write_tmp4(os.path.join(directory, f"nu_t.{ext}"), self.h.turb_visc())

    def init_particles(self, run_name: str | None = None, directory: str = "."):
        """particles_init (m_particles.f90:157-337): pt.<itmin> binary restart
        if present, else <run_name>.pt, else the lattice."""
        p = self.p
        if p.nptot <= 0:
            return
        binary = os.path.join(directory, f"pt.{p.itmin:06d}")
        if os.path.exists(binary):
            self.ipart, xyz, _ = read_particles_binary(binary, p.nx)
        else:
            text = os.path.join(directory, f"{run_name}.pt") if run_name else None
            self.ipart, xyz = particles_init(p.nptot, p.nx, text)
        self.h.particles_set(xyz)

    def write_particles(self, directory: str = "."):
        """particles_restart_write_binary (m_particles.f90:558-648)."""
        xyz, _ = self.h.particles_get()
        write_particles_binary(os.path.join(directory, f"pt.{self.itime:06d}"), self.ipart, xyz,
                               self.p.nx, self.time)

    def move_particles(self):
        """The parts task of main.f90:235-247: velocity of this time level
        (fields_to_parts), then particles_move with this step's dt."""
        p = self.p
        if self.ipart is None or self.time < p.starttime_particles:
            return
        self.int_particles = True
        self.h.fields_to_parts()
        self.h.particles_move(self.dt, p.particles_tracking_scheme)

    def step(self):
        """main.f90:90-207 for one iteration (no rescaling, restart or output);
        particles, when present, move first with the velocity and dt the
        hydro step starts from."""
        p = self.p
        self.move_particles()
        self.itime += 1
        self.courant, self.energy, self.fov, self.fos = self.h.step(
            self.dt, self.fov, self.fos, int_scalars=self.int_scalars, precision=self.precision)
        self.time += self.dt
        if p.variable_dt:
            self.dt = my_dt(self.dt, self.courant, self.int_scalars, p.sc)
        if (self.itime % p.iprint1 == 0 or self.itime % p.iwrite4 == 0):
            if p.n_scalars > 0 and not self.int_scalars and self.time > p.tscalar:
                self.int_scalars = True  # main.f90:199-204
                self.h.init_scalars(p.rn2, p.peak_wavenum_sc)

    def run(self, run_name: str, directory: str = "."):
        """The whole driver loop with its output, as main.f90:56-297 (and
        hit3d_gpu.f90) orders it: start (begin_new with its io_write_4, or
        begin_restart), particles, then per step the particle move, the
        hydro step, the restart at IPRINT2 (before my_dt), my_dt,
        statistics at IPRINT1, snapshots and the particle file at IWRITE4,
        the TMAX and TSCALAR checks; a final restart file."""
        p = self.p
        if p.itmin == 0:
            self.begin_new()
            self.itime = 0
            self.write_snapshot(directory)
        else:
            self.begin_restart(run_name, directory)
        self.divergence()
        if p.nptot > 0:
            self.init_particles(run_name, directory)
        itime = p.itmin
        for itime in range(p.itmin + 1, p.itmax + 1):
            self.move_particles()
            self.itime = itime
            self.courant, self.energy, self.fov, self.fos = self.h.step(
                self.dt, self.fov, self.fos, int_scalars=self.int_scalars,
                precision=self.precision)
            self.time += self.dt
            if itime % p.iprint2 == 0:
                self.restart_write(run_name, directory)
            if p.variable_dt:
                self.dt = my_dt(self.dt, self.courant, self.int_scalars, p.sc)
            if itime % p.iprint1 == 0:
                write_stats(self.stats(), directory)
                if self.int_scalars:
                    write_scalar_stats(self.scalar_stats(), itime, self.time, directory)
            if itime % p.iwrite4 == 0:
                self.write_snapshot(directory)
                if self.int_particles:
                    self.write_particles(directory)
            if itime % p.iprint1 == 0 or itime % p.iwrite4 == 0:
                if self.time > p.tmax:
                    break
                if p.n_scalars > 0 and not self.int_scalars and self.time > p.tscalar:
                    self.int_scalars = True  # main.f90:199-204
                    self.h.init_scalars(p.rn2, p.peak_wavenum_sc)
        self.itime = itime
        if p.itmax > p.itmin and itime % p.iprint2 != 0:
            self.restart_write(run_name, directory)

    def scalar_stats(self) -> dict:
        """stat_scalars (m_stats.f90:295-366); the reference runs it only once
        scalars are integrated (:62)."""
        return self.h.stat_scalars()

    def divergence(self):
        """pressure.f90:53-93, run by main.f90:66 before the loop."""
        return self.h.divergence()

    def stats(self) -> StatLine:
        e, ens = self.h.stats()
        return derived_stats(e, ens, self.p.nu, self.p.re, self.p.kmax, self.itime, self.time)
