"""ctypes binding of libhit3d_hip.so (the C ABI declared in include/hit3d.h).

The library is the product: every call here runs HIP kernels on the GPU.
There is no CPU fallback; a missing library or a missing GPU raises.

Method names follow the reference's Fortran subroutines (x_fftw.f90,
rhs_velocity.f90, rhs_scalars.f90, m_force.f90, pressure.f90,
dealias_all.f90, m_stats.f90, init_velocity.f90); slots and array ids have
the reference's meaning (fields 1..nf, wrk 0..nw, rhs_old 1..nf).
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhit3d_hip.so")

FIELDS, WRK, RHS_OLD = 1, 2, 3
_AXES = {"x": 1, "y": 2, "z": 3}

_I64P = ctypes.POINTER(ctypes.c_int64)
_F64P = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "hit3d_init": [_I64P] * 6,
    "hit3d_set_physics": [_F64P, _I64P, _I64P, _F64P, _F64P, _I64P],
    "hit3d_set_reaction": [_F64P],
    "hit3d_finalize": [],
    "hit3d_put": [_I64P, _I64P, _F64P],
    "hit3d_get": [_I64P, _I64P, _F64P],
    "hit3d_copy": [_I64P] * 4,
    "hit3d_xfft3d": [_I64P, _I64P],
    "hit3d_xfft3d_fields": [_I64P, _I64P],
    "hit3d_x_derivative": [_I64P, _I64P, _I64P],
    "hit3d_dealias_all": [],
    "hit3d_rhs_scalars": [_I64P],
    "hit3d_rhs_velocity": [_F64P, _F64P],
    "hit3d_force_velocity": [_F64P],
    "hit3d_advance": [_I64P, _I64P, _F64P],
    "hit3d_pressure": [],
    "hit3d_step": [_F64P, _I64P, _I64P, _I64P, _I64P, _F64P, _F64P],
    "hit3d_set_fusion": [_I64P],
    "hit3d_set_nonlinear": [_I64P],
    "hit3d_set_graphs": [_I64P],
    "hit3d_graph_stats": [_I64P, _I64P],
    "hit3d_mlp_set_params": [ctypes.POINTER(ctypes.c_float)],
    "hit3d_mlp_get_params": [ctypes.POINTER(ctypes.c_float)],
    "hit3d_mlp_set_norm": [_F64P] * 4,
    "hit3d_mlp_set_data": [_I64P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "hit3d_mlp_train": [_I64P, _I64P, _F64P, _F64P],
    "hit3d_mlp_grads": [_I64P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                        ctypes.POINTER(ctypes.c_float), _F64P],
    "hit3d_mlp_forward": [_I64P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "hit3d_mlp_grads_static": [_I64P, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                               _F64P],
    "hit3d_mlp_fallbacks": [_I64P],
    "hit3d_mlp_set_mix": [_F64P],
    "hit3d_box_filter": [_I64P, _F64P, _F64P],
    "hit3d_sgs_features": [_F64P, ctypes.POINTER(ctypes.c_float)],
    "hit3d_set_les": [_I64P, _F64P],
    "hit3d_stats": [_F64P, _F64P],
    "hit3d_stat_scalars": [_F64P] * 6,
    "hit3d_divergence": [_F64P, _F64P],
    "hit3d_skewness": [_F64P],
    "hit3d_particles_set": [_I64P, _F64P, _F64P],
    "hit3d_particles_get": [_F64P, _F64P],
    "hit3d_fields_to_parts": [],
    "hit3d_particles_move": [_F64P, _I64P],
    "hit3d_init_velocity": [_F64P, _I64P, _I64P, _F64P],
    "hit3d_random_draws": [_F64P, _I64P, _I64P, _F64P],
    "hit3d_init_gaussian": [_I64P, _F64P, _I64P, _F64P],
    "hit3d_init_scalars": [_F64P, _F64P],
    "hit3d_init_scalar": [_I64P, _F64P, _F64P],
    "hit3d_rng_state": [_I64P],
    "hit3d_snapshot": [_I64P, ctypes.POINTER(ctypes.c_float)],
    "hit3d_turb_visc": [ctypes.POINTER(ctypes.c_float)],
    "hit3d_scalar_gaussian": [_I64P, _F64P, _I64P, _F64P, _I64P],
    "hit3d_velocity_rng_state": [_I64P, _F64P, _I64P],
    "hit3d_get_ialias": [ctypes.POINTER(ctypes.c_int8)],
    "hit3d_timing_enable": [_I64P],
    "hit3d_timing_read": [_I64P, _F64P, _I64P],
    "hit3d_timing_names": [ctypes.c_char_p, _I64P],
    "hit3d_synchronize": [],
    "hit3d_last_error": [ctypes.c_char_p, _I64P],
}


class Hit3dError(RuntimeError):
    """A non-zero status from the library (the reference would call my_exit(-1))."""


_lib = None


def load(path: str = LIB_PATH):
    """Load the shared library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise Hit3dError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = ctypes.CDLL(path)
        for name, args in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def _i(v):
    return ctypes.byref(ctypes.c_int64(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _dp(a):
    return a.ctypes.data_as(_F64P)


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        _lib.hit3d_last_error(buf, _i(1024))
        raise Hit3dError(buf.value.decode(errors="replace"))


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().hit3d_last_error(buf, _i(1024))
    return buf.value.decode(errors="replace")


def random_draws(seed: float, skip: int, count: int) -> np.ndarray:
    """RANDu.f:12-52: fresh generator, random(-seed), skip, then count draws."""
    lib = load()
    out = np.empty(count, np.float64)
    _check(lib.hit3d_random_draws(_d(seed), _i(skip), _i(count), _dp(out)))
    return out


def init_gaussian(n: int, rn1: float, c: int) -> np.ndarray:
    """init_velocity.f90:87-113 Gaussian field of component c (1..3), host code."""
    lib = load()
    out = np.empty((n, n, n + 2))
    _check(lib.hit3d_init_gaussian(_i(n), _d(rn1), _i(c), _dp(out)))
    return out


def scalar_gaussian(n: int, rn2: float, iy_prev: int):
    """init_scalars.f90:105-133 Gaussian field and the generator's saved iy
    afterwards, from a generator whose saved iy is iy_prev."""
    lib = load()
    out = np.zeros((n, n, n + 2))
    nxt = ctypes.c_int64(0)
    _check(lib.hit3d_scalar_gaussian(_i(n), _d(rn2), _i(iy_prev), _dp(out), ctypes.byref(nxt)))
    return out, nxt.value


def velocity_rng_state(n: int, rn1: float) -> int:
    lib = load()
    iy = ctypes.c_int64(0)
    _check(lib.hit3d_velocity_rng_state(_i(n), _d(rn1), ctypes.byref(iy)))
    return iy.value


_open_instance = None   # weak reference to the open Hit3D, if any


class Hit3D:
    """The device-resident state of one run.  The library holds one grid
    per process (hit3d_init replaces whatever state was there), so a second
    instance while one is open is refused rather than left to drive the
    first one's state."""

    def __init__(self, n: int, n_scalars: int = 0, n_les: int = 0, dealias: int = 0):
        global _open_instance
        other = _open_instance() if _open_instance is not None else None
        if other is not None and not other._closed:
            raise Hit3dError("the library holds one grid per process: close the open Hit3D "
                             f"(n={other.n}) before opening another")
        self._closed = True
        self.lib = load()
        self.n = n
        self.n_scalars = n_scalars
        self.n_les = n_les
        self.dealias = dealias
        self.nf = 3 + n_scalars + n_les
        self.nw = max(6, 3 + n_scalars + n_les + 2)
        _check(self.lib.hit3d_init(_i(n), _i(n), _i(n), _i(n_scalars), _i(n_les), _i(dealias)))
        self._closed = False
        _open_instance = weakref.ref(self)

    @property
    def kmax(self) -> int:
        if self.dealias == 0:
            return self.n // 3
        return int(np.floor(float(self.n) / 3.0 * np.sqrt(2.0)))

    def close(self):
        self._closed = True
        _check(self.lib.hit3d_finalize())

    # -- parameters and staging ---------------------------------------------
    def set_physics(self, nu, flow_type=0, kfmax=0, famp=0.0, pe=(), scalar_type=()):
        pe_a = np.ascontiguousarray(pe, np.float64) if len(pe) else np.zeros(1)
        st_a = np.ascontiguousarray(scalar_type, np.int64) if len(scalar_type) else np.zeros(1, np.int64)
        _check(self.lib.hit3d_set_physics(_d(nu), _i(flow_type), _i(kfmax), _d(famp), _dp(pe_a),
                                          st_a.ctypes.data_as(_I64P)))

    def set_reaction(self, reac_sc):
        """Reaction rates of scalars of type >= 100 (m_parameters.f90:445)."""
        r = np.ascontiguousarray(reac_sc, np.float64) if len(reac_sc) else np.zeros(1)
        _check(self.lib.hit3d_set_reaction(_dp(r)))

    def put(self, array_id: int, slot: int, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        if a.shape != (self.n, self.n, self.n + 2):
            raise ValueError("slot shape must be (n, n, n+2)")
        _check(self.lib.hit3d_put(_i(array_id), _i(slot), _dp(a)))

    def get(self, array_id: int, slot: int) -> np.ndarray:
        out = np.empty((self.n, self.n, self.n + 2))
        _check(self.lib.hit3d_get(_i(array_id), _i(slot), _dp(out)))
        return out

    def copy(self, dst_array, dst_slot, src_array, src_slot):
        _check(self.lib.hit3d_copy(_i(dst_array), _i(dst_slot), _i(src_array), _i(src_slot)))

    # -- x_fftw ----------------------------------------------------------------
    def xFFT3d(self, flag: int, n: int):
        _check(self.lib.hit3d_xfft3d(_i(flag), _i(n)))

    def xFFT3d_fields(self, flag: int, n: int):
        _check(self.lib.hit3d_xfft3d_fields(_i(flag), _i(n)))

    def x_derivative(self, n: int, axis: str, nto: int):
        if axis not in _AXES:
            raise Hit3dError(f"x_derivative: wrong value of axis: {axis}")
        _check(self.lib.hit3d_x_derivative(_i(n), _i(_AXES[axis]), _i(nto)))

    def ialias(self) -> np.ndarray:
        out = np.empty((self.n, self.n, self.n + 2), np.int8)
        _check(self.lib.hit3d_get_ialias(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))))
        return out

    # -- physics -----------------------------------------------------------------
    def dealias_all(self):
        _check(self.lib.hit3d_dealias_all())

    def rhs_scalars(self, int_scalars: bool = True):
        _check(self.lib.hit3d_rhs_scalars(_i(1 if int_scalars else 0)))

    def rhs_velocity(self, dt: float) -> float:
        c = ctypes.c_double(0.0)
        _check(self.lib.hit3d_rhs_velocity(_d(dt), ctypes.byref(c)))
        return c.value

    def force_velocity(self) -> float:
        e = ctypes.c_double(0.0)
        _check(self.lib.hit3d_force_velocity(ctypes.byref(e)))
        return e.value

    def advance(self, which: int, first_order: bool, dt: float) -> bool:
        f = ctypes.c_int64(1 if first_order else 0)
        _check(self.lib.hit3d_advance(_i(which), ctypes.byref(f), _d(dt)))
        return bool(f.value)

    def pressure(self):
        _check(self.lib.hit3d_pressure())

    def step(self, dt: float, fov: bool, fos: bool, int_scalars: bool = False, precision: int = 0):
        """main.f90:124-167 fused.  Returns (courant, energy, fov, fos)."""
        fv = ctypes.c_int64(1 if fov else 0)
        fs = ctypes.c_int64(1 if fos else 0)
        c = ctypes.c_double(0.0)
        e = ctypes.c_double(0.0)
        _check(self.lib.hit3d_step(_d(dt), ctypes.byref(fv), ctypes.byref(fs),
                                   _i(1 if int_scalars else 0), _i(precision), ctypes.byref(c),
                                   ctypes.byref(e)))
        return c.value, e.value, bool(fv.value), bool(fs.value)

    def set_les(self, les_model: int, c_smag: float = 0.18):
        _check(self.lib.hit3d_set_les(_i(les_model), _d(c_smag)))

    def set_nonlinear(self, form: int):
        """Nonlinear term of the fused steps: 0 the divergence form (the
        reference's), 1 the rotational u x omega form (2/3 rule), 2 five
        products instead of six: the trace-free tensor u_i u_j - delta_ij
        |u|^2/3 (2/3 rule) or u_i u_j - delta_ij w w (phase shift)."""
        _check(self.lib.hit3d_set_nonlinear(_i(form)))

    def set_fusion(self, on: bool):
        _check(self.lib.hit3d_set_fusion(_i(1 if on else 0)))

    def set_graphs(self, on: bool):
        """hit3d_step from captured HIP graphs (default on) or direct launches."""
        _check(self.lib.hit3d_set_graphs(_i(1 if on else 0)))

    def graph_stats(self):
        """(step graphs held, steps run from a graph so far)."""
        c, r = ctypes.c_int64(0), ctypes.c_int64(0)
        _check(self.lib.hit3d_graph_stats(ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def stats(self):
        """(e_spec[1..kmax], enstrophy) -- m_stats.f90:195-289."""
        e = np.zeros(self.kmax)
        ens = ctypes.c_double(0.0)
        _check(self.lib.hit3d_stats(_dp(e), ctypes.byref(ens)))
        return e, ens.value

    def stat_scalars(self) -> dict:
        """Per-scalar statistics of m_stats.f90:295-437 (stat_scalars,
        get_scalar_spectra): arrays over n = 1..n_scalars of sc_diss, mean,
        variance, min, max, and es_spec[n-1, 0:kmax] (already doubled)."""
        ns = self.n_scalars
        out = {k: np.zeros(ns) for k in ("sc_diss", "mean", "variance", "sc_min", "sc_max")}
        es = np.zeros((ns, self.kmax))
        if ns:
            _check(self.lib.hit3d_stat_scalars(_dp(out["sc_diss"]), _dp(out["mean"]),
                                               _dp(out["variance"]), _dp(out["sc_min"]),
                                               _dp(out["sc_max"]), _dp(es)))
        out["es_spec"] = es
        return out

    def skewness(self) -> float:
        """Mean longitudinal velocity-derivative skewness."""
        v = ctypes.c_double(0.0)
        _check(self.lib.hit3d_skewness(ctypes.byref(v)))
        return v.value

    def divergence(self):
        """(dmin, dmax) of div u in physical space -- pressure.f90:53-93."""
        lo, hi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(self.lib.hit3d_divergence(ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    # -- particles (m_particles.f90) ------------------------------------------
    def particles_set(self, xyzp: np.ndarray, uvwp: np.ndarray | None = None):
        """xyzp[np, 3] in grid units (1-based, as the reference's xyzp(3, np));
        uvwp None means -999 (first move is Euler)."""
        x = np.ascontiguousarray(xyzp, dtype=np.float64).reshape(-1, 3)
        u = None if uvwp is None else np.ascontiguousarray(uvwp, dtype=np.float64).reshape(-1, 3)
        if u is not None and u.shape != x.shape:
            raise ValueError("uvwp must have the shape of xyzp")
        self.np = x.shape[0]
        _check(self.lib.hit3d_particles_set(_i(self.np), _dp(x), None if u is None else _dp(u)))

    def particles_get(self):
        n = getattr(self, "np", 0)
        x = np.zeros((n, 3))
        u = np.zeros((n, 3))
        _check(self.lib.hit3d_particles_get(_dp(x), _dp(u)))
        return x, u

    def fields_to_parts(self):
        _check(self.lib.hit3d_fields_to_parts())

    def particles_move(self, dt: float, scheme: int):
        _check(self.lib.hit3d_particles_move(_d(dt), _i(scheme)))

    def init_velocity(self, rn1: float, isp_type: int, ir_exp: int, peak: float):
        _check(self.lib.hit3d_init_velocity(_d(rn1), _i(isp_type), _i(ir_exp), _d(peak)))

    def init_scalars(self, rn2: float, peak_sc=()):
        """init_scalars.f90:6-383 for every scalar (types from set_physics)."""
        pk = np.ascontiguousarray(peak_sc, np.float64) if len(peak_sc) else np.zeros(1)
        _check(self.lib.hit3d_init_scalars(_d(rn2), _dp(pk)))

    def init_scalar(self, n_scalar: int, rn2: float, peak: float):
        """init_scalar(n) (init_scalars.f90:47-79), no ialias mask."""
        _check(self.lib.hit3d_init_scalar(_i(n_scalar), _d(rn2), _d(peak)))

    def snapshot(self, n_field: int) -> np.ndarray:
        """io_write_4 for one field: float32 physical array [z, y, x]."""
        out = np.zeros((self.n, self.n, self.n), np.float32)
        _check(self.lib.hit3d_snapshot(_i(n_field), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def turb_visc(self) -> np.ndarray:
        """The Smagorinsky turbulent viscosity of the last LES right-hand
        side as io_write_4 writes it (nu_t.<ext>): float32 [z, y, x]."""
        out = np.zeros((self.n, self.n, self.n), np.float32)
        _check(self.lib.hit3d_turb_visc(out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    # -- learned closure --------------------------------------------------------
    NPARAM = 6288

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def mlp_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32).ravel()
        if p.size != self.NPARAM:
            raise ValueError(f"expected {self.NPARAM} parameters")
        _check(self.lib.hit3d_mlp_set_params(self._fp(p)))

    def mlp_get_params(self) -> np.ndarray:
        p = np.zeros(self.NPARAM, np.float32)
        _check(self.lib.hit3d_mlp_get_params(self._fp(p)))
        return p

    def mlp_set_norm(self, xmean, xstd, ymean, ystd):
        a = [np.ascontiguousarray(v, np.float64) for v in (xmean, xstd, ymean, ystd)]
        _check(self.lib.hit3d_mlp_set_norm(*[_dp(v) for v in a]))

    def mlp_set_data(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._mlp_data = (X, T)
        _check(self.lib.hit3d_mlp_set_data(_i(X.shape[0]), self._fp(X), self._fp(T)))

    def mlp_train(self, steps: int, batch: int, lr: float) -> np.ndarray:
        losses = np.zeros(steps)
        _check(self.lib.hit3d_mlp_train(_i(steps), _i(batch), _d(lr), _dp(losses)))
        return losses

    def mlp_grads(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.NPARAM, np.float32)
        loss = ctypes.c_double(0.0)
        _check(self.lib.hit3d_mlp_grads(_i(X.shape[0]), self._fp(X), self._fp(T), self._fp(g),
                                         ctypes.byref(loss)))
        return g, loss.value

    def mlp_grads_static(self, X: np.ndarray, T: np.ndarray):
        """mlp_grads through the allocation-free entry point (at most 8192
        samples per call)."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.NPARAM, np.float32)
        loss = ctypes.c_double(0.0)
        _check(self.lib.hit3d_mlp_grads_static(_i(X.shape[0]), self._fp(X), self._fp(T),
                                                self._fp(g), ctypes.byref(loss)))
        return g, loss.value

    def mlp_forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 6), np.float32)
        _check(self.lib.hit3d_mlp_forward(_i(X.shape[0]), self._fp(X), self._fp(Y)))
        return Y

    def mlp_set_mix(self, weight: float):
        """LES stress of les_model 11: (1 - weight) of the network's plus
        weight times the Smagorinsky stress of the set_les constant."""
        _check(self.lib.hit3d_mlp_set_mix(_d(weight)))

    def mlp_fallbacks(self) -> int:
        c = ctypes.c_int64(0)
        _check(self.lib.hit3d_mlp_fallbacks(ctypes.byref(c)))
        return c.value

    def box_filter(self, ratio: int):
        m = self.n // ratio
        ubar = np.zeros((3, m, m, m))
        tau = np.zeros((6, m, m, m))
        _check(self.lib.hit3d_box_filter(_i(ratio), _dp(ubar), _dp(tau)))
        return ubar, tau

    def sgs_features(self, delta: float) -> np.ndarray:
        X = np.zeros((self.n ** 3, 10), np.float32)
        _check(self.lib.hit3d_sgs_features(_d(delta), self._fp(X)))
        return X

    def rng_state(self) -> int:
        iy = ctypes.c_int64(0)
        _check(self.lib.hit3d_rng_state(ctypes.byref(iy)))
        return iy.value

    # -- timing -------------------------------------------------------------------
    def timing_enable(self, on: bool):
        _check(self.lib.hit3d_timing_enable(_i(1 if on else 0)))

    def timing_read(self) -> dict:
        buf = ctypes.create_string_buffer(512)
        _check(self.lib.hit3d_timing_names(buf, _i(512)))
        names = buf.value.decode().split(",")
        ms = np.zeros(len(names))
        cnt = np.zeros(len(names), np.int64)
        _check(self.lib.hit3d_timing_read(_i(len(names)), _dp(ms), cnt.ctypes.data_as(_I64P)))
        return {nm: (float(m), int(c)) for nm, m, c in zip(names, ms, cnt)}

    def synchronize(self):
        _check(self.lib.hit3d_synchronize())
