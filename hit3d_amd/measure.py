"""Measurement entry points of the library (include/hit3d_measure.h): the
HBM yardsticks the time step's kernels are set beside."""
import ctypes

from .lib import Hit3D, _check, _i, load

_SIGNATURES = {
    "hit3d_stream_copy": [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
}


def _bind():
    lib = load()
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def exported_symbols():
    return list(_SIGNATURES)


def stream_copy(h: Hit3D, reps: int, blocks: int):
    """Mean milliseconds of a streaming copy of one slot (wrk(2) = wrk(1)),
    and the bytes it reads + writes."""
    lib = _bind()
    ms = ctypes.c_double(0.0)
    byt = ctypes.c_double(0.0)
    _check(lib.hit3d_stream_copy(_i(reps), _i(blocks), ctypes.byref(ms), ctypes.byref(byt)))
    return ms.value, byt.value
