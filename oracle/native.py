"""ctypes loader for oracle/_build/libdebye_oracle.so (test infrastructure;
see debye_oracle.c)."""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build",
                     "libdebye_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError("oracle C library not built (make -C oracle)")
        lib = ctypes.CDLL(_PATH)
        lib.oracle_flat_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p]
        lib.oracle_flat_f32.restype = ctypes.c_int
        _lib = lib
    return _lib


def flat_f32(q, scatter, qbin, q_index=None, nthreads=1, grad=True):
    """(F at q_index, grad at q_index or None) with the reference's float32
    per-pair arithmetic (cpu_flat.py + flat_multi_cpu_wrap.py)."""
    q = np.ascontiguousarray(q, dtype=np.float32)
    f = np.ascontiguousarray(scatter, dtype=np.float32)
    n, nq = f.shape
    qi = np.arange(nq, dtype=np.int32) if q_index is None else \
        np.ascontiguousarray(q_index, dtype=np.int32)
    fq = np.empty(len(qi), np.float32)
    g = np.empty((n, 3, len(qi)), np.float32) if grad else None
    rc = load().oracle_flat_f32(q.ctypes.data, f.ctypes.data, n, nq, float(np.float32(qbin)),
                                qi.ctypes.data, len(qi), int(nthreads), fq.ctypes.data,
                                None if g is None else g.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_flat_f32 failed")
    return fq, g
