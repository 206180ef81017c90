"""ctypes loader for oracle/_build/libdebye_oracle.so (test infrastructure;
see debye_oracle.c)."""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build",
                     "libdebye_oracle.so")
_libs = {}


def load(path=_PATH):
    """The C restatement at ``path`` (default: the portable build)."""
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("oracle C library not built (make -C oracle): %s" % path)
        lib = ctypes.CDLL(path)
        lib.oracle_flat_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p]
        lib.oracle_flat_f32.restype = ctypes.c_int
        lib.oracle_sums_f64.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_int,
                                                                 ctypes.c_void_p, ctypes.c_int,
                                                                 ctypes.c_int, ctypes.c_void_p,
                                                                 ctypes.c_void_p]
        lib.oracle_sums_f64.restype = ctypes.c_int
        _libs[path] = lib
    return _libs[path]


def build_native(outdir):
    """Builds the C restatement with -march=native for this machine into
    ``outdir`` (the Makefile's ``native`` target) and returns its path, or
    None when the build fails."""
    import subprocess
    path = os.path.join(outdir, "libdebye_oracle_native.so")
    here = os.path.dirname(os.path.abspath(__file__))
    try:
        subprocess.run(["make", "-s", "-C", here, "native", "NATIVE=" + path], check=True,
                       capture_output=True, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return None
    return path if os.path.exists(path) else None


def flat_f32(q, scatter, qbin, q_index=None, nthreads=1, grad=True, lib=None):
    """(F at q_index, grad at q_index or None) with the reference's float32
    per-pair arithmetic (cpu_flat.py + flat_multi_cpu_wrap.py); ``lib``: a
    path to another build of the same source."""
    q = np.ascontiguousarray(q, dtype=np.float32)
    f = np.ascontiguousarray(scatter, dtype=np.float32)
    n, nq = f.shape
    qi = np.arange(nq, dtype=np.int32) if q_index is None else \
        np.ascontiguousarray(q_index, dtype=np.int32)
    fq = np.empty(len(qi), np.float32)
    g = np.empty((n, 3, len(qi)), np.float32) if grad else None
    rc = load(lib or _PATH).oracle_flat_f32(q.ctypes.data, f.ctypes.data, n, nq, float(np.float32(qbin)),
                                qi.ctypes.data, len(qi), int(nthreads), fq.ctypes.data,
                                None if g is None else g.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_flat_f32 failed")
    return fq, g


def fp64_subset(q, elem_index, table, qbin, q_index, rows=None, nthreads=1):
    """FP64 F(Q) and -1/2 dF/dq (oracle/scatter.py fq / grad_fq) on the bins
    ``q_index`` and gradient rows ``rows`` (None: no gradient), threaded C.
    ``table`` is the per-element form factor table (E x Qb), ``elem_index``
    the element row of every atom; positions are taken as float32 (the
    reference's cast)."""
    from . import scatter as S
    q = np.ascontiguousarray(np.asarray(q, np.float32), dtype=np.float64)
    elem = np.ascontiguousarray(elem_index, dtype=np.int32)
    qi = np.asarray(q_index)
    ft = np.asarray(table, np.float32).astype(np.float64)
    fsel = np.ascontiguousarray(ft[:, qi])
    Q = np.ascontiguousarray(S.q_values(qbin, ft.shape[1])[qi])
    n = len(q)
    rw = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    fsum = np.empty(len(qi))
    gsum = None if rw is None else np.empty((len(rw), 3, len(qi)))
    rc = load().oracle_sums_f64(q.ctypes.data, elem.ctypes.data, fsel.ctypes.data, Q.ctypes.data,
                                n, len(qi), None if rw is None else rw.ctypes.data,
                                0 if rw is None else len(rw), int(nthreads), fsum.ctypes.data,
                                None if gsum is None else gsum.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_sums_f64 failed")
    # na = ((sum_e c_e f_e)^2 - sum_e c_e f_e^2) / (N - 1)  (closed form of N * mean f_i f_j)
    counts = np.bincount(elem, minlength=ft.shape[0]).astype(np.float64)
    s1 = counts @ fsel
    s2 = counts @ (fsel * fsel)
    with np.errstate(divide="ignore", invalid="ignore"):
        na = (s1 * s1 - s2) / (n - 1) if n > 1 else np.zeros_like(s1)
        fq = np.nan_to_num(2.0 * fsum / na, nan=0.0, posinf=0.0, neginf=0.0)
        grad = None if gsum is None else np.nan_to_num(-gsum / na, nan=0.0, posinf=0.0,
                                                        neginf=0.0)
    return fq, grad
