"""ctypes binding of ``libpyiid_hip.so`` (C ABI declared in include/pyiid_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
pyiid_amd/csrc``) into ``pyiid_amd/_lib/libpyiid_hip.so``.  There is no
fallback: if the library or a HIP device is missing, every entry point raises.
"""
import ctypes
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpyiid_hip.so")

PYIID_DEVICE = 1
PYIID_F64 = 2
SLOT_FQ = 0
SLOT_PDF = 1
POT_RW = 0
POT_CHI_SQ = 1
PRECISION_DEFAULT = 0
PRECISION_COMPENSATED = 1
GRAD_AUTO = 0
GRAD_ORDERED = 1
GRAD_SYMMETRIC = 2

KERNEL_FUSED = 0
KERNEL_FQ = 1
KERNEL_GEMM = 2
KERNEL_PAIRS = 3
KERNEL_PAIR_FORCE = 4

_ERRORS = {
    -1: ValueError,
    -2: RuntimeError,
    -3: RuntimeError,
    -4: MemoryError,
    -5: RuntimeError,
}

# (name, argtypes) for every entry point of include/pyiid_hip.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_SIGNATURES = [
    ("pyiid_abi_version", []),
    ("pyiid_device_count", [ctypes.POINTER(_I)]),
    ("pyiid_create", [_I, ctypes.POINTER(_P)]),
    ("pyiid_destroy", [_P]),
    ("pyiid_set_stream", [_P, _P]),
    ("pyiid_synchronize", [_P]),
    ("pyiid_enable_timing", [_P, _I]),
    ("pyiid_kernel_time", [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    ("pyiid_fq_form_counts", [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    ("pyiid_set_fq_bins_per_wave", [_P, _I]),
    ("pyiid_set_grad_bins_per_wave", [_P, _I]),
    ("pyiid_set_grad_tiles_per_item", [_P, _I]),
    ("pyiid_set_grad_sym_bins_per_wave", [_P, _I]),
    ("pyiid_set_grad_kernel", [_P, _I]),
    ("pyiid_grad_shape", [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_int)]),
    ("pyiid_set_autotune", [_P, _I]),
    ("pyiid_set_precision", [_P, _I]),
    ("pyiid_set_positions", [_P, _P, _I64, _I]),
    ("pyiid_set_elements", [_P, _P, _I64, _I]),
    ("pyiid_set_scatter", [_P, _I, _P, _I32, _I32, _I]),
    ("pyiid_fq", [_P, _I, _D, _P, _I]),
    ("pyiid_grad_fq", [_P, _I, _D, _P, _P, _I]),
    ("pyiid_pair_geometry", [_P, _I64, _I64, _P, _P, _I]),
    ("pyiid_normalization", [_P, _I, _P, _I]),
    ("pyiid_set_transform", [_P, _D, _D, _P, _I32, _D, _I32]),
    ("pyiid_pdf", [_P, _P, _P, _I]),
    ("pyiid_grad_pdf", [_P, _P, _I64, _P, _I]),
    ("pyiid_potential", [_P, _I, _P, _P, _I32, ctypes.POINTER(_D), ctypes.POINTER(_D), _I]),
    ("pyiid_potential_grad", [_P, _I, _P, _I64, _P, _P, _I32, _P, _I]),
    ("pyiid_pdf_energy_forces", [_P, _I, _D, _P, _I, _D, ctypes.POINTER(_D), _P, _I]),
    ("pyiid_fq_energy_forces", [_P, _I, _D, _I32, _P, _I, _D, ctypes.POINTER(_D), _P, _I]),
    ("pyiid_set_batch", [_P, _P, _I64, _P, _I32, _I]),
    ("pyiid_get_positions", [_P, _P, _I]),
    ("pyiid_hmc_setup", [_P, _I, _D, _P, _I, _D, _P, _I]),
    ("pyiid_hmc_run", [_P, _I, _I, _D, _D, ctypes.c_uint64, _I, _P, _P, _P, _I]),
    ("pyiid_sur_setup", [_P, _P, _P, _P, _P, _P, _P, _P, _I32]),
    ("pyiid_sur_set_output", [_P, _P, _I]),
    ("pyiid_sur_descriptors", [_P, _P, _I]),
    ("pyiid_sur_forces", [_P, _P, _P, _I]),
    ("pyiid_sur_features", [_P, _P, _I]),
    ("pyiid_sur_feature_energies", [_P, _P, _I]),
    ("pyiid_sur_get_output", [_P, _P, _I]),
    ("pyiid_sur_set_energy_weight", [_P, _D]),
    ("pyiid_sur_train", [_P, _D, _D, _P, _I]),
    ("pyiid_sur_hmc_run", [_P, _I, _I, _D, _D, ctypes.c_uint64, _I, _I, _D, _D, _D, _D, _P, _P, _P,
                           _P, _P, _I]),
    ("pyiid_fq_noalloc", [_I, _P, _P, _P, _I64, _P, _I32, _I32, _D, _P]),
]
EXPORTED = [name for name, _ in _SIGNATURES] + ["pyiid_last_error"]

_lib = None
_lock = threading.Lock()


def load(path=LIB_PATH):
    """Load the shared library (once) and declare every entry point."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                "libpyiid_hip.so is not built (%s); run __graft_entry__.build()" % path)
        # PyTorch-ROCm wheels bundle their own libamdhip64 / libhsa-runtime64
        # under the same sonames as /opt/rocm's.  Whichever is mapped first is
        # the one every later library binds to; two HSA runtimes in one process
        # leave torch without devices.  Map torch's first when it is present so
        # this library and torch share one runtime (and device pointers).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        for name, argtypes in _SIGNATURES:
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        lib.pyiid_last_error.argtypes = []
        lib.pyiid_last_error.restype = ctypes.c_char_p
        _lib = lib
        return lib


def check(rc):
    if rc != 0:
        msg = load().pyiid_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)("pyiid_hip error %d: %s" % (rc, msg))
    return rc


def device_count():
    n = ctypes.c_int(0)
    check(load().pyiid_device_count(ctypes.byref(n)))
    return n.value


def _ptr(a):
    """Host pointer of a contiguous numpy array, or a device pointer (int)."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


class Context:
    """One device context: a HIP stream plus the device-resident structure,
    scatter tables, transform and workspaces (``pyiid_ctx``).

    Launch shapes are autotuned per problem shape; the environment variable
    PYIID_AUTOTUNE=0 starts every context with fixed shapes instead
    (results that repeat bit for bit), as ``set_autotune(False)`` does."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.pyiid_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        if os.environ.get("PYIID_AUTOTUNE", "1") == "0":
            self.set_autotune(False)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.pyiid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- plumbing ---------------------------------------------------------
    def set_stream(self, stream_handle):
        check(self.lib.pyiid_set_stream(self.handle, stream_handle))

    def synchronize(self):
        check(self.lib.pyiid_synchronize(self.handle))

    def enable_timing(self, on=True):
        check(self.lib.pyiid_enable_timing(self.handle, 1 if on else 0))

    def set_autotune(self, on=True):
        """Autotune launch shapes per problem shape (default), or use fixed
        shapes (bit-reproducible results)."""
        check(self.lib.pyiid_set_autotune(self.handle, 1 if on else 0))

    def set_precision(self, compensated):
        """Compensated FP32 sums in the Debye kernels (True), or plain FP32
        runs of at most 2,048 j atoms with FP64 across runs (False, default)."""
        check(self.lib.pyiid_set_precision(self.handle, PRECISION_COMPENSATED if compensated
                                           else PRECISION_DEFAULT))

    def set_fq_bins_per_wave(self, m):
        """Fix the F-only kernel's Q bins per wave (64, 72, 80, 88, 96, 112), or 0 to autotune."""
        check(self.lib.pyiid_set_fq_bins_per_wave(self.handle, int(m)))

    def set_grad_bins_per_wave(self, m):
        """Fix the fused F + gradient kernel's Q bins per wave (16, 32), or 0 to autotune."""
        check(self.lib.pyiid_set_grad_bins_per_wave(self.handle, int(m)))

    def set_grad_tiles_per_item(self, k):
        """Fix the symmetric F + gradient kernel's tiles per item (4, 8, 16, 32), or 0 to autotune."""
        check(self.lib.pyiid_set_grad_tiles_per_item(self.handle, int(k)))

    def set_grad_sym_bins_per_wave(self, m):
        """Fix the symmetric F + gradient kernel's Q bins per wave (12, 16), or 0 to autotune."""
        check(self.lib.pyiid_set_grad_sym_bins_per_wave(self.handle, int(m)))

    def set_grad_kernel(self, kernel):
        """GRAD_AUTO, GRAD_ORDERED or GRAD_SYMMETRIC: the kernel of grad_fq."""
        check(self.lib.pyiid_set_grad_kernel(self.handle, int(kernel)))

    def grad_shape(self):
        """Launch shape of the last grad_fq: (symmetric, bins per wave, group), where group is
        the symmetric kernel's tiles per item or the ordered kernel's j ranges per tile."""
        a, b, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.pyiid_grad_shape(self.handle, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(g)))
        return bool(a.value), b.value, g.value

    def kernel_time(self, kind):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.pyiid_kernel_time(self.handle, kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def fq_form_counts(self):
        """(three-term, Reinsch) (wave, j atom) steps of the F-only kernel
        since timing was enabled."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.pyiid_fq_form_counts(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # --- structure ---------------------------------------------------------
    def set_positions(self, pos, flags=0):
        if flags & PYIID_DEVICE:
            n = pos.shape[0]
        else:
            pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
            n = pos.shape[0]
        check(self.lib.pyiid_set_positions(self.handle, _ptr(pos), n, flags))

    def set_batch(self, positions_list):
        """Load a batch of structures (list of (N_b, 3) arrays) back to back;
        returns the offsets (B + 1)."""
        sizes = [len(p) for p in positions_list]
        offsets = np.zeros(len(sizes) + 1, np.int64)
        offsets[1:] = np.cumsum(sizes)
        pos = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1, 3)
                                                   for p in positions_list]), np.float32)
        check(self.lib.pyiid_set_batch(self.handle, _ptr(pos), int(offsets[-1]), _ptr(offsets),
                                       len(sizes), 0))
        return offsets

    def get_positions(self, n):
        out = np.empty((n, 3), np.float32)
        check(self.lib.pyiid_get_positions(self.handle, _ptr(out), 0))
        return out

    def set_elements(self, elem, flags=0):
        if not flags & PYIID_DEVICE:
            elem = np.ascontiguousarray(elem, dtype=np.int32)
        check(self.lib.pyiid_set_elements(self.handle, _ptr(elem), elem.shape[0], flags))

    def set_scatter(self, slot, table, flags=0):
        if not flags & PYIID_DEVICE:
            table = np.ascontiguousarray(table, dtype=np.float32)
        e, q = table.shape
        check(self.lib.pyiid_set_scatter(self.handle, slot, _ptr(table), e, q, flags))

    # --- Debye -------------------------------------------------------------
    def fq(self, slot, qbin, nq, out=None, flags=0, n_chain=1):
        if out is None:
            out = np.empty(nq if n_chain == 1 else (n_chain, nq), np.float32)
        check(self.lib.pyiid_fq(self.handle, slot, float(qbin), _ptr(out), flags))
        return out

    def grad_fq(self, slot, qbin, n, nq, out=None, fq_out=None, f64=False, flags=0):
        if f64:
            flags |= PYIID_F64
        if out is None:
            out = np.empty((n, 3, nq), np.float64 if f64 else np.float32)
        check(self.lib.pyiid_grad_fq(self.handle, slot, float(qbin), _ptr(out), _ptr(fq_out),
                                     flags))
        return out

    def pair_geometry(self, k_offset, k_count, want_d=True, want_r=True):
        d = np.empty((k_count, 3), np.float32) if want_d else None
        r = np.empty(k_count, np.float32) if want_r else None
        check(self.lib.pyiid_pair_geometry(self.handle, int(k_offset), int(k_count), _ptr(d),
                                           _ptr(r), 0))
        return d, r

    def normalization(self, slot, nq, n_chain=1):
        out = np.empty(nq if n_chain == 1 else (n_chain, nq), np.float64)
        check(self.lib.pyiid_normalization(self.handle, slot, _ptr(out), 0))
        return out

    # --- transform ---------------------------------------------------------
    def set_transform(self, rstep, qstep, rgrid, qmin, nq):
        rgrid = np.ascontiguousarray(rgrid, dtype=np.float64)
        check(self.lib.pyiid_set_transform(self.handle, float(rstep), float(qstep), _ptr(rgrid),
                                           len(rgrid), float(qmin), int(nq)))

    def pdf(self, fq, n_r, out=None, flags=0):
        if not flags & PYIID_DEVICE:
            fq = np.ascontiguousarray(fq, dtype=np.float32)
        if out is None:
            out = np.empty(n_r, np.float64)
        check(self.lib.pyiid_pdf(self.handle, _ptr(fq), _ptr(out), flags))
        return out

    def grad_pdf(self, grad_fq, n_r, out=None, flags=0):
        if not flags & PYIID_DEVICE:
            if grad_fq.dtype == np.float64:
                flags |= PYIID_F64
                grad_fq = np.ascontiguousarray(grad_fq)
            else:
                grad_fq = np.ascontiguousarray(grad_fq, dtype=np.float32)
        rows = grad_fq.shape[0] * grad_fq.shape[1]
        if out is None:
            out = np.empty(tuple(grad_fq.shape[:-1]) + (n_r,), np.float64)
        check(self.lib.pyiid_grad_pdf(self.handle, _ptr(grad_fq), rows, _ptr(out), flags))
        return out

    # --- potentials --------------------------------------------------------
    def potential(self, potential, gobs, gcalc):
        gobs = np.ascontiguousarray(gobs, dtype=np.float64)
        gcalc = np.ascontiguousarray(gcalc, dtype=np.float64)
        v = ctypes.c_double()
        s = ctypes.c_double()
        check(self.lib.pyiid_potential(self.handle, potential, _ptr(gobs), _ptr(gcalc),
                                       len(gobs), ctypes.byref(v), ctypes.byref(s), 0))
        return v.value, s.value

    def potential_grad(self, potential, grad_gcalc, gobs, gcalc):
        gobs = np.ascontiguousarray(gobs, dtype=np.float64)
        gcalc = np.ascontiguousarray(gcalc, dtype=np.float64)
        gg = np.ascontiguousarray(grad_gcalc, dtype=np.float64)
        rows = int(np.prod(gg.shape[:-1]))
        out = np.empty(gg.shape[:-1], np.float64)
        check(self.lib.pyiid_potential_grad(self.handle, potential, _ptr(gg), rows, _ptr(gobs),
                                            _ptr(gcalc), len(gobs), _ptr(out), 0))
        return out

    def pdf_energy_forces(self, slot, qbin, gobs, potential, conv, n_atoms, forces=True):
        """Energy and reference-convention forces.  ``gobs`` is (n_r,) for one
        structure (returns a float energy) or (B, n_r) for a batch (returns B
        energies)."""
        gobs = np.ascontiguousarray(gobs, dtype=np.float64)
        n_chain = 1 if gobs.ndim == 1 else gobs.shape[0]
        e = np.zeros(n_chain, np.float64)
        f = np.empty((n_atoms, 3), np.float64) if forces else None
        check(self.lib.pyiid_pdf_energy_forces(
            self.handle, slot, float(qbin), _ptr(gobs), potential, float(conv),
            e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(f), 0))
        return (float(e[0]) if gobs.ndim == 1 else e), f

    def hmc_setup(self, slot, qbin, gobs, potential, conv, masses):
        gobs = np.ascontiguousarray(gobs, dtype=np.float64)
        masses = np.ascontiguousarray(masses, dtype=np.float64)
        check(self.lib.pyiid_hmc_setup(self.handle, slot, float(qbin), _ptr(gobs), potential,
                                       float(conv), _ptr(masses), 0))

    def hmc_run(self, n_proposals, n_leapfrog, step, beta, seed, n_chain, use_graph=True,
                times=False):
        e = np.empty((n_proposals, n_chain), np.float64)
        a = np.empty((n_proposals, n_chain), np.int32)
        t = np.empty(n_proposals, np.float64) if times else None
        check(self.lib.pyiid_hmc_run(self.handle, int(n_proposals), int(n_leapfrog), float(step),
                                     float(beta), int(seed) & (2 ** 64 - 1),
                                     1 if use_graph else 0, _ptr(e), _ptr(a), _ptr(t), 0))
        return (e, a, t) if times else (e, a)

    # ---- learned integrator ----
    def sur_setup(self, params, w1, b1, w2, b2, d_mean, d_inv_scale, max_nbr):
        f = [np.ascontiguousarray(x, dtype=np.float32) for x in
             (params, w1, b1, w2, b2, d_mean, d_inv_scale)]
        if f[0].size != 5 or f[1].shape != (16, 64) or f[2].size != 64 or f[3].shape != (64, 64) \
                or f[4].size != 64 or f[5].size != 16 or f[6].size != 16:
            raise ValueError("surrogate parameter shapes: params 5, w1 16x64, b1 64, w2 64x64, "
                             "b2 64, d_mean 16, d_inv_scale 16")
        check(self.lib.pyiid_sur_setup(self.handle, *[_ptr(x) for x in f], int(max_nbr)))

    def sur_set_output(self, w3):
        w3 = np.ascontiguousarray(w3, dtype=np.float32)
        check(self.lib.pyiid_sur_set_output(self.handle, _ptr(w3), 0))

    def sur_descriptors(self, n):
        d = np.empty((n, 16), np.float32)
        check(self.lib.pyiid_sur_descriptors(self.handle, _ptr(d), 0))
        return d

    def sur_forces(self, n, energies=False):
        f = np.empty((n, 3), np.float64)
        e = np.empty(n, np.float32) if energies else None
        check(self.lib.pyiid_sur_forces(self.handle, _ptr(f), _ptr(e), 0))
        return (f, e) if energies else f

    def sur_features(self, n):
        phi = np.empty((64, 3 * n), np.float32)
        check(self.lib.pyiid_sur_features(self.handle, _ptr(phi), 0))
        return phi

    def sur_feature_energies(self, n_chain):
        e = np.empty((n_chain, 64), np.float64)
        check(self.lib.pyiid_sur_feature_energies(self.handle, _ptr(e), 0))
        return e

    def sur_get_output(self, n_chain):
        w = np.empty((n_chain, 64), np.float32)
        check(self.lib.pyiid_sur_get_output(self.handle, _ptr(w), 0))
        return w

    def sur_set_energy_weight(self, weight):
        check(self.lib.pyiid_sur_set_energy_weight(self.handle, float(weight)))

    def sur_train(self, n_chain, forget=0.0, ridge=1e-6):
        r2 = np.empty(n_chain, np.float64)
        check(self.lib.pyiid_sur_train(self.handle, float(forget), float(ridge), _ptr(r2), 0))
        return r2

    def sur_hmc_run(self, n_proposals, n_leapfrog, step, beta, seed, n_chain, use_graph=True,
                    train_every=0, forget=0.5, ridge=1e-6, min_r2=-np.inf, min_acceptance=0.0):
        e = np.empty((n_proposals, n_chain), np.float64)
        a = np.empty((n_proposals, n_chain), np.int32)
        m = np.empty(n_proposals, np.int32)
        t = np.empty(n_proposals, np.float64)
        v = np.empty(n_chain, np.int32)
        check(self.lib.pyiid_sur_hmc_run(
            self.handle, int(n_proposals), int(n_leapfrog), float(step), float(beta),
            int(seed) & (2 ** 64 - 1), 1 if use_graph else 0, int(train_every), float(forget),
            float(ridge), float(min_r2), float(min_acceptance), _ptr(e), _ptr(a), _ptr(m), _ptr(t),
            _ptr(v), 0))
        return {"energies": e, "accepted": a, "exact_mode": m, "times_ms": t, "violations": v}

    def fq_energy_forces(self, slot, qbin, q_first, gobs, potential, conv, n_atoms, forces=True):
        gobs = np.ascontiguousarray(gobs, dtype=np.float64)
        e = ctypes.c_double()
        f = np.empty((n_atoms, 3), np.float64) if forces else None
        check(self.lib.pyiid_fq_energy_forces(self.handle, slot, float(qbin), int(q_first),
                                              _ptr(gobs), potential, float(conv),
                                              ctypes.byref(e), _ptr(f), 0))
        return e.value, f
