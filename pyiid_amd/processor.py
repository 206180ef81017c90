"""The 'MI355X' processor: drop-in replacements for the three callables that
``ElasticScatter.set_processor`` binds (elasticscatter/__init__.py:206-292):

* ``wrap_fq(atoms, qbin=.1, sum_type='fq')`` -> float32 F(Q) on the full grid
  from Q = 0 (flat_multi_cpu_wrap.py:22-60, gpu_wrap.py:119-156);
* ``wrap_fq_grad(atoms, qbin=.1, sum_type='fq')`` -> (N, 3, Q) gradient in the
  reference convention (-1/2 dF/dq), float64 like the GPU processor
  (gpu_wrap.py:159-194);
* ``grad_pdf(grad_fq, rstep, qstep, rgrid, qmin)`` -> (N, 3, R) float64
  (master_kernel.py:272-290, gpu_wrap.py:197-284); like the reference it zeroes
  ``grad_fq[..., :ceil(qmin/qstep)]`` in place.

All arithmetic runs in the gfx950 kernels of ``libpyiid_hip.so``; there is no
CPU fallback.  ``sum_type == 'fq'`` selects the atoms' 'F(Q) scatter' table,
anything else 'PDF scatter' (flat_multi_cpu_wrap.py:13-16).
"""
import contextlib
import math
import threading

import numpy as np

from . import _lib

_ctx = {}
_ctx_lock = threading.RLock()


class DeviceState:
    """A context plus host-side caches of what it holds."""

    def __init__(self, device=0):
        self.ctx = _lib.Context(device)
        self.lock = threading.RLock()
        self.elements = None
        self.tables = [None, None]
        self.transform_key = None

    def load_structure(self, positions, elements):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.ctx.set_positions(pos)
        elements = np.ascontiguousarray(elements, dtype=np.int32)
        if self.elements is None or len(self.elements) != len(elements) or \
                not np.array_equal(self.elements, elements):
            self.ctx.set_elements(elements)
            self.elements = elements.copy()

    def load_table(self, slot, table):
        table = np.ascontiguousarray(table, dtype=np.float32)
        cached = self.tables[slot]
        if cached is None or cached.shape != table.shape or not np.array_equal(cached, table):
            self.ctx.set_scatter(slot, table)
            self.tables[slot] = table.copy()

    def load_transform(self, rstep, qstep, rgrid, qmin, nq):
        rgrid = np.ascontiguousarray(rgrid, dtype=np.float64)
        key = (float(rstep), float(qstep), float(qmin), int(nq), len(rgrid),
               rgrid.tobytes() if len(rgrid) < 1 << 16 else hash(rgrid.tobytes()))
        if key != self.transform_key:
            self.ctx.set_transform(rstep, qstep, rgrid, qmin, nq)
            self.transform_key = key


def device_available():
    try:
        return _lib.device_count() > 0
    except (RuntimeError, OSError):
        return False


def state(device=0):
    """The process-wide device state of ``device`` (created on first use)."""
    with _ctx_lock:
        if device not in _ctx:
            _ctx[device] = DeviceState(device)
        return _ctx[device]


def compact_scatter(atoms, name):
    """(table E x Q, element index N) for the per-atom scatter array ``name``.

    Uses the compact form stored by ``ElasticScatter._wrap_atoms`` when the
    atoms still hold the very array it stored; otherwise deduplicates rows.
    """
    arr = atoms.arrays[name]
    cache = atoms.info.get("pyiid_compact", {}).get(name)
    if cache is not None and cache[2] is arr:
        return cache[0], cache[1]
    table, index = np.unique(np.asarray(arr, dtype=np.float32), axis=0, return_inverse=True)
    return table, index.reshape(-1).astype(np.int32)


def _host_out(shape, dtype):
    """A new host array for a result copied from the device.  With torch and a
    device present it is page-locked memory from torch's caching host
    allocator, so the device-to-host copy runs at DMA speed instead of through
    pageable staging (the 200 MB float64 gradient of a 6,699-atom particle);
    otherwise plain memory.  Every call returns a new array, as the reference's
    wrappers do."""
    try:
        import torch
        if torch.cuda.is_available():
            tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
            return torch.empty(shape, dtype=tdt, pin_memory=True).numpy()
    except (ImportError, RuntimeError):
        pass
    return np.empty(shape, dtype)


def _slot(sum_type):
    return _lib.SLOT_FQ if sum_type == "fq" else _lib.SLOT_PDF


@contextlib.contextmanager
def loaded(atoms, sum_type, device=0):
    """Uploads the structure and the scatter table of ``atoms`` to the shared
    context of ``device`` and yields (state, slot, n_q).  The context's lock
    is held from the upload to the end of the ``with`` block, so another
    thread cannot replace the structure before the caller's compute runs."""
    name = "F(Q) scatter" if sum_type == "fq" else "PDF scatter"
    table, index = compact_scatter(atoms, name)
    slot = _slot(sum_type)
    st = state(device)
    with st.lock:
        st.load_structure(atoms.get_positions(), index)
        st.load_table(slot, table)
        yield st, slot, table.shape[1]


def wrap_fq(atoms, qbin=.1, sum_type="fq"):
    """Reduced structure factor F(Q) (float32, floor(qmax/qbin) bins)."""
    with loaded(atoms, sum_type) as (st, slot, nq):
        return st.ctx.fq(slot, qbin, nq)


def wrap_fq_grad(atoms, qbin=.1, sum_type="fq"):
    """-1/2 dF/dq as an (N, 3, Q) float64 array (the reference convention)."""
    with loaded(atoms, sum_type) as (st, slot, nq):
        out = _host_out((len(atoms), 3, nq), np.float64)
        return st.ctx.grad_fq(slot, qbin, len(atoms), nq, out=out, f64=True)


def grad_pdf(grad_fq, rstep, qstep, rgrid, qmin):
    """grad G(r) = T . grad F(Q) row by row (FP32 MFMA GEMM on the device)."""
    grad_fq[..., :int(math.ceil(qmin / qstep))] = 0.0
    st = state()
    with st.lock:
        st.load_transform(rstep, qstep, rgrid, qmin, grad_fq.shape[-1])
        out = _host_out(tuple(grad_fq.shape[:-1]) + (len(rgrid),), np.float64)
        return st.ctx.grad_pdf(grad_fq, len(rgrid), out=out)


def pdf_at_qmin(fq, rstep, qstep, rgrid, qmin):
    """G(r) from F(Q) (``get_pdf_at_qmin``, master_kernel.py:39-104) on the device.
    Like the reference it zeroes ``fq[:ceil(qmin/qstep)]`` in place."""
    fq[:int(math.ceil(qmin / qstep))] = 0.0
    st = state()
    with st.lock:
        st.load_transform(rstep, qstep, rgrid, qmin, len(fq))
        return st.ctx.pdf(fq, len(rgrid))
