"""pyiid_fq_noalloc: F(Q) whose intermediates live in a static device array.

The entry point is called through the C ABI, on device buffers and torch's
stream.  Its output is checked against pyiid_fq (the context path) on the
same inputs, and at a subset of bins against the FP64 oracle.  The tolerance
is max|x - x*| / max|x*| <= 1e-5, the bound of every F(Q) path.

Two calls run in a row on one device, the second with a smaller input.
  * The first call is an 8,000-atom Pt/Ni sphere on 4,096 bins.  Its 2,080
    work items fill three passes of partial sums (1,024 items each).
  * The second call is a 55-atom Au cuboctahedron on 250 bins: fewer atoms,
    items and bins, with one element (the unweighted kernels).
What the first call left in the array must not reach the second.  A third
call gives an element index outside the table and must return NaN in every
bin.  The host-side limits are checked without a device.
"""
import ctypes

import numpy as np
import pytest

from oracle import native
from pyiid_amd import _lib, structures
from pyiid_amd.formfactors import element_table

TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def _noalloc(lib, stream, pos_d, elem_d, table_d, nq, qbin, out_d, n=None):
    return lib.pyiid_fq_noalloc(0, ctypes.c_void_p(stream), ctypes.c_void_p(pos_d.data_ptr()),
                                None if elem_d is None else ctypes.c_void_p(elem_d.data_ptr()),
                                pos_d.shape[0] if n is None else n,
                                ctypes.c_void_p(table_d.data_ptr()), table_d.shape[0], nq,
                                float(qbin), ctypes.c_void_p(out_d.data_ptr()))


def _cases():
    big = structures.jitter(structures.fcc_sphere(30.0, a=3.80), 0.05, seed=11)
    big = big[:8000].astype(np.float32)
    alloy = structures.random_alloy(len(big), 78, 28, seed=12)
    small = structures.jitter(structures.cuboctahedron(4.08), 0.05, seed=0).astype(np.float32)
    return [("ptni8000_q4096", big, alloy, 0.01, 4096),
            ("au%d_q250" % len(small), small, np.full(len(small), 79), 0.1, 250)]


@pytest.mark.gpu
def test_noalloc_matches_context_path_on_two_calls(fixed_shapes):
    import torch
    lib = _lib.load()
    ctx = _lib.Context(0)
    stream = torch.cuda.current_stream().cuda_stream
    sizes = []
    try:
        for label, pos, numbers, qbin, nq in _cases():
            print("case", label, flush=True)
            table, index, _ = element_table(numbers, qbin, nq)
            ctx.set_positions(pos)
            ctx.set_elements(index)
            ctx.set_scatter(_lib.SLOT_FQ, table)
            want = ctx.fq(_lib.SLOT_FQ, qbin, nq)
            pos_d = torch.from_numpy(pos).cuda()
            elem_d = torch.from_numpy(index.astype(np.int32)).cuda()
            table_d = torch.from_numpy(table).cuda()
            out_d = torch.full((nq,), float("nan"), dtype=torch.float32, device="cuda")
            _lib.check(_noalloc(lib, stream, pos_d, elem_d, table_d, nq, qbin, out_d))
            torch.cuda.synchronize()
            got = out_d.cpu().numpy()
            qi = np.unique(np.linspace(1, nq - 1, 24).astype(int))
            ref, _ = native.fp64_subset(pos, index, table, qbin, qi, nthreads=8)
            print("%s: vs pyiid_fq %.2e, vs oracle %.2e" % (label, rel(got, want),
                                                             rel(got[qi], ref)))
            assert np.all(np.isfinite(got))
            assert rel(got, want) < TOL
            assert rel(got[qi], ref) < TOL
            sizes.append(len(pos))
    finally:
        ctx.close()
    assert sizes[1] < sizes[0]


@pytest.mark.gpu
def test_noalloc_out_of_range_element_gives_nan():
    import torch
    lib = _lib.load()
    pos = structures.jitter(structures.cuboctahedron(4.08), 0.05, seed=1).astype(np.float32)
    table, index, _ = element_table(np.r_[np.full(len(pos) - 1, 78), 28], 0.1, 250)
    index = index.astype(np.int32)
    index[5] = table.shape[0]  # one row past the table
    pos_d = torch.from_numpy(pos).cuda()
    elem_d = torch.from_numpy(index).cuda()
    table_d = torch.from_numpy(table).cuda()
    out_d = torch.zeros(250, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_noalloc(lib, stream, pos_d, elem_d, table_d, 250, 0.1, out_d))
    torch.cuda.synchronize()
    assert np.all(np.isnan(out_d.cpu().numpy()))


def test_noalloc_rejects_sizes_beyond_its_limits():
    lib = _lib.load()
    buf = ctypes.create_string_buffer(16)
    p = ctypes.cast(buf, ctypes.c_void_p)
    for n, n_elem, nq, qbin in [(65537, 1, 10, 0.1), (10, 33, 10, 0.1), (10, 1, 4097, 0.1),
                                (-1, 1, 10, 0.1), (10, 1, 10, 0.0)]:
        rc = lib.pyiid_fq_noalloc(0, None, p, None, n, p, n_elem, nq, qbin, p)
        assert rc == -1, (n, n_elem, nq, qbin)
        assert b"limit" in lib.pyiid_last_error() or qbin == 0.0


@pytest.mark.gpu
def test_noalloc_with_no_pair_gives_zero():
    """No atoms, or one: F is 0 in every bin (no pair, na = 0 -> nan_to_num)."""
    import torch
    lib = _lib.load()
    table, _, _ = element_table(np.array([79]), 0.1, 250)
    table_d = torch.from_numpy(table).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    for n in (0, 1):
        pos_d = torch.zeros((n, 3), dtype=torch.float32, device="cuda") + 1.5
        out_d = torch.full((250,), float("nan"), dtype=torch.float32, device="cuda")
        _lib.check(lib.pyiid_fq_noalloc(0, ctypes.c_void_p(stream),
                                        ctypes.c_void_p(pos_d.data_ptr() if n else 0), None, n,
                                        ctypes.c_void_p(table_d.data_ptr()), 1, 250, 0.1,
                                        ctypes.c_void_p(out_d.data_ptr())))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_d.cpu().numpy(), np.zeros(250, np.float32))
