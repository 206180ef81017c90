"""GPU parity of the Debye kernels against the FP64 oracle (through the C ABI).

Tolerance (from the target of this path): max|x - x*| / max|x*| <= 1e-5 for
F(Q) and the gradient, with float32 positions on both sides.
"""
import numpy as np
import pytest

from oracle import scatter as oracle
from pyiid_amd import _lib, structures
from pyiid_amd.formfactors import element_table

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def load(ctx, pos, numbers, qbin, nq, slot=_lib.SLOT_FQ):
    table, index, _ = element_table(numbers, qbin, nq)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(slot, table)
    return table[index]


def cubo():
    return structures.jitter(structures.cuboctahedron(4.08), 0.05, seed=0).astype(np.float32)


@pytest.mark.parametrize("qbin,nq", [(0.1, 250), (0.0756865, 330), (0.02, 1250)])
def test_fq_cuboctahedron(ctx, qbin, nq):
    pos = cubo()
    per_atom = load(ctx, pos, np.full(len(pos), 79), qbin, nq)
    got = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    want = oracle.fq(pos, per_atom, qbin)
    print("fq N=%d nq=%d rel %.2e" % (len(pos), nq, rel(got, want)))
    assert rel(got, want) < TOL


@pytest.mark.parametrize("qbin,nq", [(0.1, 250), (0.02, 1250)])
def test_grad_fq_cuboctahedron(ctx, qbin, nq):
    pos = cubo()
    per_atom = load(ctx, pos, np.full(len(pos), 79), qbin, nq)
    fq_out = np.empty(nq, np.float32)
    got = ctx.grad_fq(_lib.SLOT_FQ, qbin, len(pos), nq, fq_out=fq_out, f64=True)
    want = oracle.grad_fq(pos, per_atom, qbin)
    print("grad N=%d nq=%d rel %.2e" % (len(pos), nq, rel(got, want)))
    assert rel(got, want) < TOL
    assert rel(fq_out, oracle.fq(pos, per_atom, qbin)) < TOL


def test_sphere_and_float32_output(ctx):
    pos = structures.jitter(structures.fcc_sphere(9.0), 0.05, seed=3).astype(np.float32)
    n, qbin, nq = len(pos), 0.02, 1250
    per_atom = load(ctx, pos, np.full(n, 79), qbin, nq)
    fq = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    qi = np.arange(0, nq, 7)
    assert rel(fq[qi], oracle.fq(pos, per_atom, qbin, q_index=qi)) < TOL
    g32 = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq)
    assert g32.dtype == np.float32
    rows = np.arange(0, n, 5)
    want = oracle.grad_fq(pos, per_atom, qbin, q_index=qi, atom_index=rows)
    assert rel(g32[rows][:, :, qi], want) < TOL


def test_alloy_weighted(ctx):
    base = structures.fcc_nearest(300, a=3.724)
    pos = structures.jitter(base, 0.05, seed=5).astype(np.float32)
    numbers = structures.random_alloy(len(pos), 78, 28, seed=7)
    qbin, nq = 0.05, 500
    per_atom = load(ctx, pos, numbers, qbin, nq)
    fq = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    assert rel(fq, oracle.fq(pos, per_atom, qbin)) < TOL
    g = ctx.grad_fq(_lib.SLOT_FQ, qbin, len(pos), nq, f64=True)
    assert rel(g, oracle.grad_fq(pos, per_atom, qbin)) < TOL
    na = ctx.normalization(_lib.SLOT_FQ, nq)
    np.testing.assert_allclose(na, oracle.normalization(per_atom), rtol=1e-12)


def test_pair_geometry(ctx):
    pos = cubo()
    load(ctx, pos, np.full(len(pos), 79), 0.1, 250)
    k = len(pos) * (len(pos) - 1) // 2
    d, r = ctx.pair_geometry(0, k)
    dw, rw = oracle.pair_geometry(pos)
    np.testing.assert_allclose(d, dw, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r, rw, rtol=1e-6)
    d2, r2 = ctx.pair_geometry(100, 200)
    np.testing.assert_array_equal(d2, d[100:300])
    np.testing.assert_array_equal(r2, r[100:300])


def test_pair_geometry_ranges_and_alignment(ctx):
    """Pair ranges of several lengths whose ends cross rows of the triangle,
    in host mode and into device buffers at 16-byte aligned and unaligned
    addresses: all equal to the full array, which matches the oracle."""
    import torch
    pos = structures.jitter(structures.fcc_nearest(97, a=4.08), 0.05, seed=9).astype(np.float32)
    load(ctx, pos, np.full(len(pos), 79), 0.1, 250)
    k = len(pos) * (len(pos) - 1) // 2
    d, r = ctx.pair_geometry(0, k)
    dw, rw = oracle.pair_geometry(pos)
    np.testing.assert_allclose(d, dw, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r, rw, rtol=1e-6)
    for off, cnt in [(0, 1), (1, 2), (5, 3), (44, 4), (45, 5), (1000, 1001), (k - 7, 7), (k - 1, 1)]:
        d2, r2 = ctx.pair_geometry(off, cnt)
        np.testing.assert_array_equal(d2, d[off:off + cnt])
        np.testing.assert_array_equal(r2, r[off:off + cnt])
    for shift in (0, 1):  # floats: 0 keeps 16-byte alignment, 1 breaks it
        dd = torch.zeros(3 * 1001 + 4, dtype=torch.float32, device="cuda")
        rr = torch.zeros(1001 + 4, dtype=torch.float32, device="cuda")
        _lib.check(ctx.lib.pyiid_pair_geometry(ctx.handle, 1000, 1001, dd[shift:].data_ptr(),
                                               rr[shift:].data_ptr(), _lib.PYIID_DEVICE))
        ctx.synchronize()
        np.testing.assert_array_equal(dd[shift:shift + 3 * 1001].cpu().numpy().reshape(-1, 3),
                                      d[1000:2001])
        np.testing.assert_array_equal(rr[shift:shift + 1001].cpu().numpy(), r[1000:2001])


@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_structures(ctx, n):
    pos = np.array([[0, 0, 0], [2.5, 0.3, 0.1]], np.float32)[:n]
    qbin, nq = 0.1, 250
    per_atom = load(ctx, pos, np.full(n, 79), qbin, nq)
    fq = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    g = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True)
    if n < 2:
        assert not np.any(fq) and not np.any(g)
    else:
        r = np.linalg.norm(pos[1].astype(np.float64) - pos[0])
        Q = oracle.q_values(qbin, nq)
        assert rel(fq, np.sin(Q * r) / r) < TOL  # dimer closed form
        assert rel(g, oracle.grad_fq(pos, per_atom, qbin)) < TOL


@pytest.mark.parametrize("k,off", [(1, 5e-4), (1, -3e-3), (1, 0.02), (1, 0.3), (2, 1e-3),
                                   (3, -0.01)])
def test_steps_near_multiples_of_pi(ctx, k, off):
    """A square of side s with qbin * s near k pi, plus an atom off its plane.
    The three-term recurrence amplifies rounding by ~min(m, 1 / |sin(qbin r)|)
    after m bins, near every multiple of pi as near 0, so such pairs take the
    accurate forms (F-only kernel: the Reinsch difference form; contracted
    forces: the rotation).  F, the gradient and the F(Q)-Rw forces (target:
    the structure shrunk by 0.1 %) against the oracle; the forces of this
    configuration move by ~1e-6 for 1e-6 noise in F or its gradient."""
    qbin, nq, q_first = 0.1, 250, 10
    side = (k * np.pi + off) / qbin
    pos = np.array([[0, 0, 0], [side, 0, 0], [0, side, 0], [side, side, 0],
                    [0.37 * side, 0.61 * side, 2.5]], np.float32)
    n = len(pos)
    per_atom = load(ctx, pos, np.full(n, 79), qbin, nq)
    f_want = oracle.fq(pos, per_atom, qbin)
    fq = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    g_want = oracle.grad_fq(pos, per_atom, qbin)
    fq_fused = np.empty(nq, np.float32)
    g = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True, fq_out=fq_fused)
    target = oracle.fq((pos * np.float32(0.999)).astype(np.float32), per_atom, qbin)[q_first:]
    e, f = ctx.fq_energy_forces(_lib.SLOT_FQ, qbin, q_first, target, _lib.POT_RW, 1.0, n)
    f_want_forces = oracle.wrap_grad_rw(g_want[:, :, q_first:], f_want[q_first:], target)
    print("k=%d off=%g: F %.2e F(fused) %.2e grad %.2e forces %.2e" % (
        k, off, rel(fq, f_want), rel(fq_fused, f_want), rel(g, g_want), rel(f, f_want_forces)))
    assert rel(fq, f_want) < TOL and rel(fq_fused, f_want) < TOL
    assert rel(g, g_want) < TOL
    assert abs(e - oracle.get_rw(target, f_want[q_first:])[0]) < 1e-6
    assert rel(f, f_want_forces) < TOL


@pytest.mark.parametrize("qbin,nq", [(0.03, 834), (0.003, 8334)])
def test_contracted_forces_both_kernels(ctx, qbin, nq):
    """F(Q)-Rw energy and forces of a Pt/Ni alloy through the unordered-pair
    contracted-force kernel: at 834 bins in one Q group, at 8,334 bins in
    several (a wave's u and u Q in LDS hold at most 2,048 bins).  Both
    against the oracle."""
    base = structures.fcc_nearest(43, a=3.724)
    pos = structures.jitter(base, 0.05, seed=4).astype(np.float32)
    numbers = structures.random_alloy(len(pos), 78, 28, seed=6)
    n, q_first = len(pos), 20
    per_atom = load(ctx, pos, numbers, qbin, nq)
    target = oracle.fq((base * np.float32(1.002)).astype(np.float32), per_atom, qbin)[q_first:]
    e, f = ctx.fq_energy_forces(_lib.SLOT_FQ, qbin, q_first, target, _lib.POT_RW, 10.0, n)
    f_want_fq = oracle.fq(pos, per_atom, qbin)
    g_want = oracle.grad_fq(pos, per_atom, qbin)
    f_want = 10.0 * oracle.wrap_grad_rw(g_want[:, :, q_first:], f_want_fq[q_first:], target)
    e_want = 10.0 * oracle.get_rw(target, f_want_fq[q_first:])[0]
    print("nq=%d: energy %.3e, forces %.2e" % (nq, abs(e - e_want), rel(f, f_want)))
    assert abs(e - e_want) < 1e-5
    assert rel(f, f_want) < TOL


@pytest.mark.parametrize("nq", [96, 107, 120, 600, 700])
def test_fq_bin_counts(ctx, nq):
    """The F-only kernel's Q tilings: 6 waves of 48 (96 bins), 56 (107, 600)
    or 64 (120, 700) bins each, with one and with two Q groups per block
    row; a single element and a Pt/Ni alloy, against the oracle."""
    base = structures.fcc_nearest(150, a=3.9)
    pos = structures.jitter(base, 0.05, seed=nq).astype(np.float32)
    qbin = 0.05
    for numbers in (np.full(len(pos), 79), structures.random_alloy(len(pos), 78, 28, seed=nq)):
        per_atom = load(ctx, pos, numbers, qbin, nq)
        assert rel(ctx.fq(_lib.SLOT_FQ, qbin, nq), oracle.fq(pos, per_atom, qbin)) < TOL


@pytest.mark.parametrize("m", [64, 72, 80, 88, 96, 112])
def test_fq_every_bins_per_wave(ctx, m):
    """Each launch shape of the F-only kernel (m bins per wave, fixed instead
    of autotuned) against the oracle: a batch of Au chains at 330 bins (one Q
    group for m >= 88, two below) and a Pt/Ni alloy at 700 bins."""
    try:
        ctx.set_fq_bins_per_wave(m)
        chains = [structures.jitter(structures.fcc_nearest(n, a=4.08), 0.05, seed=n).astype(np.float32)
                  for n in (90, 61, 2)]
        qbin, nq = 0.0756865, 330
        table, index, _ = element_table(np.full(sum(len(c) for c in chains), 79), qbin, nq)
        ctx.set_batch(chains)
        ctx.set_elements(index)
        ctx.set_scatter(_lib.SLOT_FQ, table)
        fq = ctx.fq(_lib.SLOT_FQ, qbin, nq, n_chain=len(chains))
        for b, p in enumerate(chains):
            assert rel(fq[b], oracle.fq(p, table[np.zeros(len(p), int)], qbin)) < TOL
        pos = structures.jitter(structures.fcc_nearest(170, a=3.724), 0.05, seed=m).astype(np.float32)
        numbers = structures.random_alloy(len(pos), 78, 28, seed=m)
        per_atom = load(ctx, pos, numbers, 0.05, 700)
        assert rel(ctx.fq(_lib.SLOT_FQ, 0.05, 700), oracle.fq(pos, per_atom, 0.05)) < TOL
    finally:
        ctx.set_fq_bins_per_wave(0)


def test_fused_every_bins_per_wave(ctx):
    """Each launch shape of the fused F + gradient kernel (m bins per wave:
    32 in blocks of 8 waves, 16 in blocks of 16 waves), fixed instead of
    autotuned, against the oracle: 203 atoms (the 32-bin shape's rounds of 16
    j atoms end in an odd count, so the two-register-set loop runs its tail;
    the 16-bin shape's rounds of 8 split the tile at another place) at 330
    bins, and a Pt/Ni alloy of 170 atoms at 700 bins.  The two shapes anchor
    the recurrences at different bins, so their float32 results must differ
    in the last bits: equal outputs would mean the setting did not reach the
    launch."""
    cases = ((203, 330, 0.0756865, False), (170, 700, 0.05, True))
    out = {}
    try:
        for m in (16, 32):
            ctx.set_grad_bins_per_wave(m)
            for n, nq, qbin, alloy in cases:
                pos = structures.jitter(structures.fcc_nearest(n, a=3.9), 0.05,
                                        seed=n).astype(np.float32)
                numbers = structures.random_alloy(n, 78, 28, seed=n) if alloy else np.full(n, 79)
                per_atom = load(ctx, pos, numbers, qbin, nq)
                fq_out = np.empty(nq, np.float32)
                got = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, fq_out=fq_out)
                assert ctx.grad_shape()[:2] == (False, m)
                eg = rel(got, oracle.grad_fq(pos, per_atom, qbin))
                ef = rel(fq_out, oracle.fq(pos, per_atom, qbin))
                print("m=%d n=%d: grad %.2e F %.2e" % (m, n, eg, ef))
                assert eg < TOL and ef < TOL
                out[(m, n)] = got
        for n in (203, 170):
            assert not np.array_equal(out[(16, n)], out[(32, n)])
        with pytest.raises(ValueError, match="bins per wave"):
            ctx.set_grad_bins_per_wave(24)
    finally:
        ctx.set_grad_bins_per_wave(0)


def test_symmetric_every_group_size(ctx):
    """The symmetric (unordered-pair) F + gradient kernel at each grouping of
    64-atom tiles per work item (4, 8, 16, 32), fixed instead of autotuned,
    against the oracle:
    - 2 and 3 atoms: one tile of one or three pairs;
    - 50 and 64 atoms: one tile, the diagonal steps only (a partial and a
      full tile);
    - 65 atoms: a second tile of one atom;
    - 203 atoms at 330 bins (not a multiple of the 128 bins of a block);
    - 1,100 atoms at 200 bins: 18 tiles, so groups of 4 and 8 give items whose
      J-row slots take the adds of several I tiles;
    - a Pt/Ni alloy of 400 atoms at 300 bins: tiles and J ranges end at the
      element boundary and every slot takes f_I(Q) f_J(Q);
    - an Au/Pt/Ni particle of 300 atoms (uneven thirds, 70 Ni) at 260 bins:
      two element boundaries, one of them inside the second tile.
    Groupings sum in different orders, so the 1,100-atom gradients must
    differ in the last bits between 4 and 32 (equal outputs would mean the
    setting did not reach the launch).  The 12-bin, 12-wave shape runs every
    case at 8 tiles per item."""
    cases = ((2, 100, 0.1, False), (3, 100, 0.1, False), (50, 250, 0.1, False),
             (64, 250, 0.1, False), (65, 250, 0.1, False), (203, 330, 0.0756865, False),
             (1100, 200, 0.1, False), (400, 300, 0.08, True), (300, 260, 0.09, "three"))
    out = {}
    try:
        for k, m in ((4, 16), (8, 16), (16, 16), (32, 16), (8, 12)):
            ctx.set_grad_tiles_per_item(k)
            ctx.set_grad_sym_bins_per_wave(m)
            for n, nq, qbin, alloy in cases:
                pos = structures.jitter(structures.fcc_nearest(n, a=3.9), 0.05,
                                        seed=n).astype(np.float32)
                if alloy == "three":
                    numbers = np.array([79] * 120 + [78] * 110 + [28] * 70)[
                        np.random.RandomState(n).permutation(n)]
                elif alloy:
                    numbers = structures.random_alloy(n, 78, 28, seed=n)
                else:
                    numbers = np.full(n, 79)
                per_atom = load(ctx, pos, numbers, qbin, nq)
                fq_out = np.empty(nq, np.float32)
                got = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, fq_out=fq_out, f64=True)
                assert ctx.grad_shape() == (True, m, k)
                rows = np.arange(0, n, 7) if n > 500 else np.arange(n)
                want = oracle.grad_fq(pos, per_atom, qbin, atom_index=rows)
                eg = rel(got[rows], want)
                ef = rel(fq_out, oracle.fq(pos, per_atom, qbin))
                print("k=%d m=%d n=%d nq=%d: grad %.2e F %.2e" % (k, m, n, nq, eg, ef), flush=True)
                assert eg < TOL and ef < TOL
                out[(k, m, n)] = got
        assert not np.array_equal(out[(4, 16, 1100)], out[(32, 16, 1100)])
        with pytest.raises(ValueError, match="tiles per item"):
            ctx.set_grad_tiles_per_item(12)
    finally:
        ctx.set_grad_tiles_per_item(0)
        ctx.set_grad_sym_bins_per_wave(0)


def test_grad_kernel_choice(ctx):
    """set_grad_kernel restricts the autotuned F + gradient path to one kernel,
    which grad_shape reports; both match the oracle and each other to FP32
    rounding (a Pt/Ni alloy of 700 atoms at 400 bins)."""
    n, nq, qbin = 700, 400, 0.06
    pos = structures.jitter(structures.fcc_nearest(n, a=3.724), 0.05, seed=41).astype(np.float32)
    numbers = structures.random_alloy(n, 78, 28, seed=41)
    per_atom = load(ctx, pos, numbers, qbin, nq)
    rows = np.arange(0, n, 9)
    want = oracle.grad_fq(pos, per_atom, qbin, atom_index=rows)
    got = {}
    try:
        for kernel, sym in ((_lib.GRAD_ORDERED, False), (_lib.GRAD_SYMMETRIC, True)):
            ctx.set_grad_kernel(kernel)
            got[sym] = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True)
            assert ctx.grad_shape()[0] is sym
            assert rel(got[sym][rows], want) < TOL
        assert rel(got[True], got[False]) < 2e-6
        with pytest.raises(ValueError, match="kernel"):
            ctx.set_grad_kernel(3)
    finally:
        ctx.set_grad_kernel(_lib.GRAD_AUTO)


def test_fq_bins_per_wave_rejects_other_values(ctx):
    with pytest.raises(ValueError, match="bins per wave"):
        ctx.set_fq_bins_per_wave(70)


def test_fixed_launch_shapes_repeat_bit_for_bit():
    """With the autotuner off, two fresh contexts give the same F, gradient,
    energy and forces bit for bit; with it on, the same to FP32 rounding."""
    base = structures.fcc_nearest(300, a=3.724)
    pos = structures.jitter(base, 0.05, seed=21).astype(np.float32)
    numbers = structures.random_alloy(len(pos), 78, 28, seed=21)
    qbin, nq, q_first = 0.05, 500, 10
    table, index, _ = element_table(numbers, qbin, nq)
    target = oracle.fq((base * np.float32(1.002)).astype(np.float32), table[index], qbin)[q_first:]

    def run(autotune):
        c = _lib.Context(0)
        try:
            c.set_autotune(autotune)
            load(c, pos, numbers, qbin, nq)
            f = c.fq(_lib.SLOT_FQ, qbin, nq)
            g = c.grad_fq(_lib.SLOT_FQ, qbin, len(pos), nq, f64=True)
            e, fo = c.fq_energy_forces(_lib.SLOT_FQ, qbin, q_first, target, _lib.POT_RW, 10.0,
                                       len(pos))
            return f, g, e, fo
        finally:
            c.close()

    a, b, t = run(False), run(False), run(True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert rel(a[0], oracle.fq(pos, table[index], qbin)) < TOL
    assert rel(t[0], a[0]) < 1e-6 and rel(t[1], a[1]) < 1e-6
    assert abs(t[2] - a[2]) < 1e-6 and rel(t[3], a[3]) < 1e-5


def test_environment_switch_fixes_launch_shapes(monkeypatch):
    """PYIID_AUTOTUNE=0 gives every new context the fixed shapes: F equals,
    bit for bit, that of a context switched off by set_autotune(False)."""
    pos = structures.jitter(structures.fcc_nearest(180, a=4.08), 0.05, seed=31).astype(np.float32)
    numbers = np.full(len(pos), 79)

    def fq(env, switch):
        if env:
            monkeypatch.setenv("PYIID_AUTOTUNE", "0")
        else:
            monkeypatch.delenv("PYIID_AUTOTUNE", raising=False)
        c = _lib.Context(0)
        try:
            if switch:
                c.set_autotune(False)
            load(c, pos, numbers, 0.05, 600)
            return c.fq(_lib.SLOT_FQ, 0.05, 600)
        finally:
            c.close()

    np.testing.assert_array_equal(fq(True, False), fq(False, True))


def test_ten_elements_and_mixed_batch(ctx):
    """All ten tabulated elements in one particle, in random order: F, the
    FP64 gradient and the F(Q)-Rw forces (tiles of one element each, up to
    ten tile pairs' worth of element weights), then a batch of chains of
    different compositions (one element, two, ten, a single atom) through
    the F-only kernel, each chain against the oracle."""
    zs = np.array([6, 8, 14, 26, 28, 29, 46, 47, 78, 79])
    rs = np.random.RandomState(11)
    base = structures.fcc_nearest(201, a=3.9)
    pos = structures.jitter(base, 0.05, seed=12).astype(np.float32)
    numbers = zs[rs.randint(0, len(zs), len(pos))]
    qbin, nq, q_first = 0.02, 1700, 30
    per_atom = load(ctx, pos, numbers, qbin, nq)
    f_want = oracle.fq(pos, per_atom, qbin)
    assert rel(ctx.fq(_lib.SLOT_FQ, qbin, nq), f_want) < TOL
    g_want = oracle.grad_fq(pos, per_atom, qbin)
    assert rel(ctx.grad_fq(_lib.SLOT_FQ, qbin, len(pos), nq, f64=True), g_want) < TOL
    target = oracle.fq((base * np.float32(1.002)).astype(np.float32), per_atom, qbin)[q_first:]
    e, f = ctx.fq_energy_forces(_lib.SLOT_FQ, qbin, q_first, target, _lib.POT_RW, 10.0, len(pos))
    assert abs(e - 10.0 * oracle.get_rw(target, f_want[q_first:])[0]) < 1e-5
    assert rel(f, 10.0 * oracle.wrap_grad_rw(g_want[:, :, q_first:], f_want[q_first:], target)) < TOL

    chains = [structures.jitter(structures.fcc_nearest(n, a=3.9), 0.05, seed=20 + n).astype(np.float32)
              for n in (30, 41, 57, 1)]
    zc = [np.full(30, 79), structures.random_alloy(41, 78, 28, seed=3),
          zs[rs.randint(0, len(zs), 57)], np.array([29])]
    allz = np.concatenate(zc)
    table, index, elems = element_table(allz, qbin, nq)
    ctx.set_batch(chains)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    fq = ctx.fq(_lib.SLOT_FQ, qbin, nq, n_chain=len(chains))
    for b, (p, z) in enumerate(zip(chains, zc)):
        want = oracle.fq(p, table[np.searchsorted(elems, z)], qbin)
        if len(p) < 2:
            assert np.all(fq[b] == 0.0)
        else:
            assert rel(fq[b], want) < TOL


def test_device_pointer_mode_matches_host_mode(ctx):
    """PYIID_DEVICE (inputs and outputs as device pointers, ordered on the
    context stream; what bench.py and torch callers use) gives the same
    numbers as host mode: positions uploaded from a device tensor, F, the
    float32 and float64 gradients, G and grad G."""
    import torch
    dev = _lib.PYIID_DEVICE
    pos = structures.jitter(structures.fcc_sphere(9.0), 0.05, seed=8).astype(np.float32)
    n, qbin, nq = len(pos), 0.02, 1250
    load(ctx, pos, np.full(n, 79), qbin, nq)
    f_host = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    fq_host = np.empty(nq, np.float32)
    g_host = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, fq_out=fq_host)
    g64_host = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True)

    ctx.set_positions(torch.from_numpy(pos).cuda(), flags=dev)
    f_dev = torch.empty(nq, dtype=torch.float32, device="cuda")
    ctx.fq(_lib.SLOT_FQ, qbin, nq, out=f_dev, flags=dev)
    g_dev = torch.empty((n, 3, nq), dtype=torch.float32, device="cuda")
    fq_dev = torch.empty(nq, dtype=torch.float32, device="cuda")
    ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, out=g_dev, fq_out=fq_dev, flags=dev)
    g64_dev = torch.empty((n, 3, nq), dtype=torch.float64, device="cuda")
    ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, out=g64_dev, flags=dev | _lib.PYIID_F64)
    ctx.synchronize()
    np.testing.assert_array_equal(f_dev.cpu().numpy(), f_host)
    np.testing.assert_array_equal(fq_dev.cpu().numpy(), fq_host)
    np.testing.assert_array_equal(g_dev.cpu().numpy(), g_host)
    np.testing.assert_array_equal(g64_dev.cpu().numpy(), g64_host)

    exp, pq = oracle.make_experiment({"qmin": 0.5, "rmax": 20.0})
    nqp = oracle.qmax_bin(exp["qmax"], pq)
    r = oracle.r_grid(exp)
    ctx.set_transform(exp["rstep"], pq, r, exp["qmin"], nqp)
    fq_in = np.random.RandomState(3).randn(nqp).astype(np.float32)
    gr_host = ctx.pdf(fq_in, len(r))
    gr_dev = torch.empty(len(r), dtype=torch.float64, device="cuda")
    ctx.pdf(torch.from_numpy(fq_in).cuda(), len(r), out=gr_dev, flags=dev)
    gq = np.random.RandomState(4).randn(n, 3, nqp).astype(np.float32)
    gg_host = ctx.grad_pdf(gq, len(r))
    gg_dev = torch.empty((n, 3, len(r)), dtype=torch.float64, device="cuda")
    ctx.grad_pdf(torch.from_numpy(gq).cuda(), len(r), out=gg_dev, flags=dev)
    ctx.synchronize()
    np.testing.assert_array_equal(gr_dev.cpu().numpy(), gr_host)
    np.testing.assert_array_equal(gg_dev.cpu().numpy(), gg_host)
