"""Parity sweep on random structures: uniform random boxes of 20-60 A (so
qbin * r crosses several multiples of pi at 0.1 A^-1 per bin, and close
contacts occur), one or two elements, against the FP64 oracle: F(Q) (both
kernels) and the gradient to 1e-5 of max|x*|.  The one-pass F(Q)-Rw energy
and forces (the unordered-pair force kernel) are checked against the oracle's
Rw and Rw gradient evaluated on the device F: on these structures the Rw
forces move by up to 4e-4 for 1e-6 noise in F, so that isolates the
contraction (1e-5) and the potential from F's own rounding (the device
energy uses its FP64 F, the comparison the float32 output: 1e-7)."""
import numpy as np
import pytest

from oracle import scatter as oracle
from pyiid_amd import _lib
from pyiid_amd.formfactors import element_table

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


CASES = [  # (seed, atoms, box A, qbin, elements)
    (0, 150, 20.0, 0.1, (79,)),
    (1, 300, 40.0, 0.1, (79,)),
    (2, 400, 60.0, 0.1, (78, 28)),
    (3, 250, 60.0, 0.05, (79,)),
    (4, 200, 30.0, 0.2, (78, 28)),
]


@pytest.mark.parametrize("seed,n,box,qbin,elements", CASES)
def test_random_box(ctx, seed, n, box, qbin, elements):
    rs = np.random.RandomState(seed)
    pos = (rs.random_sample((n, 3)) * box).astype(np.float32)
    numbers = np.array(elements)[rs.randint(0, len(elements), n)]
    nq = int(np.floor(25.0 / qbin))
    table, index, _ = element_table(numbers, qbin, nq)
    per_atom = table[index]
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    f_want = oracle.fq(pos, per_atom, qbin)
    g_want = oracle.grad_fq(pos, per_atom, qbin)
    f = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    f_fused = np.empty(nq, np.float32)
    g = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True, fq_out=f_fused)
    q_first = int(np.ceil(1.0 / qbin))
    shrunk = (pos - pos.mean(0)) * np.float32(0.998) + pos.mean(0)
    target = oracle.fq(shrunk.astype(np.float32), per_atom, qbin)[q_first:]
    e, forces = ctx.fq_energy_forces(_lib.SLOT_FQ, qbin, q_first, target, _lib.POT_RW, 1.0, n)
    f_dev = f.astype(np.float64)[q_first:]
    e_want = oracle.get_rw(target, f_dev)[0]
    forces_want = oracle.wrap_grad_rw(g_want[:, :, q_first:], f_dev, target)
    errs = (rel(f, f_want), rel(f_fused, f_want), rel(g, g_want), abs(e - e_want),
            rel(forces, forces_want))
    print("seed %d n %d box %g qbin %g: F %.2e F(fused) %.2e grad %.2e Rw %.1e forces %.2e"
          % ((seed, n, box, qbin) + errs))
    assert errs[0] < TOL and errs[1] < TOL and errs[2] < TOL
    assert errs[3] < 1e-7
    assert errs[4] < TOL
