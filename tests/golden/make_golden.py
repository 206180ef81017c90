"""Writes tests/golden/cuboctahedron55.npz from the FP64 oracle.

Workload: 55-atom Au FCC cuboctahedron (a = 4.08 A), positions jittered by
seeded Gaussian sigma = 0.05 A (RandomState(0)), float32; default experiment
(Q 0-25 step 0.1, r 0-40 step 0.01); F(Q), grad F(Q), G(r) of the jittered
and unjittered particles, Rw of jittered vs unjittered, and the Calc1D Rw
forces (reference convention, conv = 1).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import scatter as S  # noqa: E402
from pyiid_amd import structures  # noqa: E402
from pyiid_amd.formfactors import element_table  # noqa: E402


def main():
    base = structures.cuboctahedron(4.08).astype(np.float32)
    pos = structures.jitter(base.astype(np.float64), 0.05, seed=0).astype(np.float32)
    n = len(pos)
    exp, pq = S.make_experiment()
    nq = S.qmax_bin(exp["qmax"], exp["qbin"])
    nqp = S.qmax_bin(exp["qmax"], pq)
    r = S.r_grid(exp)
    tab_f, idx, _ = element_table(np.full(n, 79), exp["qbin"], nq)
    tab_p, _, _ = element_table(np.full(n, 79), pq, nqp)
    sf, sp = tab_f[idx], tab_p[idx]
    fq = S.fq(pos, sf, exp["qbin"])
    gfq = S.grad_fq(pos, sf, exp["qbin"])
    fq_pdf = S.fq(pos, sp, pq)
    gr = S.get_pdf_at_qmin(fq_pdf, exp["rstep"], pq, r, exp["qmin"])
    gr0 = S.get_pdf_at_qmin(S.fq(base, sp, pq), exp["rstep"], pq, r, exp["qmin"])
    rw, scale = S.get_rw(gr0, gr)
    gfq_pdf = S.grad_fq(pos, sp, pq)
    gpdf = S.grad_pdf(gfq_pdf, exp["rstep"], pq, r, exp["qmin"])
    forces = S.wrap_grad_rw(gpdf, gr, gr0)
    np.savez_compressed(os.path.join(HERE, "cuboctahedron55.npz"), positions=pos,
                        positions_ideal=base, fq=fq, grad_fq=gfq, fq_pdf=fq_pdf, gr=gr,
                        gr_ideal=gr0, rw=rw, scale=scale, forces_rw=forces,
                        table_fq=tab_f, table_pdf=tab_p)
    print("wrote cuboctahedron55.npz: Rw %.6f scale %.6f" % (rw, scale))


if __name__ == "__main__":
    main()
