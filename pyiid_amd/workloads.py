"""Seeded sampler workloads (structures and PDF targets) used by the benchmark
and tests.

* ``batched_chains``: 256 Au FCC particles, sizes uniform in [100, 3000]
  (seeded), each the n sites nearest an atom; chain b starts from its ideal
  particle with sigma = 0.05 A seeded jitter and is refined against the
  G(r) of an independently rattled copy (sigma = 0.1 A).
* ``large_chains``: 64 Au particles of 5,000 atoms, same protocol.

Targets are computed on the device (``pyiid_pdf_energy_forces`` needs a G(r)
target; the batch F(Q) -> G(r) path produces it).
"""
import numpy as np

from . import _lib, structures
from .formfactors import element_table

EXP = {"qmin": 0.0, "qmax": 25.0, "qbin": 0.1, "rmin": 0.0, "rmax": 40.0, "rstep": 0.01}


def chain_sizes(kind, seed=0):
    if kind == "batched":
        return np.random.RandomState(seed).randint(100, 3001, size=256)
    if kind == "large":
        return np.full(64, 5000)
    if kind == "small":
        return np.random.RandomState(seed).randint(100, 401, size=16)
    raise ValueError(kind)


def build(kind, seed=0, start_sigma=0.05, target_sigma=0.1):
    """(start positions list, target positions list, atomic numbers list)."""
    sizes = chain_sizes(kind, seed)
    cache = {}
    starts, targets, numbers = [], [], []
    for b, n in enumerate(sizes):
        n = int(n)
        if n not in cache:
            cache[n] = structures.fcc_nearest(n)
        base = cache[n]
        starts.append(structures.jitter(base, start_sigma, seed=1000 * seed + 2 * b))
        targets.append(structures.jitter(base, target_sigma, seed=1000 * seed + 2 * b + 1))
        numbers.append(np.full(n, 79))
    return starts, targets, numbers


def target_pdfs(scat, target_positions, numbers, device=0):
    """G(r) of every target structure on scat's grid, computed on the device."""
    ctx = _lib.Context(device)
    ctx.set_batch(target_positions)
    nq = int(np.floor(scat.exp["qmax"] / scat.pdf_qbin))
    allz = np.concatenate(numbers)
    table, index, _ = element_table(allz, scat.pdf_qbin, nq, scat.form_factor_source)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_PDF, table)
    fq = ctx.fq(_lib.SLOT_PDF, scat.pdf_qbin, nq, n_chain=len(target_positions))
    r = scat.get_r()
    ctx.set_transform(scat.exp["rstep"], scat.pdf_qbin, r, scat.exp["qmin"], nq)
    out = np.stack([ctx.pdf(np.ascontiguousarray(fq[b] if fq.ndim == 2 else fq), len(r))
                    for b in range(len(target_positions))])
    ctx.close()
    return out
