"""Deterministic nanoparticle generators for the workloads (FCC clusters).

Sites are generated on the FCC lattice in units of a/2 (integer coordinates
with an even sum), centred on an atom at the origin.
"""
import numpy as np


def fcc_sites(nmax):
    """Integer FCC sites (x + y + z even) with |x|,|y|,|z| <= nmax, in a/2 units."""
    r = np.arange(-nmax, nmax + 1)
    x, y, z = np.meshgrid(r, r, r, indexing="ij")
    keep = (x + y + z) % 2 == 0
    return np.stack([x[keep], y[keep], z[keep]], axis=1)


def cuboctahedron(a=4.08, shells=2):
    """FCC cuboctahedron with ``shells`` closed shells around the centre atom
    (shells=2: 55 atoms), lattice constant ``a`` (A)."""
    s = fcc_sites(shells)
    keep = (np.abs(s).max(axis=1) <= shells) & (np.abs(s).sum(axis=1) <= 2 * shells)
    return s[keep] * (a / 2.0)


def fcc_sphere(radius, a=4.08):
    """All FCC sites within ``radius`` (A) of the atom at the origin."""
    s = fcc_sites(int(np.ceil(2 * radius / a)) + 1) * (a / 2.0)
    d = np.linalg.norm(s, axis=1)
    return s[d <= radius + 1e-9]


def fcc_nearest(n, a=4.08):
    """The ``n`` FCC sites nearest the origin (ties broken by generation order)."""
    nmax = int(np.ceil((3.0 * n / (2 * np.pi)) ** (1.0 / 3.0))) + 3
    while True:
        s = fcc_sites(nmax) * (a / 2.0)
        d = np.linalg.norm(s, axis=1)
        order = np.argsort(d, kind="stable")
        if len(s) > n and d[order[n - 1]] < (nmax - 1) * a / 2.0:
            return s[order[:n]]
        nmax += 2


def jitter(positions, sigma, seed):
    """Positions displaced by seeded Gaussian noise (``RandomState(seed).normal``)."""
    rs = np.random.RandomState(seed)
    return positions + rs.normal(0.0, sigma, positions.shape)


def random_alloy(n, z_a, z_b, seed, fraction_a=0.5):
    """Atomic numbers for an exact ``round(n * fraction_a)`` / rest split of two
    elements, assigned by a seeded permutation."""
    rs = np.random.RandomState(seed)
    na = int(round(n * fraction_a))
    numbers = np.array([z_a] * na + [z_b] * (n - na))
    return numbers[rs.permutation(n)]


def octahedron(length, a=4.08):
    """Regular FCC octahedron with ``length`` atoms on every edge (the
    ``ase.cluster.Octahedron(symbol, length)`` construction without cutoff):
    sites (i, j, k) a/2 with |i| + |j| + |k| <= length - 1 and i + j + k of
    the parity of length - 1; 6, 19, 44, 85, ... atoms."""
    L = int(length) - 1
    r = np.arange(-L, L + 1)
    i, j, k = np.meshgrid(r, r, r, indexing="ij")
    keep = (np.abs(i) + np.abs(j) + np.abs(k) <= L) & ((i + j + k - L) % 2 == 0)
    return np.stack([i[keep], j[keep], k[keep]], axis=1).astype(np.float64) * (a / 2.0)
