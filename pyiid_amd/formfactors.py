"""Per-element X-ray form factors f(Q).

The reference evaluates ``xraylib.FF_Rayl(Z, Q / 4 pi)`` per element and bin
(``pyiid/experiments/elasticscatter/kernels/master_kernel.py:14-36``).  Two
sources are provided:

* ``"xraylib"``: that same call, used when xraylib is installed;
* ``"cromer-mann"``: the International Tables (Vol. C, Table 6.1.1.4)
  Cromer-Mann parameterisation

      f(s) = sum_{i=1..4} a_i exp(-b_i s^2) + c,   s = sin(theta)/lambda = Q / (4 pi)

  for the elements used by the workloads, with no dependency.

``"auto"`` picks xraylib when it can be imported.  For a single-element
structure the table cancels exactly against the normalisation, so F(Q), G(r)
and their gradients do not depend on the source.
"""
import numpy as np

SOURCES = ("auto", "xraylib", "cromer-mann")

# Z: (symbol, (a1, b1, a2, b2, a3, b3, a4, b4, c))
CROMER_MANN = {
    6: ("C", (2.31000, 20.8439, 1.02000, 10.2075, 1.58860, 0.568700, 0.865000, 51.6512,
              0.215600)),
    8: ("O", (3.04850, 13.2771, 2.28680, 5.70110, 1.54630, 0.323900, 0.867000, 32.9089,
              0.250800)),
    14: ("Si", (6.29150, 2.43860, 3.03530, 32.3337, 1.98910, 0.678500, 1.54100, 81.6937,
                1.14070)),
    26: ("Fe", (11.7695, 4.76110, 7.35730, 0.307200, 3.52220, 15.3535, 2.30450, 76.8805,
                1.03690)),
    28: ("Ni", (12.8376, 3.87850, 7.29200, 0.256565, 4.44380, 12.1763, 2.38000, 66.3421,
                1.03410)),
    29: ("Cu", (13.3380, 3.58280, 7.16760, 0.247000, 5.61580, 11.3966, 1.67350, 64.8126,
                1.19100)),
    46: ("Pd", (19.3319, 0.698655, 15.5017, 7.98929, 5.29537, 25.2052, 0.605844, 76.8986,
                5.26593)),
    47: ("Ag", (19.2808, 0.644600, 16.6885, 7.47260, 4.80450, 24.6605, 1.04630, 99.8156,
                5.17900)),
    78: ("Pt", (27.0059, 1.51293, 17.7639, 8.81174, 15.7131, 0.424593, 5.78370, 38.6103,
                11.6883)),
    79: ("Au", (16.8819, 0.461100, 18.5913, 8.62160, 25.5582, 1.48260, 5.86000, 36.3956,
                12.0658)),
}

SYMBOL_TO_Z = {sym: z for z, (sym, _) in CROMER_MANN.items()}
# atomic masses (amu) of the same elements, for the sampler
MASSES = {6: 12.011, 8: 15.999, 14: 28.085, 26: 55.845, 28: 58.6934, 29: 63.546, 46: 106.42,
          47: 107.8682, 78: 195.084, 79: 196.966569}


def _xraylib():
    """The xraylib module when it is installed, else None."""
    try:
        import xraylib
    except ImportError:
        return None
    return xraylib


def resolve_source(source):
    """The concrete source for ``source``: "auto" becomes "xraylib" when it
    can be imported and "cromer-mann" otherwise.  Raises ValueError for an
    unknown name and ImportError for "xraylib" without xraylib."""
    if source not in SOURCES:
        raise ValueError("form factor source must be one of %s, got %r" % (SOURCES, source))
    if source == "auto":
        return "xraylib" if _xraylib() is not None else "cromer-mann"
    if source == "xraylib" and _xraylib() is None:
        raise ImportError("form factor source 'xraylib' requested but xraylib is not installed")
    return source


def form_factor(z, q, source="cromer-mann"):
    """f(Q) of element ``z`` on the Q values ``q`` (A^-1), float64, from
    ``source`` (see ``resolve_source``)."""
    source = resolve_source(source)
    if source == "xraylib":
        xrl = _xraylib()
        # xraylib takes s = sin(theta)/lambda = Q / (4 pi), divided in the
        # reference's order (master_kernel.py:35-36)
        return np.array([xrl.FF_Rayl(int(z), float(qk) / 4 / np.pi)
                         for qk in np.asarray(q, dtype=np.float64).ravel()],
                        dtype=np.float64).reshape(np.shape(q))
    try:
        _, p = CROMER_MANN[int(z)]
    except KeyError:
        raise KeyError("no Cromer-Mann parameters for Z=%d (have %s)"
                       % (int(z), sorted(CROMER_MANN))) from None
    s2 = (np.asarray(q, dtype=np.float64) / (4.0 * np.pi)) ** 2
    f = np.full(s2.shape, p[8])
    for i in range(4):
        f += p[2 * i] * np.exp(-p[2 * i + 1] * s2)
    return f


def element_table(numbers, qbin, nq, source="cromer-mann"):
    """Table f[e][m] (float32, E x nq) on Q_m = m * qbin for the elements of
    ``numbers`` in ascending Z order, plus the per-atom element index."""
    numbers = np.asarray(numbers)
    elems, index = np.unique(numbers, return_inverse=True)
    q = np.arange(nq, dtype=np.float64) * qbin
    if len(elems) == 0:
        return np.ones((1, nq), np.float32), index.astype(np.int32), elems
    table = np.stack([form_factor(z, q, source) for z in elems]).astype(np.float32)
    return table, index.astype(np.int32), elems
