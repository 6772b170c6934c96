"""Gamma-centred k-point grids in the reference's KPOINTS text format.

The reference reads k-points from a text file (import_kpt,
src/pmx_basis.cpp:200-245): a line with NKPT0 (integration points in the whole
zone), a line with NKPT (points listed), then NKPT lines `kx ky kz w` in units
of 2pi/a, with the weight divided by NKPT0 on import.  The shipped
include/kpoints/KPOINTS_NxNxN.txt files are Gamma-centred N^3 grids
sum_j (n_j/N) b_j folded into the first Brillouin zone with unit weights.

`gamma_grid` builds the same kind of grid for any crystal the CLI knows, with
independent divisions along b1, b2, b3 (16x16x12 for a hexagonal cell), and
`write_kpoints` writes it in that format.
"""
import itertools
import math

import numpy as np

_R3 = math.sqrt(3.0)


def reciprocal(crystal, f=math.sqrt(8.0 / 3.0)):
    """Reciprocal vectors in 2pi/a (init_crystal, src/pmx_io.cpp:126-229)."""
    if crystal in ("zb", "rs"):
        return np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], float)
    if crystal in ("wz", "te"):
        return np.array([[1, -1 / _R3, 0], [1, 1 / _R3, 0], [0, 0, 1.0 / f]], float)
    raise ValueError("k-point grids are defined for zb, rs, wz and te cells")


def gamma_grid(crystal, n1, n2=None, n3=None, f=math.sqrt(8.0 / 3.0)):
    """Gamma-centred n1 x n2 x n3 grid folded into the first zone.

    Returns K [n1*n2*n3, 3] (2pi/a).  The fastest index runs along b1, as in
    the shipped KPOINTS files.  Each point is moved by the reciprocal lattice
    vector that brings it closest to Gamma; on a zone-face tie the first such
    vector in a fixed search order wins, so the output is deterministic.
    """
    n2 = n1 if n2 is None else n2
    n3 = n1 if n3 is None else n3
    b = reciprocal(crystal, f)
    shifts = np.array([i * b[0] + j * b[1] + k * b[2]
                       for i, j, k in itertools.product(range(-2, 3), repeat=3)])
    pts = []
    for i3 in range(n3):
        for i2 in range(n2):
            for i1 in range(n1):
                k = (i1 / n1) * b[0] + (i2 / n2) * b[1] + (i3 / n3) * b[2]
                d = np.einsum("ij,ij->i", k - shifts, k - shifts)
                best = np.flatnonzero(d <= d.min() + 1e-9)[0]
                pts.append(k - shifts[best])
    return np.array(pts)


def write_kpoints(path, K, weights=None, nkpt0=None):
    """Write K (and weights, default 1) in the import_kpt text format."""
    K = np.asarray(K, float).reshape(-1, 3)
    w = np.ones(len(K)) if weights is None else np.asarray(weights, float)
    nk0 = len(K) if nkpt0 is None else int(nkpt0)
    with open(path, "w") as fh:
        fh.write("%d\n%d\n" % (nk0, len(K)))
        for k, wk in zip(K, w):
            fh.write("%.17g %.17g %.17g %.17g\n" % (k[0], k[1], k[2], wk))


def read_kpoints(path):
    """Parse a KPOINTS file; returns (K [NKPT,3], KW [NKPT]) with KW = w/NKPT0."""
    K, KW = [], []
    nk0 = nk = None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if not line or line[0] == "#":
                continue
            if nk0 is None:
                nk0 = int(line.split()[0])
            elif nk is None:
                nk = int(line.split()[0])
            elif len(K) < nk:
                t = line.split()
                K.append([float(t[0]), float(t[1]), float(t[2])])
                KW.append(float(t[3]) / nk0)
    return np.array(K, float).reshape(-1, 3), np.array(KW, float)
