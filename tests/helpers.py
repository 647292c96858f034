"""Shared fixtures for the parity tests: seeded conformations and imports of the
product binding (smdlt_abi) and the checker (oracle)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "smdlt-ca-chrg-flow_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402  (the checker)
import smdlt_abi as S  # noqa: E402  (the product binding)


def random_walk(nb, nc, L, seed, bond=2.0, min_sep=1.6, tries=200):
    """Seeded random-walk chains with bond length `bond`, rejecting steps closer
    than min_sep to earlier beads of the same chain; positions wrapped into
    [-L/2, L/2)^3 (like initChains, S/main26.c:251-508, without its RNG)."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((nb * nc, 3))
    for c in range(nc):
        base = c * nb
        pos[base] = rng.uniform(-L / 2, L / 2, 3)
        for j in range(1, nb):
            for _ in range(tries):
                d = rng.normal(size=3)
                d *= bond / np.linalg.norm(d)
                p = pos[base + j - 1] + d
                if j < 2 or np.min(np.linalg.norm(pos[base:base + j - 1] - p, axis=1)) > min_sep:
                    break
            pos[base + j] = p
    pos = pos - L * np.round(pos / L)
    return pos


def cell_from_lattice_row(row, L):
    """Oracle cell from smdlt_update_lattice's row (L1x L1y L2x L2y L1xp L2xp L2yp delx)."""
    return O.Cell(row[4], 0.0, row[5], row[6], L)
