"""Lattices: site count plus edge list.

``chain`` and ``kagome`` produce the same edge sets as the reference's
factories (codes/pygfmc/hamiltonian/lattice.py:36-46 and :186-262); the edge
sets have to match so that a model built on ``kagome(12, 12)`` is the same
model.  ``square`` is new: the reference has no square-lattice factory.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Lattice:
    """A finite lattice (mirrors ``Lattice``, lattice.py:9-33)."""

    name: str
    N: int
    edges: list
    J_x: object = None
    J_y: object = None
    J_z: object = None
    extras: dict = field(default_factory=dict, compare=False, hash=False)


def chain(n_sites: int, periodic: bool = True) -> Lattice:
    """Nearest-neighbour chain; the closing bond ``(n-1, 0)`` when periodic and
    ``n > 2``; suggested couplings J_x = J_y = -1, J_z = 1 (lattice.py:36-46)."""
    n = int(n_sites)
    edges = [(i, i + 1) for i in range(n - 1)]
    if periodic and n > 2:
        edges.append((n - 1, 0))
    return Lattice(name="chain", N=n, edges=edges, J_x=-1.0, J_y=-1.0, J_z=1.0)


def square(Lx: int, Ly: int, periodic: bool = True) -> Lattice:
    """Nearest-neighbour square lattice of ``Lx`` x ``Ly`` sites, site index
    ``x + Lx*y``.  Edges are the distinct ``(min, max)`` pairs, sorted, so a
    periodic side of length 2 does not double a bond.  N = Lx*Ly; 2N edges
    when periodic with Lx, Ly >= 3."""
    Lx, Ly = int(Lx), int(Ly)
    if Lx < 1 or Ly < 1:
        raise ValueError(f"square needs Lx, Ly >= 1, got {Lx}, {Ly}")
    out = set()
    for y in range(Ly):
        for x in range(Lx):
            i = x + Lx * y
            for dx, dy in ((1, 0), (0, 1)):
                xx, yy = x + dx, y + dy
                if not periodic and (xx >= Lx or yy >= Ly):
                    continue
                j = (xx % Lx) + Lx * (yy % Ly)
                if i != j:
                    out.add((min(i, j), max(i, j)))
    pos = np.array([[x, y] for y in range(Ly) for x in range(Lx)], dtype=float)
    extras = dict(positions=pos, supercell_dims=np.array([Lx, Ly]))
    if periodic:
        extras["translations"] = square_translations(Lx, Ly)
    return Lattice(name="square", N=Lx * Ly, edges=sorted(out), extras=extras)


def square_translations(Lx: int, Ly: int) -> np.ndarray:
    """The Lx*Ly translations of the periodic square lattice as site
    permutations: row r = dx + Lx*dy maps site x + Lx*y to
    (x + dx) % Lx + Lx*((y + dy) % Ly)."""
    x = np.arange(Lx * Ly) % Lx
    y = np.arange(Lx * Ly) // Lx
    return np.array([((x + dx) % Lx) + Lx * ((y + dy) % Ly)
                     for dy in range(Ly) for dx in range(Lx)], dtype=np.int64)


def kagome(Lx: int, Ly: int, periodic: bool = True) -> Lattice:
    """Kagome lattice of ``Lx`` x ``Ly`` three-site cells (a1 = (2, 0),
    a2 = (1, sqrt 3), sites at 0, a1/2, a2/2; site index 3*(x + Lx*y) + b).

    Bonds per cell: the up triangle (b0-b1, b0-b2, b1-b2) and the down
    triangle joining b1 to b0 of cell (x+1, y), b2 to b0 of cell (x, y+1) and
    b1 to b2 of cell (x+1, y-1).  Edges are distinct (min, max) pairs, sorted
    (the same set as lattice.py:186-262).  Periodic needs Lx, Ly >= 2."""
    Lx, Ly = int(Lx), int(Ly)
    if periodic and (Lx < 2 or Ly < 2):
        raise ValueError(
            f"kagome requires Lx, Ly >= 2 for periodic boundary conditions "
            f"(got {Lx}, {Ly}); smaller tori wrap a bond onto itself.")

    def site(x, y, b):
        if not periodic and not (0 <= x < Lx and 0 <= y < Ly):
            return None
        return 3 * ((x % Lx) + Lx * (y % Ly)) + b

    bonds = [((0, 0, 0), (0, 0, 1)), ((0, 0, 0), (0, 0, 2)), ((0, 0, 1), (0, 0, 2)),
             ((0, 0, 1), (1, 0, 0)), ((0, 0, 2), (0, 1, 0)), ((0, 0, 1), (1, -1, 2))]
    out = set()
    for x in range(Lx):
        for y in range(Ly):
            for (dx1, dy1, b1), (dx2, dy2, b2) in bonds:
                i = site(x + dx1, y + dy1, b1)
                j = site(x + dx2, y + dy2, b2)
                if i is None or j is None:
                    continue
                out.add((min(i, j), max(i, j)))
    a1 = np.array([2.0, 0.0])
    a2 = np.array([1.0, np.sqrt(3.0)])
    basis = np.array([[0.0, 0.0], a1 / 2, a2 / 2])
    pos = np.array([x * a1 + y * a2 + basis[b]
                    for y in range(Ly) for x in range(Lx) for b in range(3)])
    return Lattice(name="kagome", N=3 * Lx * Ly, edges=sorted(out),
                   extras=dict(positions=pos, primitive_vectors=np.array([a1, a2]),
                               supercell_dims=np.array([Lx, Ly]),
                               sublattice_offsets=basis))
