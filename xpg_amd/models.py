"""Crystal models for the per-block refinement before integration (float64 numpy, host).

Mirrors xpg/models.py:1-111: the free cell parameters of each lattice system, the Busing-Levy B matrix, and a Crystal
held as orientation U and cell (A* = U B).  The integrator reads only ``Astar`` and ``space_group``, so these crystals
go wherever geometry.Crystal goes.  Every function is batched over leading dimensions: the refinement evaluates a
parameter vector and its forward-difference perturbations together.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

# free cell parameters per lattice system, conventional settings (xpg/models.py:10-18)
CELL_PARAMS = {
    "a": ["a", "b", "c", "al", "be", "ga"],
    "m": ["a", "b", "c", "be"],
    "o": ["a", "b", "c"],
    "t": ["a", "c"],
    "h": ["a", "c"],
    "c": ["a"],
}


def cell_from_free(system, v):
    """Full cell (..., 6) from the free parameters (..., n) of a lattice system (xpg/models.py:21-35)."""
    v = np.asarray(v, float)
    e = [v[..., i] for i in range(v.shape[-1])]
    if system not in CELL_PARAMS:
        raise ValueError(system)
    if system == "a":
        cols = e[:6]
    else:
        n90 = np.full(v.shape[:-1], 90.0)
        cols = {"m": lambda: [e[0], e[1], e[2], n90, e[3], n90],
                "o": lambda: [e[0], e[1], e[2], n90, n90, n90],
                "t": lambda: [e[0], e[0], e[1], n90, n90, n90],
                "h": lambda: [e[0], e[0], e[1], n90, n90, np.full(v.shape[:-1], 120.0)],
                "c": lambda: [e[0], e[0], e[0], n90, n90, n90]}[system]()
    return np.stack(cols, -1)


def free_from_cell(system, cell):
    """Free parameters of a cell (xpg/models.py:50-53): averages of the lengths the system ties together."""
    a, b, c, al, be, ga = (float(x) for x in cell)
    return np.array({"a": [a, b, c, al, be, ga], "m": [a, b, c, be], "o": [a, b, c], "t": [(a + b) / 2, c],
                     "h": [(a + b) / 2, c], "c": [(a + b + c) / 3]}[system], float)


def symmetrize_cell(system, cell):
    return cell_from_free(system, free_from_cell(system, cell))


def b_matrix(cell):
    """Busing-Levy B (columns a*, b*, c*) of cells (..., 6) -> (..., 3, 3): the inverse of the real basis with a
    along x and b in the xy plane (xpg/models.py:64-78)."""
    cell = np.asarray(cell, float)
    a, b, c = cell[..., 0], cell[..., 1], cell[..., 2]
    al, be, ga = (np.radians(cell[..., i]) for i in (3, 4, 5))
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz = np.sqrt(np.maximum(c * c - cx * cx - cy * cy, 1e-12))
    z = np.zeros_like(a)
    real = np.stack([np.stack([a, z, z], -1), np.stack([b * np.cos(ga), b * np.sin(ga), z], -1),
                     np.stack([cx, cy, cz], -1)], -2)
    return np.linalg.inv(real)


def skew(u):
    """Cross-product matrices of (..., 3) vectors."""
    u = np.asarray(u, float)
    k = np.zeros(u.shape[:-1] + (3, 3))
    k[..., 0, 1], k[..., 0, 2], k[..., 1, 2] = -u[..., 2], u[..., 1], -u[..., 0]
    k[..., 1, 0], k[..., 2, 0], k[..., 2, 1] = u[..., 2], -u[..., 1], u[..., 0]
    return k


def rotvec_to_matrix(r):
    """Rotation matrices of rotation vectors (axis x angle, (..., 3)) by Rodrigues' formula, with the series
    1 - t^2/6 and 1/2 - t^2/24 below t^2 = 1e-12 (xpg/geometry.py:52-65).  (t = sqrt(t^2 + 1e-30) keeps both
    quotients finite, so no floating-point warning can arise.)"""
    r = np.asarray(r, float)
    th2 = np.sum(r * r, -1)[..., None, None]
    th = np.sqrt(th2 + 1e-30)
    k = skew(r)
    small = th2 < 1e-12
    a = np.where(small, 1.0 - th2 / 6.0, np.sin(th) / th)
    b = np.where(small, 0.5 - th2 / 24.0, (1.0 - np.cos(th)) / (th2 + 1e-30))
    return np.eye(3) + a * k + b * (k @ k)


def real_basis(Astar):
    return np.linalg.inv(Astar)


def cell_from_Astar(Astar):
    R = real_basis(Astar)
    a, b, c = (np.linalg.norm(R[i]) for i in range(3))
    ang = lambda u, v, nu, nv: np.degrees(np.arccos(np.dot(u, v) / (nu * nv)))
    return np.array([a, b, c, ang(R[1], R[2], b, c), ang(R[0], R[2], a, c), ang(R[0], R[1], a, b)])


def U_from_Astar(Astar, cell=None):
    """Orientation U with A* = U B(cell), the rotation nearest to A* B^-1 (polar decomposition by SVD)."""
    if cell is None:
        cell = cell_from_Astar(Astar)
    u, _, vt = np.linalg.svd(np.asarray(Astar, float) @ np.linalg.inv(b_matrix(cell)))
    U = u @ vt
    if np.linalg.det(U) < 0:
        raise ValueError("left-handed basis")
    return U


@dataclass
class Crystal:
    """Orientation and cell (xpg/models.py:81-104); A* = U B(cell)."""
    U: np.ndarray
    cell: np.ndarray
    system: str = "a"
    centring: str = "P"
    space_group: str = "P 1"
    mosaicity: float = 0.1
    divergence: float = 0.03

    @property
    def Astar(self):
        return self.U @ b_matrix(self.cell)

    @classmethod
    def from_Astar(cls, Astar, system="a", **kw):
        cell = symmetrize_cell(system, cell_from_Astar(Astar))
        return cls(U=U_from_Astar(Astar, cell), cell=cell, system=system, **kw)

    def copy(self, **kw):
        return replace(self, **kw)
