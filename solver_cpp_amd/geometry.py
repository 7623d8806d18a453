"""Seeded synthetic microstructures: periodic Voronoi grains and uniform
orientations on SO(3).

The reference reads grains from a VTI file (discretization_grid.cpp:8-84)
whose ``material`` array indexes the ``material`` list of material.yaml; the
grid examples are made by DAMASK's periodic Voronoi tessellation of seeded
points.  These helpers produce the same kind of input from a seed:
``material_at`` (0-based grain index per cell, x fastest) and one unit
quaternion (w, x, y, z) per grain.
"""
from __future__ import annotations

import numpy as np


def cell_centres(cells, size) -> np.ndarray:
    """(N, 3) cell-centre coordinates, x fastest (cells x, y, z)."""
    cells = [int(c) for c in cells]
    axes = [(np.arange(n) + 0.5) * (float(s) / n) for n, s in zip(cells, size)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def voronoi(cells, size, n_grains: int, seed: int = 0) -> np.ndarray:
    """Periodic Voronoi tessellation of ``n_grains`` uniform random seeds.

    Returns material_at (int32, N = prod(cells), x fastest): the index of the
    nearest seed under periodic distance.  Ties go to the lower seed index.
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    size = np.asarray(size, dtype=np.float64)
    seeds = rng.random((int(n_grains), 3)) * size
    tree = cKDTree(seeds, boxsize=size)
    _, idx = tree.query(cell_centres(cells, size) % size, k=1)
    return np.ascontiguousarray(idx.astype(np.int32))


def uniform_orientations(n: int, seed: int = 0) -> np.ndarray:
    """(n, 4) unit quaternions (w, x, y, z), uniform on SO(3) (Shoemake 1992), w >= 0."""
    rng = np.random.default_rng(seed)
    u = rng.random((int(n), 3))
    a = np.sqrt(1.0 - u[:, 0])
    b = np.sqrt(u[:, 0])
    q = np.stack([b * np.cos(2 * np.pi * u[:, 2]), a * np.sin(2 * np.pi * u[:, 1]),
                  a * np.cos(2 * np.pi * u[:, 1]), b * np.sin(2 * np.pi * u[:, 2])], axis=1)
    q[q[:, 0] < 0] *= -1.0
    return np.ascontiguousarray(q / np.linalg.norm(q, axis=1, keepdims=True))


def polycrystal(cells, n_grains: int, seed: int = 0, size=None):
    """(material_at, orientations, size) of a seeded periodic polycrystal; the
    default size is one unit length per cell edge scaled to 1e-5 m (the grid
    examples' 16 cells over 1e-5 m per edge)."""
    if size is None:
        size = [1.0e-5 * c / 16.0 for c in cells]
    return voronoi(cells, size, n_grains, seed), uniform_orientations(n_grains, seed + 7919), list(size)
