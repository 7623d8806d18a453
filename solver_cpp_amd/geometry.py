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


def polycrystal(cells, n_grains: int, seed: int = 0, size=None, ori_seed=None):
    """(material_at, orientations, size) of a seeded periodic polycrystal in a
    1e-5 m cube (the grid examples' box) unless ``size`` is given.  Seeds:
    ``seed`` for the Voronoi points, ``ori_seed`` (default seed + 1) for the
    orientations."""
    if size is None:
        size = [1.0e-5, 1.0e-5, 1.0e-5]
    if ori_seed is None:
        ori_seed = seed + 1
    return voronoi(cells, size, n_grains, seed), uniform_orientations(n_grains, ori_seed), list(size)


# The measured configurations: grid edge, grains, Voronoi seed, orientation seed,
# load-case duration (s) and increments (tension at dot_F_xx = 1e-3/s, dt = 0.5 s)
CONFIGS = {
    "C1": dict(edge=16, grains=20, seed=1, ori_seed=2, t=10.0, N=20),
    "C2": dict(edge=64, grains=200, seed=3, ori_seed=4, t=20.0, N=40),
    "C3": dict(edge=128, grains=1000, seed=5, ori_seed=6, t=30.0, N=60),
}

# examples/grid/numerics.yaml: N_iter_min = 4, N_iter_max = 40
EXAMPLE_NUMERICS = dict(itmin=4, itmax=40)
