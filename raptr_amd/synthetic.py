"""Synthetic phantoms and list-mode counts for the benchmark workloads.

Counts are drawn Poisson (numpy's Philox generator, seeded) from the FP64
forward projection of the phantom, computed matrix-free on the GPU over all
channels, scaled to the requested total; the measurement list is the set of
channels with a non-zero count (wrappers.hpp:353-368 keeps only y != 0).
"""
import numpy as np

from . import sysmat


def spheres_phantom(grid, spheres=(((-10.0, 0.0, 0.0), 15.0, 1.0),
                                   ((20.0, 0.0, 0.0), 8.0, 4.0))):
    """Uniform spheres (centre mm, radius mm, activity) on zero background,
    voxelised by centre inclusion; overlapping spheres add."""
    c = grid.centres()
    x = np.zeros(grid.nvox, np.float64)
    for centre, r, act in spheres:
        x[np.linalg.norm(c - np.asarray(centre), axis=1) < r] += act
    return x


def random_ellipsoids(grid, n=6, seed=2, semi=(5.0, 25.0), act=(0.5, 8.0), background=0.1):
    """n axis-aligned ellipsoids with semi-axes U[semi], activity
    log-uniform in act, centres uniform in the inner 60% of the grid, on a
    uniform background; returns (image, parameters)."""
    rng = np.random.default_rng(seed)
    c = grid.centres()
    lo = np.array(grid.o) + 0.2 * np.array(grid.d) * np.array(grid.n)
    hi = np.array(grid.o) + 0.8 * np.array(grid.d) * np.array(grid.n)
    x = np.full(grid.nvox, background, np.float64)
    params = []
    for _ in range(n):
        ctr = rng.uniform(lo, hi)
        ax = rng.uniform(semi[0], semi[1], 3)
        a = float(np.exp(rng.uniform(np.log(act[0]), np.log(act[1]))))
        inside = np.sum(((c - ctr) / ax) ** 2, axis=1) <= 1.0
        x[inside] += a
        params.append((ctr, ax, a))
    return x, params


def poisson_counts(ybar, total, seed=1):
    """Poisson counts with mean ybar scaled to sum `total` (Philox stream)."""
    ybar = np.asarray(ybar, np.float64)
    lam = ybar * (float(total) / ybar.sum())
    rng = np.random.Generator(np.random.Philox(key=seed))
    return rng.poisson(lam)


def measurement(setup, grid, nrays, image, total, seed=1, ray_seed=1):
    """(channel list of y > 0, counts there as float32, full FP64 projection)."""
    ybar = sysmat.forward_project(setup, grid, np.asarray(image, np.float32), nrays,
                                  seed=ray_seed)
    y = poisson_counts(ybar, total, seed)
    meas = np.nonzero(y > 0)[0].astype(np.int32)
    return meas, y[meas].astype(np.float32), ybar
