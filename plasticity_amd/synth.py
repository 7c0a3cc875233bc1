"""Seeded synthetic microstructures and the FCC/BCC material tables of the benchmark workloads.

Voronoi grains: nearest of `n_grains` seed points drawn uniformly in the domain (numpy PCG64,
seed recorded), ties to the lowest seed id, evaluated at the voxel centres of a grid equal to the
element grid (so every cell centre is a voxel centre and the nearest-voxel lookup has no ties).
Orientations: uniform random rotations by Shoemake's method, written as Rodrigues vectors
r = q_vec / q0 with q0 >= 0 (the reference's orientation-file convention, rotationOperations.cc:4-51).
"""
import os

import numpy as np

FCC_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "fcc")


def voronoi_voxels(nv, n_grains, seed, span=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0.0, 1.0, size=(n_grains, 3)) * np.asarray(span)
    axes = [(np.arange(n) + 0.5) * (span[d] / n) for d, n in enumerate(nv)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    out = np.empty(len(pts), dtype=np.uint32)
    for s in range(0, len(pts), 65536):
        p = pts[s: s + 65536]
        d = ((p[:, None, :] - seeds[None, :, :]) ** 2).sum(axis=2)
        out[s: s + 65536] = np.argmin(d, axis=1)  # first minimum = lowest seed id
    return (out + 1).reshape(nv)  # grain ids start at 1 as in the reference files


def random_rodrigues(n, seed):
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.uniform(size=(3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    q[q[:, 0] < 0] *= -1.0
    q[:, 0] = np.maximum(q[:, 0], 1e-12)
    return q[:, 1:] / q[:, :1]


def fcc_tables():
    """12 {111}<110> systems and the coplanar/off-plane latent ratios 1 / 1.4 of the FCC workload."""
    m = np.loadtxt(os.path.join(FCC_DIR, "slipDirections.txt"))
    n = np.loadtxt(os.path.join(FCC_DIR, "slipNormals.txt"))
    q = np.loadtxt(os.path.join(FCC_DIR, "LatentHardeningRatio.txt"))
    C = np.zeros((6, 6))
    C[:3, :3] = 124e3
    np.fill_diagonal(C[:3, :3], 170e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 75e3
    return dict(m=m, n=n, q=q, C=C, s0=[16.0] * 12, h0=[180.0] * 12, ss=[148.0] * 12, pw=[2.25] * 12, tol=1e-9)


def fcc_workload(n, n_grains, steps, strain_per_step, seed=2024):
    """Uniaxial tension of a unit cube, n^3 Q1 elements: face x- fixed, face x+ pulled in x."""
    from .model import make_material, make_problem
    t = fcc_tables()
    mat = make_material(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], t["tol"])
    total_u = strain_per_step * steps
    bcs = [(1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, total_u)]
    pb = make_problem((n, n, n), (1.0, 1.0, 1.0), steps, 1.0, bcs)
    vox = voronoi_voxels((n, n, n), n_grains, seed)
    rod = np.zeros((n_grains + 1, 3))
    rod[1:] = random_rodrigues(n_grains, seed + 1)
    return mat, pb, vox, rod, bcs, t
