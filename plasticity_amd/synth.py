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


def _mb_dir(u, v, t, w, ca):
    """Cartesian vector of the Miller-Bravais direction [u v t w] (a1 along x, c along z, a = 1)."""
    return np.array([u - 0.5 * v - 0.5 * t, 0.5 * np.sqrt(3.0) * (v - t), ca * w])


def _mb_normal(h, k, i, l, ca):
    """Cartesian normal of the Miller-Bravais plane (h k i l): h a1* + k a2* + l c*."""
    return np.array([h, (h + 2.0 * k) / np.sqrt(3.0), l / ca])


def _mb_family(first3, last):
    """Distinct (up to sign) index vectors whose first three entries permute `first3` (signs kept,
    also negated) and whose last is +-last."""
    import itertools
    out = []
    for p in set(itertools.permutations(first3)):
        for s in (1, -1):
            for ls in ((1, -1) if last else (1,)):
                v = tuple(s * a for a in p) + (ls * last,)
                if tuple(-a for a in v) not in out and v not in out:
                    out.append(v)
    return out


def hcp24_tables(ca=1.624):
    """HCP with 24 slip systems, c/a = 1.624: basal {0001}<11-20> (3), prismatic {10-10}<11-20> (3),
    pyramidal <a> {10-11}<11-20> (6), first-order pyramidal <c+a> {10-11}<11-23> (12); a system is a
    plane of the family and a direction of the family lying in it (h u + k v + i t + l w = 0), each
    pair counted once up to sign.  Elastic constants (C11 59.3, C12 25.7, C13 21.4, C33 61.5,
    C44 16.4 GPa, C66 = (C11 - C12) / 2) and hardening per family from the reference's
    hcp/simpleTension input (its pyramidal <c+a> values for all 12); latent ratio 1."""
    a_dirs = _mb_family((2, -1, -1), 0)
    ca_dirs = _mb_family((1, 1, -2), 3)
    fams = [([(0, 0, 0, 1)], a_dirs, (45.0, 150.0, 0.0, 95.0)),
            (_mb_family((1, -1, 0), 0), a_dirs, (105.0, 150.0, 6.0, 135.0)),
            (_mb_family((1, -1, 0), 1), a_dirs, (7000.0, 0.0, 6.0, 12000.0)),
            (_mb_family((1, -1, 0), 1), ca_dirs, (210.0, 150.0, 6.0, 420.0))]
    m, n, s0, h0, pw, ss = [], [], [], [], [], []
    for planes, dirs, (a, b, c, d) in fams:
        for pl in planes:
            for dr in dirs:
                if pl[0] * dr[0] + pl[1] * dr[1] + pl[2] * dr[2] + pl[3] * dr[3] != 0:
                    continue
                nv, mv = _mb_normal(*pl, ca), _mb_dir(*dr, ca)
                n.append(nv / np.linalg.norm(nv))
                m.append(mv / np.linalg.norm(mv))
                s0.append(a); h0.append(b); pw.append(c); ss.append(d)
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = 59.3e3
    C[0, 1] = C[1, 0] = 25.7e3
    C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = 21.4e3
    C[2, 2] = 61.5e3
    C[3, 3] = C[4, 4] = 16.4e3
    C[5, 5] = 16.8e3
    return dict(m=np.array(m), n=np.array(n), q=np.ones((len(m), len(m))), C=C, s0=s0, h0=h0, ss=ss, pw=pw, tol=1e-9)


def bcc48_tables():
    """BCC with 48 slip systems: {110}<111> (12), {112}<111> (12), {123}<111> (24), every plane of
    the family with every <111> direction in it (each pair once up to sign); C11/C12/C44 =
    231/135/117 GPa; the reference's bcc/simpleTension hardening (s0 200, h0 1500, a 1, s_s 500 MPa);
    latent ratio 1 on coplanar systems, 1.4 otherwise."""
    import itertools

    def family(hkl):
        out = []
        for p in set(itertools.permutations(hkl)):
            for sg in itertools.product((1, -1), repeat=3):
                v = tuple(a * b for a, b in zip(p, sg))
                if v not in out and tuple(-a for a in v) not in out:
                    out.append(v)
        return sorted(out)
    dirs = family((1, 1, 1))
    m, n = [], []
    for fam in ((1, 1, 0), (1, 1, 2), (1, 2, 3)):
        for nv in family(fam):
            for d in dirs:
                if np.dot(nv, d) == 0:
                    n.append(np.array(nv, float) / np.linalg.norm(nv))
                    m.append(np.array(d, float) / np.linalg.norm(d))
    m, n = np.array(m), np.array(n)
    q = np.where(np.abs(n @ n.T - 1.0) < 1e-12, 1.0, 1.4)
    C = np.zeros((6, 6))
    C[:3, :3] = 135e3
    np.fill_diagonal(C[:3, :3], 231e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 117e3
    k = len(m)
    return dict(m=m, n=n, q=q, C=C, s0=[200.0] * k, h0=[1500.0] * k, ss=[500.0] * k, pw=[1.0] * k, tol=1e-9)


def batch_workload(n_rve=64, n_min=8, n_max=24, grains=16, steps=10, strain_per_step=1e-3, seed=2024):
    """The batched-RVE workload: n_rve unit cubes with edge n ~ U{n_min..n_max} elements (seeded),
    alternately BCC with 48 systems and HCP with 24 (bcc48_tables, hcp24_tables), each with its own
    `grains` seeded Voronoi grains and random orientations; uniaxial tension (face x- fixed, face
    x+ pulled in x) in `steps` steps of `strain_per_step`.  Returns (materials, rve_mat, rve_nel,
    problem, grain maps per RVE (element order x fastest), orientations per RVE)."""
    from .model import make_material, make_problem
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_rve)
    tabs = [bcc48_tables(), hcp24_tables()]
    mats = [make_material(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], t["tol"]) for t in tabs]
    bcs = [(1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, strain_per_step * steps)]
    pb = make_problem((1, 1, 1), (1.0, 1.0, 1.0), steps, 1.0, bcs, lin_tol=1e-10)
    rve_mat = np.arange(n_rve) % 2
    rve_nel = np.stack([sizes, sizes, sizes], axis=1)
    egs, rods = [], []
    for r in range(n_rve):
        n = int(sizes[r])
        vox = voronoi_voxels((n, n, n), grains, seed + 101 + r)
        egs.append(vox.transpose(2, 1, 0).ravel().astype(np.uint32))
        rod = np.zeros((grains + 1, 3))
        rod[1:] = random_rodrigues(grains, seed + 5003 + r)
        rods.append(rod)
    return mats, rve_mat, rve_nel, pb, egs, rods
