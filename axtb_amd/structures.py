"""Seeded synthetic molecules of the benchmark workloads (positions in bohr).

- water clusters: O on a cubic lattice (x fastest), H at 0.9572 Å and 104.52°, a uniform random
  orientation per water (Shoemake quaternion);
- all-trans n-alkanes CkH2k+2: C–C 1.53 Å, C–H 1.09 Å, tetrahedral angles, CH3 caps, optional
  Gaussian coordinate noise;
- the 4096-molecule batch (water clusters of 3–20 waters or alkanes C3–C20, 50/50) and the
  C30H62 conformer set;
- ``alkane_bench``: axtb's own benchmark alkane geometry (bench/final_timings.py:97-107), the
  geometry of its published C20H42 / C100H202 energies.
"""
import math

import numpy as np

AATOAU = 1.0 / (5.29177210544e-11 * 1e10)  # CODATA 2022, axtb base/src/units.rs:18-20
TET = math.radians(109.4712)


def _quat_rot(u):
    """Rotation matrix of a uniform random unit quaternion from three uniforms (Shoemake)."""
    u1, u2, u3 = u
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(nwat, lattice, spacing=2.9, rng=None, noise=0.0):
    """(H2O)n with O on the first n sites of a lattice^3 cubic grid; returns (numbers, bohr)."""
    rng = np.random.default_rng(0) if rng is None else rng
    roh, ang = 0.9572, math.radians(104.52)
    h1 = np.array([roh * math.sin(ang / 2), 0.0, roh * math.cos(ang / 2)])
    h2 = np.array([-roh * math.sin(ang / 2), 0.0, roh * math.cos(ang / 2)])
    z, xyz = [], []
    for s in range(nwat):
        i, j, k = s % lattice, (s // lattice) % lattice, s // (lattice * lattice)
        o = np.array([i, j, k], dtype=float) * spacing
        R = _quat_rot(rng.random(3))
        z += [8, 1, 1]
        xyz += [o, o + R @ h1, o + R @ h2]
    xyz = np.array(xyz)
    if noise > 0:
        xyz = xyz + rng.normal(0.0, noise, xyz.shape)
    return np.array(z, dtype=np.int32), xyz * AATOAU


def alkane(k, rng=None, noise=0.0):
    """All-trans CkH2k+2 zig-zag in the xy plane; returns (numbers, bohr)."""
    cc, ch = 1.53, 1.09
    half = TET / 2
    dx, dy = cc * math.sin(half), cc * math.cos(half)
    cpos = [np.array([i * dx, (i % 2) * dy, 0.0]) for i in range(-1, k + 1)]  # with virtual ends
    z, xyz = [], []
    for i in range(1, k + 1):
        c = cpos[i]
        a = cpos[i - 1] - c
        b = cpos[i + 1] - c
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        bis = -(a + b)
        bis /= np.linalg.norm(bis)
        hdirs = [bis * math.cos(half) + np.array([0, 0, 1.0]) * math.sin(half),
                 bis * math.cos(half) - np.array([0, 0, 1.0]) * math.sin(half)]
        if i == 1:
            hdirs.append(a)
        if i == k:
            hdirs.append(b)
        z.append(6)
        xyz.append(c)
        for h in hdirs:
            z.append(1)
            xyz.append(c + ch * h)
    xyz = np.array(xyz)
    if noise > 0:
        rng = np.random.default_rng(0) if rng is None else rng
        xyz = xyz + rng.normal(0.0, noise, xyz.shape)
    return np.array(z, dtype=np.int32), xyz * AATOAU


def alkane_bench(n):
    """axtb's benchmark alkane CnH2n+2 (bench/final_timings.py:97-107), bohr."""
    atoms = []
    for i in range(n):
        x, y, s = i * 1.24946158, (0.0 if i % 2 == 0 else 0.88303214), (-1.0 if i % 2 == 0 else 1.0)
        atoms += [(6, x, y, 0.0), (1, x, y + s * 0.51, 0.89), (1, x, y + s * 0.51, -0.89)]
        if i == 0:
            atoms.append((1, x - 1.03, y + s * 0.36, 0.0))
        if i == n - 1:
            atoms.append((1, x + 1.03, y + s * 0.36, 0.0))
    z = np.array([a[0] for a in atoms], dtype=np.int32)
    return z, np.array([a[1:] for a in atoms]) * AATOAU


def read_xyz(path):
    sym = {"H": 1, "C": 6, "N": 7, "O": 8, "F": 9, "S": 16, "P": 15, "Cl": 17}
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    z = np.array([sym[l.split()[0].capitalize()] for l in lines[2:2 + n]], dtype=np.int32)
    x = np.array([[float(v) for v in l.split()[1:4]] for l in lines[2:2 + n]])
    return z, x * AATOAU


def config(name, seed=0):
    """The named single-molecule workloads: water20, water200, water1000, c100h202, c30h62."""
    rng = np.random.default_rng(seed)
    if name == "water20":
        return water_cluster(20, 3, rng=rng)
    if name == "water200":
        return water_cluster(200, 6, rng=rng)
    if name == "water1000":
        return water_cluster(1000, 10, rng=rng)
    if name == "c100h202":
        return alkane(100, rng=rng, noise=0.02)
    if name == "c30h62":
        return alkane(30, rng=rng, noise=0.1)
    raise KeyError(name)


def batch(nmol=4096, seed=0, noise=0.05):
    """The uneven batch: per molecule, at random, (H2O)w with w in 3..20 (lattice ceil(w^(1/3)))
    or CkH2k+2 with k in 3..20, plus Gaussian noise of `noise` Å."""
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(nmol):
        if rng.random() < 0.5:
            w = int(rng.integers(3, 21))
            lat = math.ceil(w ** (1.0 / 3.0) - 1e-9)
            mols.append(water_cluster(w, lat, rng=rng, noise=noise))
        else:
            k = int(rng.integers(3, 21))
            mols.append(alkane(k, rng=rng, noise=noise))
    return mols


def conformers(n=2048, seed=0, k=30, noise=0.1):
    rng = np.random.default_rng(seed)
    return [alkane(k, rng=rng, noise=noise) for _ in range(n)]
