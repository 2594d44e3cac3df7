"""Synthetic molecules of the benchmark workloads.

All generators work in Angstrom and return bohr (CODATA-2018:
1 bohr = 0.529177210903 A). The reference itself takes coordinates as bohr
with no conversion (``shelled.py:38``, ``main.py:38``), so the bohr arrays
returned here are what its drivers would be given.

Random draws use ``numpy.random.default_rng(seed)`` (PCG64) in a fixed order,
and random rotations are uniform unit quaternions (Shoemake 1992).
"""
from __future__ import annotations

import numpy as np

BOHR_PER_ANG = 1.0 / 0.529177210903

# water internal geometry: O at the origin, H at (+-0.757, 0.586, 0) A
WATER_ANG = np.array([[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]])
WATER_Z = [8, 1, 1]


def quaternion_rotation(rng) -> np.ndarray:
    """Uniform random rotation matrix from three uniforms (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    w, x, y, z = (b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2),
                  a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3))
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def water():
    """Single water, O at the origin (Z list, bohr coordinates)."""
    return list(WATER_Z), WATER_ANG * BOHR_PER_ANG


def lattice_sites(shape, spacing):
    nx, ny, nz = shape
    return np.array([[i * spacing, j * spacing, k * spacing]
                     for i in range(nx) for j in range(ny) for k in range(nz)])


def water_cluster(shape=(4, 2, 2), spacing=3.1, seed=0, sites=None, rng=None):
    """(H2O)_n on lattice sites, each water rotated about O, then translated.

    One rotation is drawn per site in i-major site order. With ``sites`` the
    cluster uses that subset of the lattice (in the given order).
    """
    rng = np.random.default_rng(seed) if rng is None else rng
    all_sites = lattice_sites(shape, spacing)
    use = all_sites if sites is None else all_sites[np.asarray(sites)]
    Z, xyz = [], []
    for site in use:
        R = quaternion_rotation(rng)
        xyz.append(WATER_ANG @ R.T + site)
        Z.extend(WATER_Z)
    return Z, np.concatenate(xyz) * BOHR_PER_ANG


def alkane(n_carbon: int, cc=1.54, ch=1.09):
    """All-trans CnH(2n+2): planar C zig-zag, tetrahedral angles, staggered CH3.

    Returns (Z list, Angstrom coordinates); atoms are ordered C1..Cn, then the
    hydrogens of C1, C2, ... in turn.
    """
    half = np.deg2rad(109.4712206) / 2.0
    dx, dy = cc * np.sin(half), cc * np.cos(half)
    C = np.array([[k * dx, (k % 2) * dy, 0.0] for k in range(n_carbon)])
    H = []
    zhat = np.array([0.0, 0.0, 1.0])
    for k in range(n_carbon):
        nbrs = [C[j] - C[k] for j in (k - 1, k + 1) if 0 <= j < n_carbon]
        units = [v / np.linalg.norm(v) for v in nbrs]
        if len(units) == 2:
            b = units[0] + units[1]
            b /= np.linalg.norm(b)
            n = np.cross(units[0], units[1])
            n /= np.linalg.norm(n)
            ca, sa = np.cos(half), np.sin(half)
            for s in (1.0, -1.0):
                H.append(C[k] + ch * (-b * ca + s * n * sa))
        else:
            u = units[0]
            # H anti to the next C-C bond along the chain (dihedral 180 deg)
            j2 = 2 if k == 0 else n_carbon - 3
            if 0 <= j2 < n_carbon and n_carbon > 2:
                w = C[j2] - C[k + (1 if k == 0 else -1)]
            else:
                w = np.array([0.0, 1.0, 0.0])
            w = w - np.dot(w, u) * u
            e1 = -w / np.linalg.norm(w)
            e2 = np.cross(u, e1)
            for phi in (0.0, 2 * np.pi / 3, 4 * np.pi / 3):
                d = -u / 3.0 + np.sqrt(8.0 / 9.0) * (np.cos(phi) * e1 + np.sin(phi) * e2)
                H.append(C[k] + ch * d)
            if n_carbon == 1:
                H.append(C[k] + ch * u * -1.0)
    Z = [6] * n_carbon + [1] * len(H)
    return Z, np.vstack([C, np.array(H)])


def alkane_bohr(n_carbon: int, noise_sigma_ang=0.0, seed=None):
    Z, xyz = alkane(n_carbon)
    if noise_sigma_ang:
        rng = np.random.default_rng(seed)
        xyz = xyz + rng.normal(0.0, noise_sigma_ang, size=xyz.shape)
    return Z, xyz * BOHR_PER_ANG


def synthetic_spd_basis(Z, seed=1, nprim_range=(3, 6), exp_range=(0.1, 1e4)):
    """Per-atom synthetic s/p/d basis in the reference JSON format.

    Per C: 3 s + 2 p + 1 d shells; per H: 2 s shells. Each shell draws, in
    this order, nprim ~ U{3..6}, exponents log-uniform in [0.1, 1e4] and
    coefficients U[0.1, 1] scaled to unit 2-norm. Draw order is atom, then
    shell. Returned as one basis entry list per atom (atoms may share Z but
    get different shells, so the result is keyed by atom index).
    """
    rng = np.random.default_rng(seed)
    lo, hi = np.log(exp_range[0]), np.log(exp_range[1])
    per_atom = []
    for z in Z:
        ls = [0, 0, 0, 1, 1, 2] if int(z) == 6 else [0, 0]
        entries = []
        for l in ls:
            npr = int(rng.integers(nprim_range[0], nprim_range[1] + 1))
            ex = np.exp(rng.uniform(lo, hi, size=npr))
            co = rng.uniform(0.1, 1.0, size=npr)
            co /= np.linalg.norm(co)
            entries.append([[l], ex.tolist(), [co.tolist()]])
        per_atom.append(entries)
    return per_atom


def batch_of_clusters(n_mol=64, seed=2, lattice=(3, 3, 3), spacing=3.1, nmax=24):
    """n_mol water clusters (H2O)_n, n ~ U{1..nmax}, on random lattice subsets.

    Draw order per molecule: n, then the site subset (without replacement),
    then one rotation per chosen site.
    """
    rng = np.random.default_rng(seed)
    nsite = int(np.prod(lattice))
    mols = []
    for _ in range(n_mol):
        n = int(rng.integers(1, nmax + 1))
        sites = rng.choice(nsite, size=n, replace=False)
        mols.append(water_cluster(lattice, spacing, sites=sites, rng=rng))
    return mols


def methane():
    """CH4, tetrahedral, C-H 1.09 A (Z list, bohr)."""
    d = 1.09 / np.sqrt(3.0)
    xyz = np.array([[0, 0, 0], [d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d]])
    return [6, 1, 1, 1, 1], xyz * BOHR_PER_ANG
