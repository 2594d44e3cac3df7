"""Orbital and density values on a grid (the reference's post-processing).

shelled_eval.py:27-133 builds a box grid around the molecule, evaluates the
basis on it in chunks of 1e6 points, forms the MOs and writes per-orbital
densities to ``preprocess/*.npy`` FP32 memmaps; main.py:880-924 writes the
total and spin densities from the density matrices. Here the AO values come
from the HIP kernel behind ``dft_ao_grid`` (``Plan.ao_values``) and the
contractions are plain GEMMs on the device; the files have the reference's
names, dtypes and shapes, so visualize.py reads them unchanged.
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch


def reference_grid(centers, spacing=0.008, y_padding=-0.75, z_padding=1.0):
    """The box of shelled_eval.py:27-57: x from min to max - 1.0, y and z
    padded, linspace axes of ceil(extent / spacing) + 1 points; returns the
    FP32 (npts, 3) grid (x slowest) and its (nx, ny, nz) shape."""
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    lo = [float(c[:, 0].min()), float(c[:, 1].min()) - y_padding, float(c[:, 2].min()) - z_padding]
    hi = [float(c[:, 0].max()) - 1.0, float(c[:, 1].max()) + y_padding, float(c[:, 2].max()) + z_padding]
    return box_grid(lo, hi, spacing)


def box_grid(lo, hi, spacing):
    """linspace box with ceil((hi - lo) / spacing) + 1 points per axis."""
    axes = [np.linspace(a, b, math.ceil((b - a) / spacing) + 1) for a, b in zip(lo, hi)]
    shape = tuple(len(a) for a in axes)
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).astype(np.float32)
    return g.reshape(-1, 3), shape


def _chunks(n, size):
    for a in range(0, n, size):
        yield a, min(a + size, n)


def write_orbital_densities(plan, grid, C_a, C_b, N_a, N_b, A=None, out_dir="preprocess", chunk=10**6):
    """shelled_eval.py:61-133. With nmo = max(N_a, N_b) and the first nmo
    columns of C_a, C_b (rounded to FP32 as there):
      grid.npy           FP32 (npts, 3)
      total_density.npy  FP32 (nmo, npts): occ_a |psi_a|^2 + occ_b |psi_b|^2 per orbital
      phase_density.npy  FP32 (nmo, npts): psi_a in rows < N_a
    ``A`` (n_sph x n_cart) maps Cartesian AO values to the spherical basis of
    C (shelled.py:878-920); None for a Cartesian C."""
    os.makedirs(out_dir, exist_ok=True)
    grid = np.ascontiguousarray(np.asarray(grid, dtype=np.float32).reshape(-1, 3))
    npts = grid.shape[0]
    nmo = int(max(N_a, N_b))
    dev = plan.device
    Ca = torch.as_tensor(np.asarray(C_a)[:, :nmo].astype(np.float32).T, device=dev).to(torch.float64)
    Cb = torch.as_tensor(np.asarray(C_b)[:, :nmo].astype(np.float32).T, device=dev).to(torch.float64)
    At = None if A is None else torch.as_tensor(np.asarray(A, dtype=np.float64), device=dev)
    np.save(os.path.join(out_dir, "grid.npy"), grid)
    total = np.lib.format.open_memmap(os.path.join(out_dir, "total_density.npy"), mode="w+", dtype=np.float32,
                                      shape=(nmo, npts))
    phase = np.lib.format.open_memmap(os.path.join(out_dir, "phase_density.npy"), mode="w+", dtype=np.float32,
                                      shape=(nmo, npts))
    occ_a = torch.zeros(nmo, dtype=torch.float64, device=dev)
    occ_b = torch.zeros_like(occ_a)
    occ_a[:int(N_a)] = 1.0
    occ_b[:int(N_b)] = 1.0
    for a, b in _chunks(npts, chunk):
        ao = plan.ao_values(grid[a:b])
        if At is not None:
            ao = At @ ao
        psi_a = Ca @ ao
        psi_b = Cb @ ao
        phase[:int(N_a), a:b] = psi_a[:int(N_a)].to(torch.float32).cpu().numpy()
        dens = occ_a[:, None] * psi_a * psi_a + occ_b[:, None] * psi_b * psi_b
        total[:, a:b] = dens.to(torch.float32).cpu().numpy()
    total.flush()
    phase.flush()
    return nmo, npts


def write_densities(plan, grid, P_a, P_b, C_a, A=None, out_dir="preprocess", chunk=10**6):
    """main.py:880-924: total_density.npy and spin_density.npy FP32 (npts,)
    with rho_s(r) = sum_mn P^s_mn phi_m(r) phi_n(r) (total = a + b, spin =
    a - b), and orbital_density.npy FP32 (norb, npts) = (C_a^T phi)^2 for
    every column of C_a. P and C are in the basis of ``A @ phi`` (Cartesian
    if A is None)."""
    os.makedirs(out_dir, exist_ok=True)
    grid = np.ascontiguousarray(np.asarray(grid, dtype=np.float32).reshape(-1, 3))
    npts = grid.shape[0]
    dev = plan.device
    Pa = torch.as_tensor(np.asarray(P_a, dtype=np.float64), device=dev)
    Pb = torch.as_tensor(np.asarray(P_b, dtype=np.float64), device=dev)
    CaT = torch.as_tensor(np.asarray(C_a, dtype=np.float64).T.copy(), device=dev)
    At = None if A is None else torch.as_tensor(np.asarray(A, dtype=np.float64), device=dev)
    np.save(os.path.join(out_dir, "grid.npy"), grid)
    tot = np.lib.format.open_memmap(os.path.join(out_dir, "total_density.npy"), mode="w+", dtype=np.float32,
                                    shape=(npts,))
    spin = np.lib.format.open_memmap(os.path.join(out_dir, "spin_density.npy"), mode="w+", dtype=np.float32,
                                     shape=(npts,))
    orb = np.lib.format.open_memmap(os.path.join(out_dir, "orbital_density.npy"), mode="w+", dtype=np.float32,
                                    shape=(CaT.shape[0], npts))
    for a, b in _chunks(npts, chunk):
        ao = plan.ao_values(grid[a:b])
        if At is not None:
            ao = At @ ao
        ra = ((Pa @ ao) * ao).sum(0)
        rb = ((Pb @ ao) * ao).sum(0)
        tot[a:b] = (ra + rb).to(torch.float32).cpu().numpy()
        spin[a:b] = (ra - rb).to(torch.float32).cpu().numpy()
        psi = CaT @ ao
        orb[:, a:b] = (psi * psi).to(torch.float32).cpu().numpy()
    for m in (tot, spin, orb):
        m.flush()
    return npts
