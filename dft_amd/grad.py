"""Analytic nuclear gradients behind torch.autograd.

``scf_energy(coords, Z, basis)`` is differentiable with respect to the
nuclear coordinates (bohr): its forward pass runs the SCF on the GPU, its
backward pass calls the derivative-integral kernels of libdft_hip.so
(``dft_grad``: derivative ERIs contracted with the density pair, one-electron
derivatives with P and the energy-weighted density W) and adds the nuclear
repulsion gradient.

The reference has no gradients; this module adds them (checked against
finite differences of the SCF energy).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from .basis import build_shells, build_shells_from_entries
from .engine import Plan
from .linalg import JACOBI_NMAX
from .scf import electron_counts, sad_guess, scf_diis, scf_diis_graph


def nuclear_repulsion_gradient(Z, xyz):
    """d/dR_A sum_{i<j} Z_i Z_j / |R_i - R_j| (shelled.py:157-163 differentiated)."""
    Z = np.asarray(Z, dtype=np.float64)
    R = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    d = R[:, None, :] - R[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(r, np.inf)
    return -np.sum((Z[:, None] * Z[None, :] / r ** 3)[:, :, None] * d, axis=1)


def shell_atoms(plan: Plan) -> np.ndarray:
    """Nucleus index of each plan segment (from the shells' atom_idx)."""
    out = []
    for sh, segs in zip(plan.shells, plan.segments):
        out.extend([sh["atom_idx"]] * len(segs))
    return np.ascontiguousarray(np.asarray(out, dtype=np.int32))


def electronic_gradient(plan: Plan, C_a, eps_a, N_a, C_b, eps_b, N_b):
    """sum P dH + sum Gamma d(ab|cd) - sum W dS from converged orbitals
    (Cartesian, device tensors): P_s = C_occ C_occ^T, W = sum_s C_occ eps C_occ^T."""
    Ca, Cb = C_a[:, :N_a], C_b[:, :N_b]
    P_a = (Ca @ Ca.T).contiguous()
    P_b = (Cb @ Cb.T).contiguous()
    P = (P_a + P_b).contiguous()
    W = ((Ca * eps_a[:N_a]) @ Ca.T + (Cb * eps_b[:N_b]) @ Cb.T).contiguous()
    g = torch.empty(3 * len(plan.Z), dtype=torch.float64, device=plan.device)
    rhf = torch.equal(P_a, P_b)
    sa = shell_atoms(plan)
    s = torch.cuda.current_stream(plan.device)
    _lib.check(plan.lib.dft_grad(plan.handle, _lib.ptr(sa, C.c_int32), P.data_ptr(), P_a.data_ptr(),
                                 None if rhf else P_b.data_ptr(), W.data_ptr(), g.data_ptr(),
                                 C.c_void_p(s.cuda_stream)))
    return g.reshape(-1, 3)


def _shells(Z, xyz, basis):
    if isinstance(basis, dict):
        return build_shells(basis, Z, xyz)
    return build_shells_from_entries(basis, xyz)  # per-atom entries


def energy_and_gradient(Z, xyz, basis, charge=0, mult=1, schwarz_tau=1e-12, e_tol=1e-11, p_tol=1e-9, guess=None):
    """(E, dE/dR [natom, 3], SCF result) at bohr coordinates ``xyz``.
    ``guess`` = (P_a, P_b) starts the SCF from those densities; None starts
    from the superposition of atomic densities (``scf.sad_guess``, atomic
    densities cached per element and basis), "core" from the core
    Hamiltonian (the reference's start). The converged energy and gradient
    do not depend on the start."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    shells = _shells(Z, xyz, basis)
    N_a, N_b = electron_counts(Z, charge, mult)
    if isinstance(guess, str):
        if guess != "core":
            raise ValueError("guess must be (P_a, P_b), None or 'core'")
        guess = None
    elif guess is None:
        guess = sad_guess(shells, Z, N_a, N_b)
    plan = Plan(shells, Z, xyz, schwarz_tau=schwarz_tau, fp32_tau=0.0, uhf=N_a != N_b)
    # undamped DIIS from the first iteration: the gradient needs a tightly
    # converged density, which damping (the reference's schedule) slows down.
    # Closed shells up to the eigensolver's order run the loop as one
    # captured hipGraph per iteration (same iteration as scf_diis).
    if N_a == N_b and plan.nbf <= JACOBI_NMAX:
        res = scf_diis_graph(plan, Z, xyz, N_a, e_tol=e_tol, p_tol=p_tol, max_count=300, guess=guess, rate=1.0,
                             diis_threshold=float("inf"))
    else:
        res = scf_diis(plan, Z, xyz, N_a, N_b, e_tol=e_tol, p_tol=p_tol, max_count=300, guess=guess, rate=1.0,
                       diis_threshold=float("inf"))
    if not res.converged:
        raise RuntimeError("SCF did not converge; the analytic gradient needs a stationary density")
    g = electronic_gradient(plan, res.C_a, res.eps_a, N_a, res.C_b, res.eps_b, N_b).cpu().numpy()
    return res.energy, g + nuclear_repulsion_gradient(Z, xyz), res


class SCFEnergy(torch.autograd.Function):
    """E(R) with an analytic backward (derivative-ERI kernels). ``state``
    (a dict or None) carries the converged densities from one call to the
    next as the SCF guess (geometry optimisation, finite differences)."""

    @staticmethod
    def forward(ctx, coords, Z, basis, charge, mult, state=None):
        guess = state.get("P") if state is not None else None
        E, g, res = energy_and_gradient(Z, coords.detach().cpu().numpy(), basis, charge, mult, guess=guess)
        if state is not None:
            state["P"] = (res.P_a, res.P_b)
            state["iterations"] = res.iterations
        ctx.save_for_backward(torch.as_tensor(g, dtype=coords.dtype, device=coords.device))
        return coords.new_tensor(E)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None, None, None


def scf_energy(coords: torch.Tensor, Z, basis, charge=0, mult=1, state=None) -> torch.Tensor:
    return SCFEnergy.apply(coords, Z, basis, charge, mult, state)


def steepest_descent(Z, xyz, basis, steps=10, step_size=0.2, charge=0, mult=1):
    """Fixed-step steepest descent in bohr; returns per-step (E, |g|max, s)."""
    coords = torch.tensor(np.asarray(xyz, dtype=np.float64).reshape(-1, 3), requires_grad=True)
    history = []
    state = {}  # previous step's densities seed the next SCF
    for _ in range(steps):
        t0 = time.perf_counter()
        E = scf_energy(coords, Z, basis, charge, mult, state)
        E.backward()
        with torch.no_grad():
            g = coords.grad.clone()
            coords -= step_size * g
        coords.grad = None
        history.append((float(E.detach()), float(g.abs().max()), time.perf_counter() - t0,
                        state.get("iterations")))
    return coords.detach().numpy(), history
