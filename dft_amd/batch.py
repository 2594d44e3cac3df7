"""Batched SCF over many independent molecules.

One ``BatchPlan`` holds every molecule's pair data and all quartet classes of
all molecules; one J/K build (captured once as a hipGraph and relaunched each
iteration) serves the whole batch. The per-molecule linear algebra (S^-1/2,
Fock diagonalisation, density, energy) stays FP64 and is batched by padding
every molecule to the largest basis: the padded block of S is the identity
and that of F a large diagonal, so padded orbitals are never occupied.
Reference loop per molecule: main.py:753-828 (plain Roothaan).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .engine import BatchPlan, Graph
from .scf import nuclear_repulsion

PAD_DIAG = 1.0e4


@dataclass
class BatchResult:
    energies: np.ndarray
    converged: np.ndarray
    iterations: int
    fock_times: list = field(default_factory=list)
    iter_times: list = field(default_factory=list)
    P: object = None   # packed total density after the last iteration
    Pa: object = None  # packed alpha density after the last iteration


class _Pad:
    def __init__(self, plan: BatchPlan):
        self.plan = plan
        self.nmol = len(plan.nbfs)
        self.nmax = max(plan.nbfs)
        dev = plan.device
        idx_m, idx_r, idx_c, flat = [], [], [], []
        for m, n in enumerate(plan.nbfs):
            r, c = np.divmod(np.arange(n * n), n)
            idx_m.append(np.full(n * n, m))
            idx_r.append(r)
            idx_c.append(c)
            flat.append(plan.offsets[m] + np.arange(n * n))
        self.m = torch.as_tensor(np.concatenate(idx_m), device=dev)
        self.r = torch.as_tensor(np.concatenate(idx_r), device=dev)
        self.c = torch.as_tensor(np.concatenate(idx_c), device=dev)
        self.flat = torch.as_tensor(np.concatenate(flat), device=dev)
        mask = torch.ones((self.nmol, self.nmax), dtype=torch.float64, device=dev)
        for m, n in enumerate(plan.nbfs):
            mask[m, :n] = 0.0
        self.pad_mask = mask  # 1 on padded diagonal entries

    def pad(self, flat, diag=0.0):
        out = torch.zeros((self.nmol, self.nmax, self.nmax), dtype=torch.float64, device=flat.device)
        out[self.m, self.r, self.c] = flat[self.flat]
        if diag:
            out = out + torch.diag_embed(self.pad_mask * diag)
        return out

    def unpad(self, padded, out=None):
        out = torch.empty(self.plan.nelem, dtype=torch.float64, device=padded.device) if out is None else out
        out[self.flat] = padded[self.m, self.r, self.c]
        return out


def scf_batch(plan: BatchPlan, n_occ, mode="fp64", use_graph=True, e_tol=1e-8, p_tol=1e-6, max_iter=100,
              fixed_iters=None, diis=True, history=10):
    """Closed-shell SCF of every molecule of the batch (N_alpha = N_beta = n_occ[m]).

    With ``fixed_iters`` the loop runs exactly that many iterations (for
    timing); otherwise until every molecule has |dE| < e_tol and
    max|dP| < p_tol, or max_iter. ``diis`` extrapolates each molecule's Fock
    matrix from the commutator residuals FPS - SPF as shelled.py:1196-1203
    (B c = 1 without a Lagrange row, c normalised to sum 1), from the second
    iteration on; the energy is always that of the un-extrapolated F.
    """
    if plan.uhf:
        raise ValueError("scf_batch is closed-shell: build the BatchPlan with uhf=False")
    pad = _Pad(plan)
    S, T, V = plan.one_electron()
    H = T + V
    Sp = pad.pad(S, diag=1.0)
    w, U = torch.linalg.eigh(Sp)
    X = U @ torch.diag_embed(w.rsqrt()) @ U.transpose(1, 2)
    Hp = pad.pad(H, diag=PAD_DIAG)
    occ = torch.zeros((pad.nmol, pad.nmax), dtype=torch.float64, device=plan.device)
    for m, n in enumerate(n_occ):
        occ[m, :int(n)] = 1.0
    E_NN = torch.as_tensor([nuclear_repulsion(Z, xyz) for _, Z, xyz in plan.molecules], dtype=torch.float64,
                           device=plan.device)

    def density(Fp):
        _, Cp = torch.linalg.eigh(X.transpose(1, 2) @ Fp @ X)
        Cm = X @ Cp
        return (Cm * occ[:, None, :]) @ Cm.transpose(1, 2)

    Pa_p = density(Hp)
    Pa = pad.unpad(Pa_p)
    P = (2 * Pa).contiguous()
    J, K = plan.empty(), plan.empty()
    graph = Graph(plan, P, Pa, None, J, K, None, mode) if use_graph else None
    E_old = torch.full((pad.nmol,), -1e9, dtype=torch.float64, device=plan.device)
    fock_t, iter_t = [], []
    hist_F, hist_R = [], []
    it = 0
    converged = torch.zeros(pad.nmol, dtype=torch.bool, device=plan.device)
    while True:
        t0 = time.perf_counter()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        if graph is not None:
            graph.launch()
        else:
            plan.jk_into(P, Pa, None, J, K, None, mode)
        e1.record()
        F = H + J - K
        Fp = pad.pad(F, diag=PAD_DIAG)
        Pp = pad.pad(P)
        Pap = pad.pad(Pa)
        E = 0.5 * (Pp * Hp).sum((1, 2)) + (Pap * Fp).sum((1, 2)) + E_NN
        Fd = Fp
        if diis:
            R = Fp @ Pap @ Sp - Sp @ Pap @ Fp
            hist_F.append(Fp)
            hist_R.append(R)
            del hist_F[:-history], hist_R[:-history]
            if len(hist_F) > 1:
                Rs = torch.stack(hist_R)
                B = torch.einsum("imab,jmab->mij", Rs, Rs)
                lam = 1e-14 * B.diagonal(dim1=1, dim2=2).amax(1).clamp_min(1e-300)
                B = B + lam[:, None, None] * torch.eye(B.shape[1], dtype=B.dtype, device=B.device)
                c = torch.linalg.solve(B, torch.ones(B.shape[:2], dtype=B.dtype, device=B.device))
                c = c / c.sum(1, keepdim=True)
                Fd = torch.einsum("mi,imab->mab", c, torch.stack(hist_F))
        Pa_new_p = density(Fd)
        dP = (Pa_new_p - Pap).abs().amax((1, 2))
        dE = (E - E_old).abs()
        e1.synchronize()
        fock_t.append(e0.elapsed_time(e1) * 1e-3)
        converged = (dE < e_tol) & (dP < p_tol)
        it += 1
        done = (fixed_iters is not None and it >= fixed_iters) or \
               (fixed_iters is None and (bool(converged.all()) or it >= max_iter))
        # new density goes into the buffers the graph reads
        pad.unpad(Pa_new_p, out=Pa)
        torch.mul(Pa, 2.0, out=P)
        E_old = E
        torch.cuda.synchronize()
        iter_t.append(time.perf_counter() - t0)
        if done:
            return BatchResult(E.cpu().numpy(), converged.cpu().numpy(), it, fock_t, iter_t, P.clone(), Pa.clone())


def water_batch(n_mol=64, seed=2):
    """The 64 (H2O)_n clusters (n ~ U{1..24}, seed 2) on a 3x3x3 lattice, STO-3G."""
    from .basis import build_shells, load_basis
    from .geometry import batch_of_clusters
    basis = load_basis()
    mols = []
    for Z, xyz in batch_of_clusters(n_mol, seed):
        mols.append((build_shells(basis, Z, xyz), Z, xyz))
    return mols
