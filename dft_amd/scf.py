"""SCF drivers on top of the HIP Fock build.

``scf_plain`` is the plain UHF Roothaan loop of ``main.py:753-828`` (core
Hamiltonian guess, symmetric orthogonalisation without cutoff, aufbau
occupation, convergence on |dE| < 1e-8 and max|dP| < 1e-6). ``scf_diis`` is
the damped + DIIS loop of ``shelled.py:1103-1292``. J/K come from
``Plan.jk`` (fused direct build on the GPU); orthogonalisation,
diagonalisation, density build and energy stay FP64 on the GPU: the
eigendecompositions go through the library's Jacobi kernel (``linalg.eigh``,
started from the previous iteration's orbitals), the rest through torch. RHF (N_alpha == N_beta) uses one exchange matrix, since then
P_a = P_b and K_a = K_b exactly.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .engine import Plan
from .linalg import eigh


def nuclear_repulsion(Z, centers) -> float:
    """sum_{i<j} Z_i Z_j / |R_i - R_j| (shelled.py:157-163)."""
    Z = np.asarray(Z, dtype=np.float64)
    R = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(R[:, None, :] - R[None, :, :], axis=-1)
    iu = np.triu_indices(len(Z), 1)
    return float(np.sum(Z[iu[0]] * Z[iu[1]] / d[iu]))


def init_UHF_density(C_a, C_b, N_a, N_b):
    """P_sigma = C_occ C_occ^T (shelled.py:549-559)."""
    Ca = C_a[:, :int(N_a)]
    Cb = C_b[:, :int(N_b)]
    return Ca @ Ca.T, Cb @ Cb.T


def UHF_density(C_a, C_b, N_a, N_b, old_P_a, old_P_b, rate):
    """Damped density, rate * new + (1 - rate) * old (shelled.py:561-574)."""
    P_a, P_b = init_UHF_density(C_a, C_b, N_a, N_b)
    return rate * P_a + (1 - rate) * old_P_a, rate * P_b + (1 - rate) * old_P_b


@dataclass
class SCFResult:
    energy: float
    converged: bool
    iterations: int
    P_a: torch.Tensor
    P_b: torch.Tensor
    F_a: torch.Tensor
    F_b: torch.Tensor
    C_a: torch.Tensor
    C_b: torch.Tensor
    eps_a: torch.Tensor
    eps_b: torch.Tensor
    fock_times: list = field(default_factory=list)   # seconds per Fock build (device events)
    iter_times: list = field(default_factory=list)   # seconds per full iteration (wall)
    energies: list = field(default_factory=list)


def electron_counts(Z, charge=0, mult=1):
    n = int(round(float(np.sum(Z)))) - charge
    N_a = (n + mult - 1) // 2
    N_b = (n - mult + 1) // 2
    return N_a, N_b


class _Fock:
    """J/K on preallocated device buffers, timed with HIP events.

    With ``A`` (spherical x Cartesian, from shelled.spherical_matrix) the
    densities arrive in the spherical basis, are taken to Cartesian as
    A^T P A for the build and J/K return as A J A^T (shelled.py:1166-1177).
    """

    def __init__(self, plan: Plan, mode: str, rhf: bool, A=None):
        self.plan, self.mode, self.rhf, self.A = plan, mode, rhf, A
        n = plan.nbf
        dev = plan.device
        self.J = torch.empty((n, n), dtype=torch.float64, device=dev)
        self.Ka = torch.empty_like(self.J)
        self.Kb = None if rhf else torch.empty_like(self.J)
        self.times = []

    def __call__(self, P, P_a, P_b):
        A = self.A
        if A is not None:
            P, P_a, P_b = (A.T @ x @ A for x in (P, P_a, P_b))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        if self.rhf:
            self.plan.jk_into(P.contiguous(), P_a.contiguous(), None, self.J, self.Ka, None, self.mode)
        else:
            self.plan.jk_into(P.contiguous(), P_a.contiguous(), P_b.contiguous(), self.J, self.Ka, self.Kb,
                              self.mode)
        e1.record()
        e1.synchronize()
        self.times.append(e0.elapsed_time(e1) * 1e-3)
        J, Ka, Kb = self.J, self.Ka, (self.Ka if self.rhf else self.Kb)
        if A is not None:
            J, Ka, Kb = (A @ x @ A.T for x in (J, Ka, Kb))
        return J, Ka, Kb


def _setup(plan: Plan, Z, centers, A=None):
    S, T, V = plan.one_electron()
    H = T + V
    if A is not None:
        S, H = A @ S @ A.T, A @ H @ A.T
    w, U = eigh(S)
    X = U @ torch.diag(w.rsqrt()) @ U.T
    return S, H, X, nuclear_repulsion(Z, centers)


def _gram(R):
    """B_ij = sum_ab R_i[a,b] R_j[a,b] for a stack of residuals. Small stacks
    go through one elementwise product + reduction (a GEMM with K = n^2 and
    a 10 x 10 output runs on a handful of workgroups)."""
    h, n = R.shape[0], R.shape[1]
    if h * h * n * n <= (1 << 22):
        return (R.unsqueeze(1) * R.unsqueeze(0)).sum(dim=(2, 3))
    return torch.einsum("iab,jab->ij", R, R)


def scf_plain(plan: Plan, Z, centers, N_a, N_b, mode="fp64", e_tol=1e-8, p_tol=1e-6, max_count=100,
              p_rms_tol=None, A=None):
    """Plain UHF loop of main.py:753-828; RHF fast path when N_a == N_b.

    Exits when |dE| < e_tol and max|dP| < p_tol (and RMS dP < p_rms_tol if
    given), or after count > max_count with ``converged`` False (the
    reference silently exits there, main.py:806).
    """
    rhf = N_a == N_b and not plan.uhf
    if not rhf and not plan.uhf:
        raise ValueError("N_a != N_b needs a plan built with uhf=True (two exchange matrices)")
    S, H, X, E_NN = _setup(plan, Z, centers, A)
    _, Cp = eigh(X.T @ H @ X)
    C = X @ Cp
    P_a, P_b = init_UHF_density(C, C, N_a, N_b)
    fock = _Fock(plan, mode, rhf, A)
    Qa = Qb = Cp  # orbitals of the last diagonalisation: the next one starts from them
    E_total = -1000.0
    count = 0
    energies, iter_times = [], []
    while True:
        t0 = time.perf_counter()
        P = P_a + P_b
        J, Ka, Kb = fock(P, P_a, P_b)
        Fa = H + J - Ka
        Fb = H + J - Kb
        if not bool(torch.isfinite(Fa).all()):
            raise FloatingPointError(f"non-finite Fock matrix at iteration {count}")
        E_new = float(0.5 * torch.sum(P * H + P_a * Fa + P_b * Fb)) + E_NN
        dE = abs(E_new - E_total)
        ea, Qa = eigh(X.T @ Fa @ X, Qa)
        if rhf:
            eb, Qb = ea, Qa
        else:
            eb, Qb = eigh(X.T @ Fb @ X, Qb)
        Ca, Cb = X @ Qa, X @ Qb
        Pa_new, Pb_new = init_UHF_density(Ca, Cb, N_a, N_b)
        dPa, dPb = Pa_new - P_a, Pb_new - P_b
        dP = float(torch.maximum(dPa.abs().max(), dPb.abs().max()))
        dP_rms = float(torch.sqrt(0.5 * (dPa.square().mean() + dPb.square().mean())))
        energies.append(E_new)
        iter_times.append(time.perf_counter() - t0)
        ok = dE < e_tol and dP < p_tol and (p_rms_tol is None or dP_rms < p_rms_tol)
        if ok or count > max_count:
            return SCFResult(E_new, ok, count, P_a, P_b, Fa, Fb, Ca, Cb, ea, eb, fock.times, iter_times,
                             energies)
        E_total = E_new
        P_a, P_b = Pa_new, Pb_new
        count += 1


# Keys of the reference's SCF checkpoint (shelled.py:1220-1237, read back at
# :1136-1152).
CHECKPOINT_KEYS = ("count", "switch_to_DIIS", "DIIS_P_a_history", "DIIS_P_b_history", "DIIS_R_a_history",
                   "DIIS_R_b_history", "DIIS_Fock_a_history", "DIIS_Fock_b_history", "rate", "strikes", "P_a",
                   "P_b", "delta_E", "delta_P", "prev_delta_E", "prev_delta_P")


def _history_array(items, history, n):
    """Newest-first history list as the reference's zero-padded
    (history, n, n) array."""
    out = np.zeros((history, n, n))
    for k, v in enumerate(items[:history]):
        out[k] = v.cpu().numpy()
    return out


def scf_diis(plan: Plan, Z, centers, N_a, N_b, mode="fp64", rate=0.4, diis_threshold=1e-3, history=10,
             e_tol=1e-8, p_tol=1e-6, max_count=500, A=None, guess=None, checkpoint=None, resume=None,
             eig_tol=0.0):
    """Damped + DIIS UHF loop of shelled.py:1103-1292.

    Damping rate 0.4, divided by 1.1 every fourth strike (an iteration whose
    dE or dP grew), DIIS on the commutator residuals FPS - SPF from the
    first iteration with dE < 1e-3, history 10, no Lagrange row and the
    coefficients normalised to sum 1 (shelled.py:1196-1201). Unlike the
    reference it stops after ``max_count`` iterations with converged False.
    ``guess`` = (P_a, P_b) starts from those densities instead of the core
    Hamiltonian (e.g. the converged densities of a nearby geometry).
    ``eig_tol`` > 0: per-iteration eigensolves to that off-diagonal size, as
    in scf_diis_graph (0, the default, is exact). For RHF
    (N_a == N_b, P_a == P_b throughout) one diagonalisation serves both spins.
    ``checkpoint``: path written every 10th iteration in the reference's
    ``scf_checkpoint.npz`` layout (shelled.py:1219-1237); ``resume``: such a
    file to continue from (counter, DIIS state, damping and densities, as
    shelled.py:1136-1162; the energy of the previous iteration is not in it,
    so the first dE after a resume is infinite there too).
    """
    S, H, X, E_NN = _setup(plan, Z, centers, A)
    n = S.shape[0]
    if resume is not None:
        with np.load(resume) as ck:
            state = {k: ck[k] for k in CHECKPOINT_KEYS}
            if "E_total" in ck.files:
                state["E_total"] = ck["E_total"]
        guess = (state["P_a"], state["P_b"])
    Qa = Qb = None  # orbitals of the last diagonalisation: the next one starts from them
    if guess is not None:
        P_a, P_b = (torch.as_tensor(g, dtype=torch.float64, device=plan.device).clone() for g in guess)
    else:
        _, Cp = eigh(X.T @ H @ X)
        C = X @ Cp
        P_a, P_b = init_UHF_density(C, C, N_a, N_b)
        Qa = Qb = Cp
    rhf = N_a == N_b and not plan.uhf
    if not rhf and not plan.uhf:
        raise ValueError("N_a != N_b needs a plan built with uhf=True (two exchange matrices)")
    fock = _Fock(plan, mode, rhf, A)
    hist = {k: [] for k in ("Pa", "Pb", "Fa", "Fb", "Ra", "Rb")}
    E_total = None
    prev_dE = prev_dP = float("inf")
    strikes = 0
    use_diis = False
    count = 0
    if resume is not None:
        # the file holds iteration `count` before its damping / DIIS-switch
        # update (as written by shelled.py:1219-1237): apply that update and
        # go on with the next iteration
        count = int(state["count"])
        use_diis = bool(state["switch_to_DIIS"])
        rate = float(state["rate"])
        strikes = int(state["strikes"])
        dE, dP = float(state["delta_E"]), float(state["delta_P"])
        if (dE > float(state["prev_delta_E"]) or dP > float(state["prev_delta_P"])) and not use_diis:
            if strikes == 3:
                rate /= 1.1
                strikes = 0
            else:
                strikes += 1
        if dE < diis_threshold:
            use_diis = True
        prev_dE, prev_dP = dE, dP
        E_total = float(state["E_total"]) if "E_total" in state else None
        keep = min(count + 1, history)
        count += 1
        for k, key in (("Pa", "DIIS_P_a_history"), ("Pb", "DIIS_P_b_history"), ("Fa", "DIIS_Fock_a_history"),
                       ("Fb", "DIIS_Fock_b_history"), ("Ra", "DIIS_R_a_history"), ("Rb", "DIIS_R_b_history")):
            hist[k] = [torch.as_tensor(m, dtype=torch.float64, device=plan.device) for m in state[key][:keep]]
    energies, iter_times = [], []
    gram = _gram(torch.stack(hist["Ra"])) + _gram(torch.stack(hist["Rb"])) if hist["Ra"] else None
    while True:
        t0 = time.perf_counter()
        P = P_a + P_b
        J, Ka, Kb = fock(P, P_a, P_b)
        Fa = H + J - Ka
        Fb = H + J - Kb
        if not bool(torch.isfinite(Fa).all() & torch.isfinite(Fb).all()):
            raise FloatingPointError(f"non-finite Fock matrix at iteration {count}")
        for k, v in (("Pa", P_a), ("Pb", P_b), ("Fa", Fa), ("Fb", Fb), ("Ra", Fa @ P_a @ S - S @ P_a @ Fa),
                     ("Rb", Fb @ P_b @ S - S @ P_b @ Fb)):
            hist[k].insert(0, v)
            del hist[k][history:]
        # Gram matrix of the residual history (newest first), kept from one
        # iteration to the next: only the new residual's row is computed,
        # as elementwise products and sums (a GEMM with K = n^2 and a 10 x 10
        # output runs on a handful of workgroups)
        d = sum((torch.stack(hist[k]) * hist[k][0]).sum((1, 2)) for k in ("Ra", "Rb"))
        Bn = torch.empty((d.numel(), d.numel()), dtype=d.dtype, device=d.device)
        if gram is not None and d.numel() > 1:
            Bn[1:, 1:] = gram[:d.numel() - 1, :d.numel() - 1]
        Bn[0, :] = d
        Bn[:, 0] = d
        gram = Bn
        if use_diis:
            B = gram
            try:
                c = torch.linalg.solve(B, torch.ones(B.shape[0], dtype=B.dtype, device=B.device))
                c = c / c.sum()
            except RuntimeError:  # singular: residuals that no longer differ
                c = None
            # without a (finite) DIIS solution the iteration keeps its own
            # Fock matrix, as the device loop does
            if c is not None and bool(torch.isfinite(c).all()):
                Fa = torch.einsum("i,iab->ab", c, torch.stack(hist["Fa"]))
                Fb = torch.einsum("i,iab->ab", c, torch.stack(hist["Fb"]))
        E_new = float(0.5 * torch.sum(P * H + P_a * Fa + P_b * Fb)) + E_NN
        dE = float("inf") if E_total is None else abs(E_new - E_total)
        _, Qa = eigh(X.T @ Fa @ X, Qa, tol=eig_tol)
        Qb = Qa if rhf else eigh(X.T @ Fb @ X, Qb, tol=eig_tol)[1]
        Ca, Cb = X @ Qa, X @ Qb
        Pa_new, Pb_new = UHF_density(Ca, Cb, N_a, N_b, P_a, P_b, rate)
        dP = float(torch.maximum((Pa_new - P_a).abs().max(), (Pb_new - P_b).abs().max()))
        energies.append(E_new)
        iter_times.append(time.perf_counter() - t0)  # wall, synchronised by the float() reads
        if checkpoint is not None and count != 0 and count % 10 == 0:
            np.savez(checkpoint, count=count, switch_to_DIIS=use_diis,
                     DIIS_P_a_history=_history_array(hist["Pa"], history, n),
                     DIIS_P_b_history=_history_array(hist["Pb"], history, n),
                     DIIS_R_a_history=_history_array(hist["Ra"], history, n),
                     DIIS_R_b_history=_history_array(hist["Rb"], history, n),
                     DIIS_Fock_a_history=_history_array(hist["Fa"], history, n),
                     DIIS_Fock_b_history=_history_array(hist["Fb"], history, n),
                     rate=rate, strikes=strikes, P_a=Pa_new.cpu().numpy(), P_b=Pb_new.cpu().numpy(),
                     delta_E=dE, delta_P=dP, prev_delta_E=prev_dE, prev_delta_P=prev_dP,
                     # not in the reference's file: the energy of this
                     # iteration, so that a resume continues exactly
                     E_total=E_new)
        if (dE > prev_dE or dP > prev_dP) and not use_diis:
            if strikes == 3:
                rate /= 1.1
                strikes = 0
            else:
                strikes += 1
        if dE < diis_threshold:
            use_diis = True
        if (dE < e_tol and dP < p_tol) or count >= max_count:
            # final rebuild from the last density, as shelled.py:1261-1285
            P = P_a + P_b
            J, Ka, Kb = fock(P, P_a, P_b)
            Fa = H + J - Ka
            Fb = H + J - Kb
            E_fin = float(0.5 * torch.sum(P * H + P_a * Fa + P_b * Fb)) + E_NN
            ea, Ca = eigh(X.T @ Fa @ X, Qa)
            eb, Cb = (ea, Ca) if rhf else eigh(X.T @ Fb @ X, Qb)
            return SCFResult(E_fin, dE < e_tol and dP < p_tol, count, P_a, P_b, Fa, Fb, X @ Ca, X @ Cb, ea, eb,
                             fock.times, iter_times, energies)
        E_total = E_new
        prev_dE, prev_dP = dE, dP
        P_a, P_b = Pa_new, Pb_new
        count += 1


# Indices of the device SCF state (include/dft_hip.h, DFT_SCF_ST_*).
ST_ESUM, ST_ENN, ST_E, ST_HAS_E, ST_DE, ST_DP, ST_PREV_DE, ST_PREV_DP, ST_RATE, ST_STRIKES, ST_USE_DIIS, \
    ST_COUNT, ST_CONV, ST_ETOL, ST_PTOL, ST_DIIS_TH, ST_ENEW, ST_SLOT = range(18)
SCF_STATE_LEN = 1152


def scf_diis_graph(plan: Plan, Z, centers, N_occ, mode="fp64", rate=0.4, diis_threshold=1e-3, history=10,
                   e_tol=1e-8, p_tol=1e-6, max_count=500, guess=None, eig_tol=0.0):
    """The damped + DIIS loop of shelled.py:1103-1292 for a closed shell
    (N_alpha = N_beta = N_occ), one iteration = one replay of a captured
    hipGraph: the direct J/K build (dft_jk, its class launches on four
    streams), F = H + J - K, the residual F P S - S P F (GEMMs), the DIIS
    stage (dft_scf_diis_stage: history ring, Gram row, DIIS solve,
    extrapolation, energy sum), X^T F X in the previous orbitals' basis
    (dft_scf_gemm, dft_scf_fock_residual, dft_scf_orbital_matrix), the
    Jacobi eigensolver (dft_syevj_ws, warm start), the aufbau density and
    damping / strikes / DIIS switch / convergence (dft_scf_damp_control).
    The loop state stays on the device; the host reads it from a pinned
    copy while the next replay runs. The exit rebuild of
    shelled.py:1261-1285 runs eagerly after the loop. Orders up to 128
    (the eigensolver's limit); no checkpoint files (scf_diis writes them).
    ``guess`` = P_a (or (P_a, P_b) with P_a == P_b) starts from that density
    instead of the core Hamiltonian's; the first diagonalisation then starts
    from the core Hamiltonian's orbitals. ``eig_tol`` > 0 lets the
    per-iteration eigensolves stop at off-diagonal entries of that size
    (dft_syevj_tol; fewer Jacobi sweeps, orbitals to about eig_tol / gap):
    the converged state is the same, the intermediate energies no longer
    follow the reference loop's exact eigh to rounding (0, the default,
    keeps them exact); the exit rebuild's eigensolve is always exact. Keep
    eig_tol at or below about 1e-3 p_tol: a looser one leaves density errors
    near p_tol and slows the convergence ((H2O)16 at p_tol = 1e-8 with
    eig_tol = 1e-9: 193 iterations instead of 55)."""
    import ctypes as C
    from . import _lib
    from .linalg import scf_fock_residual_, scf_gemm_, scf_orbital_matrix_, syevj_
    if plan.uhf:
        raise ValueError("scf_diis_graph is closed-shell: build the plan with uhf=False")
    n = plan.nbf
    if n > 128:
        raise ValueError("scf_diis_graph needs nbf <= 128 (the Jacobi eigensolver's limit); use scf_diis")
    if not 1 <= history <= 16:
        raise ValueError("history must be in [1, 16]")
    lib = _lib.load()
    dev = plan.device
    S, H, X, E_NN = _setup(plan, Z, centers)
    f64 = dict(dtype=torch.float64, device=dev)
    _, Q = eigh(X.T @ H @ X)
    if guess is not None:
        g = guess[0] if isinstance(guess, (tuple, list)) else guess
        Pa = torch.as_tensor(g, **f64).clone().contiguous()
        if Pa.shape != (n, n):
            raise ValueError(f"guess density must be {n} x {n}")
    else:
        C0 = X @ Q
        Pa = (C0[:, :N_occ] @ C0[:, :N_occ].T).contiguous()
    P = (2 * Pa).contiguous()
    Q = Q.contiguous()
    J, K, F, R, Fu, PS, XQ, T, Cm, Pauf = (torch.empty((n, n), **f64) for _ in range(10))
    histF = torch.zeros((history, n, n), **f64)
    histR = torch.zeros((history, n, n), **f64)
    G = torch.zeros((history, history), **f64)
    A = torch.empty((1, n, n), **f64)
    V = torch.empty((1, n, n), **f64)
    w = torch.empty((1, n), **f64)
    st = torch.zeros(SCF_STATE_LEN, **f64)
    st[ST_ENN], st[ST_RATE], st[ST_ETOL], st[ST_PTOL], st[ST_DIIS_TH] = E_NN, rate, e_tol, p_tol, diis_threshold
    st[ST_PREV_DE] = st[ST_PREV_DP] = float("inf")
    H, X, S = H.contiguous(), X.contiguous(), S.contiguous()

    def body():
        s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        plan.jk_into(P, Pa, None, J, K, None, mode)
        # F = H + J - K, R = F Pa S - S Pa F (16 x 16-tile products)
        scf_gemm_(Pa, S, PS)
        scf_fock_residual_(H, J, K, PS, F, R)
        _lib.check(lib.dft_scf_diis_stage(n, history, F.data_ptr(), R.data_ptr(), histF.data_ptr(),
                                          histR.data_ptr(), G.data_ptr(), P.data_ptr(), H.data_ptr(),
                                          Pa.data_ptr(), st.data_ptr(), Fu.data_ptr(), s))
        # A = (X Q)^T Fu (X Q), symmetrised; V = Q
        scf_gemm_(X, Q, XQ)
        scf_gemm_(Fu, XQ, T)
        scf_orbital_matrix_(XQ, T, A[0], Q, V[0])
        syevj_(A, w, V, False, tol=eig_tol)
        # orbitals X V (and Q = V), aufbau density C_occ C_occ^T
        scf_gemm_(X, V[0], Cm, copy=(V[0], Q))
        scf_gemm_(Cm, Cm, Pauf, trans_b=True, k=N_occ)
        _lib.check(lib.dft_scf_damp_control(n, Pa.data_ptr(), Pauf.data_ptr(), P.data_ptr(), st.data_ptr(), s))

    # warm-up of the GEMM paths on other buffers, outside the capture
    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        for _ in range(2):
            tmp = X @ (Q @ S)
            _ = tmp.T @ tmp
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize(dev)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        body()
    energies, iter_times = [], []
    converged = False
    # replay k + 1 is queued before iteration k's state is read (pinned
    # double buffer), so the device does not wait for the host between
    # iterations; a replay after convergence changes neither the densities
    # nor the loop state (dft_scf_damp_control)
    pin = [torch.empty(st.shape, dtype=st.dtype).pin_memory() for _ in range(2)]
    ev = [torch.cuda.Event(), torch.cuda.Event()]
    graph.replay()
    pin[0].copy_(st, non_blocking=True)
    ev[0].record()
    k = 0
    t0 = time.perf_counter()
    while True:
        graph.replay()
        pin[(k + 1) % 2].copy_(st, non_blocking=True)
        ev[(k + 1) % 2].record()
        ev[k % 2].synchronize()
        v = pin[k % 2].numpy().copy()
        t1 = time.perf_counter()
        iter_times.append(t1 - t0)
        t0 = t1
        E = float(v[ST_ENEW])
        if not np.isfinite(E):
            torch.cuda.current_stream(dev).synchronize()
            raise FloatingPointError(f"non-finite energy at iteration {int(v[ST_COUNT])}")
        energies.append(E)
        if v[ST_CONV] != 0:
            converged = True
            break
        if int(v[ST_COUNT]) > max_count:
            break
        k += 1
    torch.cuda.current_stream(dev).synchronize()
    if not converged:
        # the iteration queued after the last one read ran in full: report it
        w = pin[(k + 1) % 2].numpy().copy()
        if int(w[ST_COUNT]) > int(v[ST_COUNT]):
            v = w
            energies.append(float(v[ST_ENEW]))
            converged = bool(v[ST_CONV] != 0)
    count = int(v[ST_COUNT])
    # exit rebuild from the last density (shelled.py:1261-1285)
    fock = _Fock(plan, mode, True)
    Jf, Kf, _ = fock(P, Pa, Pa)
    Ff = H + Jf - Kf
    E_fin = float(0.5 * torch.sum(P * H + 2 * Pa * Ff)) + E_NN
    ea, Cp = eigh(X.T @ Ff @ X, Q)
    Cf = X @ Cp
    return SCFResult(E_fin, converged, count, Pa.clone(), Pa.clone(), Ff, Ff, Cf, Cf, ea, ea, fock.times,
                     iter_times, energies)


# ---------------------------------------------------------------------------
# Superposition of atomic densities (SAD) as the SCF starting density

_ATOM_DENSITY: dict = {}


def _atom_shell_key(z, shells) -> tuple:
    return (int(z),) + tuple((int(s["angular_momentum"]), np.asarray(s["exponents"], dtype=np.float64).tobytes(),
                              np.asarray(s["coefficients"], dtype=np.float64).tobytes()) for s in shells)


def atomic_density(z, shells, max_iter=200, tol=1e-10) -> np.ndarray:
    """Total density of the neutral atom of charge ``z`` in the basis
    ``shells`` (the atom's shell dicts, any centre): restricted Hartree-Fock
    with spherically averaged fractional occupations (two electrons per
    orbital in energy order, the last set of degenerate orbitals sharing the
    remainder equally), damped. The integrals (S, T, V and the ERI tensor)
    come from the library's kernels; the atomic matrices are a few functions
    wide, so the loop itself runs in NumPy. Cached per element and basis."""
    from .basis import build_shells_from_entries
    key = _atom_shell_key(z, shells)
    hit = _ATOM_DENSITY.get(key)
    if hit is not None:
        return hit
    entries = [([int(s["angular_momentum"])], np.asarray(s["exponents"]), np.asarray(s["coefficients"]))
               for s in shells]
    ash = build_shells_from_entries([entries], np.zeros((1, 3)))
    plan = Plan(ash, [z], np.zeros((1, 3)), schwarz_tau=0.0, fp32_tau=0.0)
    S, T, V = (M.cpu().numpy() for M in plan.one_electron())
    eri = plan.eri_tensor()
    n = plan.nbf
    H = T + V
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    nel = float(z)
    P = np.zeros((n, n))
    E_old = None
    for it in range(max_iter):
        F = H + np.einsum("abcd,cd->ab", eri, P) - 0.5 * np.einsum("acbd,cd->ab", eri, P)
        E = 0.5 * float(np.sum(P * (H + F)))
        eps, Cp = np.linalg.eigh(X.T @ F @ X)
        Cm = X @ Cp
        occ = np.zeros(n)
        left, i = nel, 0
        while left > 1e-12 and i < n:
            j = i
            while j + 1 < n and abs(eps[j + 1] - eps[i]) <= 1e-6 * max(1.0, abs(eps[i])):
                j += 1
            g = j - i + 1  # degenerate set i..j
            put = min(left, 2.0 * g)
            occ[i:j + 1] = put / g
            left -= put
            i = j + 1
        Pn = (Cm * occ) @ Cm.T
        P = Pn if it == 0 else 0.5 * (P + Pn)
        if E_old is not None and abs(E - E_old) < tol and np.abs(Pn - P).max() < 1e-8:
            break
        E_old = E
    _ATOM_DENSITY[key] = P
    return P


def sad_guess(shells, Z, N_a, N_b):
    """Alpha and beta starting densities (NumPy) from the superposition of
    the neutral atoms' densities (``atomic_density``), scaled to N_a and N_b
    electrons. The atoms' functions are contiguous in the shell order."""
    n = int(shells[-1]["ao_stop"])
    P = np.zeros((n, n))
    by_atom: dict = {}
    for s in shells:
        by_atom.setdefault(int(s["atom_idx"]), []).append(s)
    nel = 0.0
    for a, sh in by_atom.items():
        lo, hi = int(sh[0]["ao_start"]), int(sh[-1]["ao_stop"])
        if sum(int(s["n_ao"]) for s in sh) != hi - lo:
            raise ValueError(f"the functions of atom {a} are not contiguous in the shell order")
        P[lo:hi, lo:hi] = atomic_density(Z[a], sh)
        nel += float(Z[a])
    if nel <= 0:
        raise ValueError("no electrons in the superposition of atomic densities")
    return P * (N_a / nel), P * (N_b / nel)
