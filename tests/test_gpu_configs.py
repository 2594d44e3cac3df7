"""Parity at the size of the benchmark workloads.

Each test runs one of the five benchmark configurations at full size on the
HIP path and checks it against the oracle (the C restatement of the
reference's per-primitive-Cartesian McMurchie-Davidson path, or the NumPy
restatement) or against size-independent properties:

* (H2O)16 / STO-3G, Schwarz 1e-10: direct J/K against the C oracle's
  fill_matrices over the identical screened quartet list; mixed precision
  against the same FP64 oracle J/K (per Fock matrix and per energy).
* C12H26 with the synthetic s/p/d basis: 1000 sampled contracted quartets
  per class combination against the C oracle (1e-10 absolute); FP64 vs mixed
  SCF energies and Fock matrices; a C2H6 fragment with the same basis rule
  through the FP64 J/K against the C oracle.
* 64 (H2O)_n clusters: the batched J/K bitwise equal to 64 single-molecule
  builds, the three largest molecules against the C oracle, mixed vs FP64
  batched SCF energies of every molecule within 1 uEh.
* C8H18 + noise / STO-3G: J/K against the C oracle, analytic gradient
  against central finite differences of the SCF energy on 12 coordinates.
* A molecule with a general contraction (two coefficient rows on one shell)
  through the direct J/K against the NumPy oracle.

Bars: ERIs 1e-10 absolute; J/K 1e-9 absolute (sums over up to ~1e4 kets of
integrals that each carry the 1e-14 primitive-quartet cut); energies 1e-8 Eh
(FP64) and 1e-6 Eh (mixed); gradients 1e-6 Eh/bohr.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

JK_TOL = 1e-9


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _screened(Q, tau):
    """Unique (a>=b, c>=d, ab>=cd) segment quartets with Q_ab Q_cd >= tau."""
    n = Q.shape[0]
    ii, jj = np.tril_indices(n)
    qp = Q[ii, jj]
    bra, ket = np.tril_indices(len(ii))
    keep = qp[bra] * qp[ket] >= tau
    bra, ket = bra[keep], ket[keep]
    return np.stack([ii[bra], jj[bra], ii[ket], jj[ket]], axis=1).astype(np.int32)


def _random_density(n, seed, scale=0.2):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    return (A + A.T) / 2


def _oracle_jk(shells, plan, tau, P, Pa):
    import oracle.ref as ref
    seg = ref.Segments(shells)
    q = _screened(plan.schwarz(), tau)
    J, K, _ = ref.build_jk(seg, q, P, Pa, None)
    return J, K, len(q)


def _check_jk(torch, plan, shells, tau, P, Pa, label):
    Pd = torch.as_tensor(P, device="cuda")
    Pad = torch.as_tensor(Pa, device="cuda")
    J, K, _ = plan.jk(Pd, Pad)
    Jr, Kr, nq = _oracle_jk(shells, plan, tau, P, Pa)
    assert nq == plan.stats()["n_quartets"], (label, nq, plan.stats()["n_quartets"])
    dj = float(np.abs(J.cpu().numpy() - Jr).max())
    dk = float(np.abs(K.cpu().numpy() - Kr).max())
    print(f"{label}: {nq} quartets, max|dJ| {dj:.2e}, max|dK| {dk:.2e}", flush=True)
    assert dj < JK_TOL and dk < JK_TOL, (label, dj, dk)
    return Jr, Kr


# --------------------------------------------------------- (H2O)16, config 1
def test_water16_jk_full_list_vs_c_oracle(torch_cuda):
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import water_cluster
    from dft_amd.scf import scf_plain
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    shells = build_shells(load_basis(), Z, xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    # a realistic density: two Roothaan iterations from the core guess
    r = scf_plain(plan, Z, xyz, 80, 80, max_count=1, e_tol=-1, p_tol=-1)
    Pa = r.P_a.cpu().numpy()
    P = 2 * Pa
    Jr, Kr = _check_jk(torch_cuda, plan, shells, 1e-10, P, Pa, "(H2O)16 fp64")
    # mixed precision against the FP64 oracle, per Fock matrix and per energy
    # functional E[P] = sum P (H + (J - K)/2)-part: 1/2 sum (P J - Pa K - Pb K)
    Pd = torch_cuda.as_tensor(P, device="cuda")
    Pad = torch_cuda.as_tensor(Pa, device="cuda")
    J32, K32, _ = plan.jk(Pd, Pad, mode="mixed")
    dF = float(np.abs((J32 - K32).cpu().numpy() - (Jr - Kr)).max())
    E2 = lambda J, K: 0.5 * float(np.sum(P * J) - 2 * np.sum(Pa * K))  # noqa: E731
    dE = abs(E2(J32.cpu().numpy(), K32.cpu().numpy()) - E2(Jr, Kr))
    print(f"(H2O)16 mixed vs FP64 oracle: max|dF| {dF:.2e} Eh, |dE| {dE:.2e} Eh", flush=True)
    assert dF < 1e-6 and dE < 1e-6


# ------------------------------------------------- C12H26 s/p/d, config 2
@pytest.fixture(scope="module")
def alkane12(torch_cuda):
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import alkane_bohr, synthetic_spd_basis
    Z, xyz = alkane_bohr(12)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    return Z, xyz, shells, plan


COMBOS = [(pb, pk) for pb in range(6) for pk in range(pb + 1)]
PC_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


@pytest.mark.parametrize("combo", COMBOS, ids=[f"{PC_L[b]}{PC_L[k]}" for b, k in COMBOS])
def test_alkane12_sampled_quartets_vs_c_oracle(alkane12, combo):
    """1000 random screened quartets of one class combination (all of them
    if fewer), as contracted blocks in the requested orientation."""
    import oracle.ref as ref
    Z, xyz, shells, plan = alkane12
    seg = ref.Segments(shells)
    l = seg.l
    Q = plan.schwarz()
    q = _screened(Q, 1e-10)
    lb_, lk_ = PC_L[combo[0]], PC_L[combo[1]]

    def cls(a, b):
        return np.where(l[a] >= l[b], l[a] * (l[a] + 1) // 2 + l[b], l[b] * (l[b] + 1) // 2 + l[a])
    pb, pk = cls(q[:, 0], q[:, 1]), cls(q[:, 2], q[:, 3])
    sel = ((pb == combo[0]) & (pk == combo[1])) | ((pb == combo[1]) & (pk == combo[0]))
    cand = q[sel]
    if len(cand) == 0:
        pytest.skip(f"no screened quartet of class {lb_}{lk_} in this molecule")
    rng = np.random.default_rng(combo[0] * 7 + combo[1])
    pick = cand[rng.choice(len(cand), size=min(1000, len(cand)), replace=False)]
    got = plan.eri_segment_blocks(pick)
    want, _ = ref.eri_blocks(seg, pick)
    err = float(np.abs(got - want).max())
    print(f"class {lb_}{lk_}: {len(pick)} of {len(cand)} quartets, max abs err {err:.2e}", flush=True)
    assert err < 1e-10


def test_alkane12_mixed_vs_fp64_scf(alkane12, torch_cuda):
    """|dE| of the converged SCF energies and max|dF| at the converged FP64
    density, mixed (FP32 quartets below Q_ab Q_cd = 1e-4) against FP64."""
    from dft_amd.scf import electron_counts, scf_diis
    from dft_amd.shelled import spherical_matrix
    torch = torch_cuda
    Z, xyz, shells, plan = alkane12
    A = torch.as_tensor(spherical_matrix(shells, shells[-1]["sphere_ao_stop"]), device="cuda")
    N_a, N_b = electron_counts(Z)
    r64 = scf_diis(plan, Z, xyz, N_a, N_b, mode="fp64", A=A, max_count=200)
    r32 = scf_diis(plan, Z, xyz, N_a, N_b, mode="mixed", A=A, max_count=200)
    assert r64.converged and r32.converged
    Pc = (A.T @ r64.P_a @ A).contiguous()
    J64, K64, _ = plan.jk(2 * Pc, Pc, mode="fp64")
    J32, K32, _ = plan.jk(2 * Pc, Pc, mode="mixed")
    dF = float((A @ ((J64 - K64) - (J32 - K32)) @ A.T).abs().max())
    dE = abs(r64.energy - r32.energy)
    print(f"C12H26 spd: E64 {r64.energy:.10f} |dE| {dE:.2e} Eh, max|dF| {dF:.2e} Eh", flush=True)
    assert dE < 1e-6 and dF < 1e-6


def test_ethane_spd_fp64_jk_vs_c_oracle(torch_cuda):
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import alkane_bohr, synthetic_spd_basis
    Z, xyz = alkane_bohr(2)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=0.0)
    Pa = _random_density(plan.nbf, 11, 0.05)
    _check_jk(torch_cuda, plan, shells, 1e-10, 2 * Pa, Pa, "C2H6 spd fp64")


# ------------------------------------------------ 64 clusters, config 3
@pytest.fixture(scope="module")
def batch64(torch_cuda):
    from dft_amd.batch import water_batch
    from dft_amd.engine import BatchPlan
    mols = water_batch(64, 2)
    return mols, BatchPlan(mols, schwarz_tau=1e-10, fp32_tau=1e-4, split=False)


def test_batch64_bitwise_equal_to_single_plans(batch64, torch_cuda):
    from dft_amd.engine import Plan
    torch = torch_cuda
    mols, bp = batch64
    dens = [_random_density(sh[-1]["ao_stop"], 100 + m, 0.1) for m, (sh, _, _) in enumerate(mols)]
    P = bp.pack([2 * d for d in dens])
    Pa = bp.pack(dens)
    J, K, _ = bp.jk(P, Pa)
    Js, Ks = bp.unpack(J), bp.unpack(K)
    for m, (sh, Z, xyz) in enumerate(mols):
        pl = Plan(sh, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4, split=False)
        j, k, _ = pl.jk(torch.as_tensor(2 * dens[m], device="cuda"), torch.as_tensor(dens[m], device="cuda"))
        assert torch.equal(j, Js[m]) and torch.equal(k, Ks[m]), m
        del pl


def test_batch64_largest_molecules_vs_c_oracle(batch64, torch_cuda):
    from dft_amd.engine import Plan
    mols, _ = batch64
    order = np.argsort([-sh[-1]["ao_stop"] for sh, _, _ in mols], kind="stable")[:3]
    for m in order:
        sh, Z, xyz = mols[m]
        plan = Plan(sh, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4, split=False)
        Pa = _random_density(plan.nbf, 200 + int(m), 0.1)
        _check_jk(torch_cuda, plan, sh, 1e-10, 2 * Pa, Pa, f"batch molecule {m} ({plan.nbf} bf)")


def test_batch64_mixed_vs_fp64_scf(batch64, torch_cuda):
    from dft_amd.batch import scf_batch
    mols, bp = batch64
    nocc = [int(sum(Z)) // 2 for _, Z, _ in mols]
    r64 = scf_batch(bp, nocc, mode="fp64", max_iter=60)
    r32 = scf_batch(bp, nocc, mode="mixed", max_iter=60)
    assert r64.converged.all() and r32.converged.all()
    dE = float(np.abs(r64.energies - r32.energies).max())
    print(f"batch of 64: {r64.iterations}/{r32.iterations} iterations, max|dE| mixed {dE:.2e} Eh", flush=True)
    assert dE < 1e-6


# --------------------------------------------------- C8H18 gradient, config 4
def test_octane_jk_and_gradient_vs_fd(torch_cuda):
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import alkane_bohr
    from dft_amd.grad import energy_and_gradient
    Z, xyz = alkane_bohr(8, noise_sigma_ang=0.05, seed=3)
    basis = load_basis()
    shells = build_shells(basis, Z, xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-12, fp32_tau=0.0)
    Pa = _random_density(plan.nbf, 7, 0.1)
    _check_jk(torch_cuda, plan, shells, 1e-12, 2 * Pa, Pa, "C8H18 fp64")
    E, g, res = energy_and_gradient(Z, xyz, basis)
    rng = np.random.default_rng(0)
    atoms = rng.choice(len(Z), 12, replace=False)
    axes = rng.integers(0, 3, 12)
    h = 1e-4
    err = 0.0
    for a, x in zip(atoms, axes):
        p, m = xyz.copy(), xyz.copy()
        p[a, x] += h
        m[a, x] -= h
        guess = (res.P_a, res.P_b)
        fd = (energy_and_gradient(Z, p, basis, guess=guess)[0] - energy_and_gradient(Z, m, basis, guess=guess)[0]) / (2 * h)
        err = max(err, abs(fd - g[a, x]))
    print(f"C8H18 gradient: max |analytic - FD| over 12 coordinates {err:.2e} Eh/bohr", flush=True)
    assert err < 1e-6
    assert np.abs(g.sum(0)).max() < 1e-8  # translational invariance


# ------------------------------------------------ general contraction
def test_general_contraction_jk_vs_numpy_oracle(torch_cuda):
    """Water whose O 2s/2p and H 1s shells carry two coefficient rows each
    (shelled.py:57-65, contraction-major AO order)."""
    import oracle.md_oracle as O
    from dft_amd.basis import build_shells_from_entries, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import water
    Z, xyz = water()
    b = load_basis()
    rng = np.random.default_rng(9)
    per_atom = []
    for z in Z:
        ents = []
        for ls, ex, co in b[str(int(z))]:
            rows = [list(r) for r in co]
            if len(ls) == 1:
                rows = rows + [list(rng.uniform(0.1, 1.0, len(ex)))]
                ents.append([ls, ex, rows])
            else:  # SP entry: s and p rows, each given a second row
                for m, r in zip(ls, rows):
                    ents.append([[m], ex, [r, list(rng.uniform(0.1, 1.0, len(ex)))]])
        per_atom.append(ents)
    shells = build_shells_from_entries(per_atom, xyz)
    assert any(sh["n_contractions"] == 2 for sh in shells)
    plan = Plan(shells, Z, xyz, schwarz_tau=0.0, fp32_tau=0.0)
    ERI = O.eri_tensor(shells)
    Pa = _random_density(plan.nbf, 3, 0.3)
    J, K, _ = plan.jk(torch_cuda.as_tensor(2 * Pa, device="cuda"), torch_cuda.as_tensor(Pa, device="cuda"))
    Jr, Kr, _ = O.jk_dense(ERI, 2 * Pa, Pa, Pa)
    dj = float(np.abs(J.cpu().numpy() - Jr).max())
    dk = float(np.abs(K.cpu().numpy() - Kr).max())
    print(f"general contraction ({plan.nbf} bf): max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
    assert dj < 1e-10 and dk < 1e-10
    eri = np.concatenate([blk.ravel() for blk in plan.eri_blocks([(0, 0, 0, 0), (2, 1, 1, 0)])])
    s = [slice(shells[k]["ao_start"], shells[k]["ao_stop"]) for k in (0, 1, 2)]
    want = np.concatenate([ERI[s[0], s[0], s[0], s[0]].ravel(), ERI[s[2], s[1], s[1], s[0]].ravel()])
    assert np.abs(eri - want).max() < 1e-10
