"""The reference's self-check invariants (main.py:967-1026) on the HIP path,
failure handling of the SCF loop, and property tests over random small
molecules (HIP ERIs against the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _dimer():
    from conftest import golden
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    g = golden("dimer_sto3g.npz")
    Z, xyz = list(g["Z"]), np.asarray(g["xyz"])
    shells = build_shells(load_basis(), Z, xyz)
    return shells, Z, xyz, Plan(shells, Z, xyz, schwarz_tau=0.0)


def test_reference_self_check_invariants(torch_cuda):
    """main.py:967-1026 on a converged water-dimer SCF of the HIP path."""
    torch = torch_cuda
    from dft_amd.grid import box_grid
    from dft_amd.scf import _setup, nuclear_repulsion, scf_diis
    shells, Z, xyz, plan = _dimer()
    S, T, V = (m.cpu().numpy() for m in plan.one_electron())
    H = T + V
    for M in (S, T, V, H):
        assert np.abs(M - M.T).max() < 1e-12                          # :971-976
    assert np.abs(np.diag(S) - 1.0).max() < 1e-3                        # :970 (contracted, not renormalised)
    assert np.linalg.eigvalsh(T).min() > 0 and np.linalg.eigvalsh(V).max() < 0   # :973, :975
    assert nuclear_repulsion(Z, xyz) > 0                                # :977
    E = np.asarray(plan.eri_tensor())
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):             # :978-980
        assert np.abs(E - E.transpose(perm)).max() < 1e-12
    n = E.shape[0]
    diag = E.reshape(n * n, n * n).diagonal().reshape(n, n)             # (ij|ij)
    assert (np.einsum("iiii->i", E) >= 0).all()                          # :986
    assert (E.reshape(n * n, n * n) ** 2 <= np.outer(diag, diag) * (1 + 1e-12) + 1e-24).all()  # :987
    res = scf_diis(plan, Z, xyz, 10, 10)
    assert res.converged
    X = _setup(plan, Z, xyz)[2].cpu().numpy()                           # X = U s^-1/2 U^T
    assert np.abs(X.T @ S @ X - np.eye(n)).max() < 1e-10                # :989
    C = res.C_a.cpu().numpy()
    assert np.abs(C.T @ S @ C - np.eye(n)).max() < 1e-10                # :992
    F = res.F_a.cpu().numpy()
    eps = res.eps_a.cpu().numpy()
    assert np.abs(F @ C - S @ C * eps[None, :]).max() < 1e-9            # :994
    P = C[:, :10] @ C[:, :10].T
    assert abs(np.trace(P @ S) - 10) < 1e-10                            # :1000-1002
    assert np.abs(P - P.T).max() < 1e-12 and np.abs(P @ S @ P - P).max() < 1e-10   # :1003-1006
    assert np.linalg.matrix_rank(P, tol=1e-8) == 10                     # :1007-1008
    assert np.abs(F @ P @ S - S @ P @ F).max() < 1e-5                   # :1020-1021 (converged to dP 1e-6)
    g, shape = box_grid(xyz.min(0) - 6.0, xyz.max(0) + 6.0, 0.1)
    dv = float(np.prod((xyz.max(0) - xyz.min(0) + 12.0) / (np.asarray(shape) - 1)))
    Pt = torch.as_tensor(2 * P, device=plan.device)
    rho = torch.cat([((Pt @ a) * a).sum(0) for a in (plan.ao_values(g[i:i + (1 << 20)])
                                                       for i in range(0, len(g), 1 << 20))])
    assert float(rho.min()) >= -1e-12                                   # :1024-1025
    assert abs(float(rho.sum()) * dv - 20.0) < 0.2                      # :1022-1023 (rtol 1e-2)


def test_nan_density_stops_the_scf_cleanly(torch_cuda):
    """A non-finite density (e.g. a corrupted resume) raises at the first
    Fock build instead of iterating on NaN."""
    from dft_amd.scf import scf_diis
    shells, Z, xyz, plan = _dimer()
    P = np.full((plan.nbf, plan.nbf), np.nan)
    with pytest.raises(FloatingPointError):
        scf_diis(plan, Z, xyz, 10, 10, guess=(P, P))


def _random_molecule(seed):
    """2-3 centres within 3 bohr, 1-2 shells each with l <= 2, 1-3
    primitives with exponents log-uniform in [0.1, 50], random coefficients."""
    from dft_amd.basis import build_shells_from_entries
    rng = np.random.default_rng(seed)
    nat = int(rng.integers(2, 4))
    xyz = rng.uniform(-1.5, 1.5, size=(nat, 3))
    entries = []
    for _ in range(nat):
        ent = []
        for _ in range(int(rng.integers(1, 3))):
            l = int(rng.integers(0, 3))
            npr = int(rng.integers(1, 4))
            ex = np.exp(rng.uniform(np.log(0.1), np.log(50.0), size=npr))
            co = rng.uniform(0.2, 1.0, size=npr)
            ent.append([[l], ex.tolist(), [co.tolist()]])
        entries.append(ent)
    return build_shells_from_entries(entries, xyz), [1] * nat, xyz


def test_random_molecules_eri_and_jk_match_oracle(torch_cuda):
    """Property test over 12 random s/p/d molecules: HIP ERI tensor and J/K
    against the oracle (1e-10 relative to the largest integral)."""
    import oracle.md_oracle as O
    from dft_amd.engine import Plan
    for seed in range(12):
        shells, Z, xyz = _random_molecule(100 + seed)
        plan = Plan(shells, Z, xyz, schwarz_tau=0.0)
        R = O.eri_tensor(shells)
        E = np.asarray(plan.eri_tensor())
        scale = max(1.0, np.abs(R).max())
        assert np.abs(E - R).max() < 1e-10 * scale, seed
        rng = np.random.default_rng(seed)
        n = R.shape[0]
        A = rng.standard_normal((n, n))
        Pa = (A + A.T) / 4
        J, K, _ = plan.jk(2 * Pa, Pa)
        Jr, Kr, _ = O.jk_dense(R, 2 * Pa, Pa, Pa)
        assert np.abs(J.cpu().numpy() - Jr).max() < 1e-10 * scale * max(1.0, np.abs(Pa).sum()), seed
        assert np.abs(K.cpu().numpy() - Kr).max() < 1e-10 * scale * max(1.0, np.abs(Pa).sum()), seed
