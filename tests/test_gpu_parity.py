"""GPU parity: the HIP path against the golden fixtures and the oracle.

Bars: ERIs within 1e-10 absolute, total energy within 1e-8 Eh,
J/K as the dense contraction of the oracle ERI tensor.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _plan(Z, xyz, tau=0.0, uhf=False, fp32_tau=0.0):
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    shells = build_shells(load_basis(), Z, xyz)
    return shells, Plan(shells, Z, xyz, schwarz_tau=tau, fp32_tau=fp32_tau, uhf=uhf)


def test_boys_fp64_vs_mpmath(torch_cuda):
    from dft_amd.engine import boys
    g = golden("boys_mpmath.npz")
    T, F = g["T"], g["F"]
    for mmax in (0, 1, 4, 8, 12, 16):
        got = boys(T, mmax).cpu().numpy()
        ref = F[:, :mmax + 1]
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        assert err.max() < 5e-14, (mmax, err.max(), T[np.unravel_index(err.argmax(), err.shape)[0]])


def test_boys_fp32_vs_mpmath(torch_cuda):
    from dft_amd.engine import boys
    g = golden("boys_mpmath.npz")
    T, F = g["T"], g["F"]
    sel = T < 1e3
    got = boys(T[sel], 8, fp32=True).cpu().numpy()
    ref = F[sel, :9]
    err = np.abs(got - ref) / np.abs(ref)
    assert err.max() < 2e-6, err.max()


@pytest.mark.parametrize("name", ["water_sto3g.npz", "dimer_sto3g.npz"])
def test_eri_tensor_vs_golden(torch_cuda, name):
    g = golden(name)
    _, plan = _plan(list(g["Z"]), g["xyz"])
    T = plan.eri_tensor()
    err = np.abs(T - g["ERI"]).max()
    assert err < 1e-10, err


@pytest.mark.parametrize("name", ["water_sto3g.npz", "dimer_sto3g.npz"])
def test_one_electron_vs_golden(torch_cuda, name):
    g = golden(name)
    _, plan = _plan(list(g["Z"]), g["xyz"])
    S, T, V = (x.cpu().numpy() for x in plan.one_electron())
    for got, ref in ((S, g["S"]), (T, g["T"]), (V, g["V"])):
        assert np.abs(got - ref).max() < 1e-10


def _rand_density(n, rng):
    A = rng.standard_normal((n, n))
    return (A + A.T) / 2


@pytest.mark.parametrize("name", ["water_sto3g.npz", "dimer_sto3g.npz"])
@pytest.mark.parametrize("uhf", [False, True])
def test_jk_vs_dense(torch_cuda, name, uhf):
    import oracle.md_oracle as O
    g = golden(name)
    ERI = g["ERI"]
    n = ERI.shape[0]
    rng = np.random.default_rng(11)
    Pa, Pb = _rand_density(n, rng), _rand_density(n, rng)
    P = Pa + Pb
    _, plan = _plan(list(g["Z"]), g["xyz"], uhf=uhf)
    J, Ka, Kb = plan.jk(P, Pa, Pb if uhf else None)
    Jr, Kar, Kbr = O.jk_dense(ERI, P, Pa, Pb)
    assert np.abs(J.cpu().numpy() - Jr).max() < 1e-10
    assert np.abs(Ka.cpu().numpy() - Kar).max() < 1e-10
    if uhf:
        assert np.abs(Kb.cpu().numpy() - Kbr).max() < 1e-10


def test_jk_deterministic(torch_cuda):
    from dft_amd.geometry import water_cluster
    Z, xyz = water_cluster()
    Z, xyz = Z[:18], xyz[:18]
    shells, plan = _plan(Z, xyz, tau=1e-10)
    n = plan.nbf
    rng = np.random.default_rng(3)
    P = _rand_density(n, rng)
    J1, K1, _ = plan.jk(P, P / 2)
    J2, K2, _ = plan.jk(P, P / 2)
    assert torch_cuda.equal(J1, J2) and torch_cuda.equal(K1, K2)


def test_scf_energy_water(torch_cuda):
    from dft_amd.scf import scf_plain
    g = golden("water_sto3g.npz")
    Z = list(g["Z"])
    _, plan = _plan(Z, g["xyz"], tau=1e-12)
    res = scf_plain(plan, Z, g["xyz"], 5, 5)
    assert res.converged
    assert abs(res.energy - float(g["E"])) < 1e-8, res.energy - float(g["E"])


def test_scf_energy_dimer(torch_cuda):
    from dft_amd.scf import scf_plain
    g = golden("dimer_sto3g.npz")
    Z = list(g["Z"])
    _, plan = _plan(Z, g["xyz"], tau=1e-12)
    res = scf_plain(plan, Z, g["xyz"], 10, 10)
    assert res.converged
    assert abs(res.energy - float(g["E"])) < 1e-8, res.energy - float(g["E"])


def test_size_consistency_far_apart_waters(torch_cuda):
    """E(A + B) = E(A) + E(B) for two waters 1000 bohr apart: their
    interaction (dipole-dipole, < 1e-9 Eh there) is below the tolerance, so
    the intermolecular nuclear repulsion, nuclear attraction and Coulomb
    terms (each ~0.1 Eh) must cancel to 1e-8."""
    from dft_amd.geometry import water
    from dft_amd.scf import scf_plain
    Z, xyz = water()
    _, p1 = _plan(Z, xyz, tau=1e-12)
    r1 = scf_plain(p1, Z, xyz, 5, 5)
    Z2 = list(Z) + list(Z)
    xyz2 = np.vstack([xyz, xyz + np.array([1000.0, 0.0, 0.0])])
    _, p2 = _plan(Z2, xyz2, tau=1e-12)
    r2 = scf_plain(p2, Z2, xyz2, 10, 10)
    assert r1.converged and r2.converged
    assert abs(r2.energy - 2 * r1.energy) < 1e-8, r2.energy - 2 * r1.energy


def test_schwarz_bound_holds(torch_cuda):
    g = golden("dimer_sto3g.npz")
    ERI = g["ERI"]
    shells, plan = _plan(list(g["Z"]), g["xyz"])
    Q = plan.schwarz()
    n = len(shells)
    for a in range(n):
        for b in range(n):
            sa = slice(shells[a]["ao_start"], shells[a]["ao_stop"])
            sb = slice(shells[b]["ao_start"], shells[b]["ao_stop"])
            diag = np.abs(np.einsum("ijij->ij", ERI[sa, sb, sa, sb])).max()
            assert abs(Q[a, b] - np.sqrt(diag)) < 1e-12


def _methane_plan(uhf=False, tau=0.0, fp32_tau=0.0):
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import methane, synthetic_spd_basis
    Z, xyz = methane()
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    return shells, Plan(shells, Z, xyz, schwarz_tau=tau, fp32_tau=fp32_tau, uhf=uhf)


def test_eri_tensor_spd_vs_golden(torch_cuda):
    g = golden("methane_spd.npz")
    _, plan = _methane_plan()
    T = plan.eri_tensor()
    ref = g["ERI"]
    err = np.abs(T - ref) / (1e-10 + 1e-11 * np.abs(ref))
    assert err.max() < 1.0, (np.abs(T - ref).max(), np.abs(ref).max())


def test_one_electron_spd_vs_golden(torch_cuda):
    g = golden("methane_spd.npz")
    _, plan = _methane_plan()
    S, T, V = (x.cpu().numpy() for x in plan.one_electron())
    for got, ref in ((S, g["S"]), (T, g["T"]), (V, g["V"])):
        assert np.abs(got - ref).max() < 1e-10 + 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("uhf", [False, True])
def test_jk_spd_vs_dense(torch_cuda, uhf):
    import oracle.md_oracle as O
    g = golden("methane_spd.npz")
    ERI = g["ERI"]
    n = ERI.shape[0]
    rng = np.random.default_rng(13)
    Pa, Pb = _rand_density(n, rng) * 0.1, _rand_density(n, rng) * 0.1
    _, plan = _methane_plan(uhf=uhf)
    J, Ka, Kb = plan.jk(Pa + Pb, Pa, Pb if uhf else None)
    Jr, Kar, Kbr = O.jk_dense(ERI, Pa + Pb, Pa, Pb)
    scale = 1e-10 * max(1.0, np.abs(Jr).max())
    assert np.abs(J.cpu().numpy() - Jr).max() < scale
    assert np.abs(Ka.cpu().numpy() - Kar).max() < scale
    if uhf:
        assert np.abs(Kb.cpu().numpy() - Kbr).max() < scale


def test_mixed_two_fp32_levels_water_cluster(torch_cuda):
    """(H2O)48 STO-3G is large enough for the second FP32 record level (over
    2^20 quartets of a class below 1e-3 of the FP32 threshold, with a 1000x
    looser primitive-pair cut). Mixed J/K stay within 1e-7 of the largest
    FP64 element (the FP64 path itself is checked against the oracle in the
    tests above)."""
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import water_cluster
    Z, xyz = water_cluster((4, 4, 3), 3.1, seed=3)
    shells = build_shells(load_basis(), Z, xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    st = plan.stats()
    assert st["n_quartets_fp32_level2"] > 0
    rng = np.random.default_rng(5)
    Pa = _rand_density(plan.nbf, rng) * 0.05
    J64, K64, _ = plan.jk(2 * Pa, Pa, mode="fp64")
    J32, K32, _ = plan.jk(2 * Pa, Pa, mode="mixed")
    dJ = (J64 - J32).abs().max().item() / J64.abs().max().item()
    dK = (K64 - K32).abs().max().item() / K64.abs().max().item()
    print(f"level-2 quartets {st['n_quartets_fp32_level2']}, rel dJ {dJ:.2e}, rel dK {dK:.2e}")
    assert dJ < 1e-7 and dK < 1e-7


def test_mixed_precision_spd_error_small(torch_cuda):
    g = golden("methane_spd.npz")
    n = g["ERI"].shape[0]
    rng = np.random.default_rng(17)
    Pa = _rand_density(n, rng) * 0.05
    _, plan = _methane_plan(tau=1e-12, fp32_tau=1e-4)
    st = plan.stats()
    assert st["n_quartets_fp32"] > 0
    J64, K64, _ = plan.jk(2 * Pa, Pa, mode="fp64")
    J32, K32, _ = plan.jk(2 * Pa, Pa, mode="mixed")
    # FP32 only on quartets with Q_ab Q_cd < 1e-4: error far below the largest element
    assert (J64 - J32).abs().max().item() < 1e-6 * max(1.0, J64.abs().max().item())
    assert (K64 - K32).abs().max().item() < 1e-6 * max(1.0, K64.abs().max().item())


# ----------------------------------------------------------- edge cases

def _dense_check(shells, plan, rng, uhf=False, tol=1e-10):
    import oracle.md_oracle as O
    ERI = O.eri_tensor(shells)
    n = ERI.shape[0]
    T = np.asarray(plan.eri_tensor())
    assert np.abs(T - ERI).max() < tol
    Pa, Pb = _rand_density(n, rng), _rand_density(n, rng)
    P = Pa + Pb
    J, Ka, Kb = plan.jk(P, Pa, Pb if uhf else None)
    Jr, Kar, Kbr = O.jk_dense(ERI, P, Pa, Pb)
    assert np.abs(J.cpu().numpy() - Jr).max() < tol
    assert np.abs(Ka.cpu().numpy() - Kar).max() < tol
    if uhf:
        assert np.abs(Kb.cpu().numpy() - Kbr).max() < tol


def test_single_shell_hydrogen_atom(torch_cuda):
    """One s shell, one basis function, one quartet (both precisions)."""
    shells, plan = _plan([1], np.zeros((1, 3)), uhf=True)
    _dense_check(shells, plan, np.random.default_rng(1), uhf=True)
    st = plan.stats()
    assert st["n_quartets"] == 1


def test_far_apart_atoms_have_empty_pairs(torch_cuda):
    """H at 60 bohr from water: every primitive pair of a cross shell pair
    is below the exp(-mu R^2) cut, so those pairs have no primitive records;
    without Schwarz screening their quartets still run (and give 0)."""
    from dft_amd.geometry import water
    Z, xyz = water()
    Z = list(Z) + [1]
    xyz = np.vstack([xyz, [[60.0, 0.0, 0.0]]])
    shells, plan = _plan(Z, xyz, tau=0.0)
    _dense_check(shells, plan, np.random.default_rng(2))
    shells, plan = _plan(Z, xyz, tau=1e-10, fp32_tau=1e-4)
    for mode in ("fp64", "mixed"):
        J, K, _ = plan.jk(np.eye(plan.nbf) * 0.1, np.eye(plan.nbf) * 0.05, mode=mode)
        assert bool(torch_cuda.isfinite(J).all()) and bool(torch_cuda.isfinite(K).all())


def test_everything_screened_gives_zero_jk(torch_cuda):
    """A Schwarz threshold above every bound leaves no quartet: J = K = 0."""
    from dft_amd.geometry import water
    Z, xyz = water()
    _, plan = _plan(Z, xyz, tau=1e6, fp32_tau=1e7)
    assert plan.stats()["n_quartets"] == 0
    n = plan.nbf
    for mode in ("fp64", "mixed"):
        J, K, _ = plan.jk(np.ones((n, n)), np.ones((n, n)) * 0.5, mode=mode)
        assert float(J.abs().max()) == 0.0 and float(K.abs().max()) == 0.0


def test_scf_checkpoint_resume_continues_exactly(torch_cuda, tmp_path):
    """scf_diis writes the reference's scf_checkpoint.npz layout every 10th
    iteration; resuming from it reproduces the uninterrupted run."""
    from dft_amd.geometry import water_cluster
    from dft_amd.scf import CHECKPOINT_KEYS, scf_diis
    Z, xyz = water_cluster((2, 2, 1), 3.1, seed=5)
    _, plan = _plan(Z, xyz, tau=1e-12)
    nocc = sum(Z) // 2
    ck = str(tmp_path / "scf_checkpoint.npz")
    full = scf_diis(plan, Z, xyz, nocc, nocc, checkpoint=ck)
    assert full.converged and full.iterations >= 10
    with np.load(ck) as f:
        assert set(CHECKPOINT_KEYS) <= set(f.files)
        assert f["DIIS_P_a_history"].shape == (10, plan.nbf, plan.nbf)
        saved = int(f["count"])
    res = scf_diis(plan, Z, xyz, nocc, nocc, resume=ck)
    assert res.converged
    assert res.iterations == full.iterations and saved <= full.iterations
    assert abs(res.energy - full.energy) < 1e-10
