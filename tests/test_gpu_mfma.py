"""The heavy d class combinations on the matrix cores (DFT_PLAN_MFMA: the
ket and bra McMurchie-Davidson contractions as FP64 MFMA products) against
the oracle and against the VALU column kernel."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _sym(n, seed, scale):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    return (A + A.T) / 2


@pytest.mark.parametrize("uhf", [False, True])
def test_mfma_jk_methane_spd_vs_dense_oracle(torch_cuda, uhf):
    import oracle.md_oracle as O
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import methane, synthetic_spd_basis
    torch = torch_cuda
    g = golden("methane_spd.npz")
    Z, xyz = methane()
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=0.0, fp32_tau=0.0, uhf=uhf, mfma=True)
    n = plan.nbf
    Pa, Pb = _sym(n, 1, 0.3), _sym(n, 2, 0.3)
    dev = lambda x: torch.as_tensor(x, device="cuda")  # noqa: E731
    J, Ka, Kb = plan.jk(dev(Pa + Pb), dev(Pa), dev(Pb) if uhf else None)
    Jr, Kar, Kbr = O.jk_dense(g["ERI"], Pa + Pb, Pa, Pb)
    err = [float(np.abs(J.cpu().numpy() - Jr).max()), float(np.abs(Ka.cpu().numpy() - Kar).max())]
    if uhf:
        err.append(float(np.abs(Kb.cpu().numpy() - Kbr).max()))
    print(f"methane s/p/d MFMA J/K vs dense oracle: {err}", flush=True)
    assert max(err) < 1e-10 * max(1.0, float(np.abs(Jr).max()))


def test_mfma_jk_alkane12_matches_column_kernel(torch_cuda):
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import alkane_bohr, synthetic_spd_basis
    from dft_amd.workloads import time_builds
    torch = torch_cuda
    Z, xyz = alkane_bohr(12)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    col = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    mf = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4, mfma=True)
    Pa = torch.as_tensor(_sym(col.nbf, 3, 0.05), device="cuda")
    P = 2 * Pa
    J0, K0, _ = col.jk(P, Pa)
    J1, K1, _ = mf.jk(P, Pa)
    dj = float((J0 - J1).abs().max())
    dk = float((K0 - K1).abs().max())
    J, K = torch.empty_like(P), torch.empty_like(P)
    t0 = time_builds(lambda: col.jk_into(P, Pa, None, J, K, None, "fp64"), 5, 2)
    t1 = time_builds(lambda: mf.jk_into(P, Pa, None, J, K, None, "fp64"), 5, 2)
    print(f"C12H26 s/p/d FP64 build: column kernel {t0:.2f} ms, MFMA {t1:.2f} ms; max|dJ| {dj:.2e} max|dK| {dk:.2e}",
          flush=True)
    assert dj < 1e-9 and dk < 1e-9


def test_tuner_keeps_mfma_only_where_faster_and_results_hold(torch_cuda):
    """dft_plan_tune times the matrix-core variant of each heavy
    combination's FP64 part and reports m = 1 where it keeps it; forcing
    every heavy combination onto it through dft_plan_set_shapes and forcing
    it off both give the untuned J/K."""
    from dft_amd import _lib
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan, set_shapes, tune
    from dft_amd.geometry import alkane_bohr, synthetic_spd_basis
    torch = torch_cuda
    Z, xyz = alkane_bohr(4)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    Pa = torch.as_tensor(_sym(plan.nbf, 4, 0.05), device="cuda")
    P = 2 * Pa
    J0, K0, _ = plan.jk(P, Pa, mode="fp64")
    rep = tune(plan, P, Pa, reps=3)
    present = [i for i in range(rep.shape[0]) if rep[i, 1]]
    heavy, light = [], []
    for i in present:  # set_shapes accepts m = 1 for the heavy combinations only
        r = rep.copy()
        r[:, 0] = 0
        r[i, 0] = 1
        try:
            set_shapes(plan, r)
            heavy.append(i)
        except _lib.DFTError:
            light.append(i)
    assert heavy and light and heavy[0] >= 8
    assert set(rep[heavy, 0].tolist()) <= {0, 1} and not rep[light, 0].any()
    print(f"tuner: matrix cores kept for combinations {[i for i in heavy if rep[i, 0]]} of {heavy}", flush=True)
    for force in (1, 0):
        r = rep.copy()
        r[heavy, 0] = force
        set_shapes(plan, r)
        J, K, _ = plan.jk(P, Pa, mode="fp64")
        assert float((J - J0).abs().max()) < 1e-9 and float((K - K0).abs().max()) < 1e-9
