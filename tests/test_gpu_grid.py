"""Grid evaluation (the reference's density post-processing): HIP AO values
against the oracle, the reference's file layouts, and the electron count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _water():
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import water
    Z, xyz = water()
    shells = build_shells(load_basis(), Z, xyz)
    return shells, Z, np.asarray(xyz), Plan(shells, Z, xyz, schwarz_tau=1e-12)


@pytest.mark.parametrize("spd", [False, True])
def test_ao_values_match_oracle(torch_cuda, spd):
    import oracle.md_oracle as O
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.engine import Plan
    from dft_amd.geometry import methane, synthetic_spd_basis
    if spd:
        Z, xyz = methane()
        xyz = np.asarray(xyz)
        shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
        plan = Plan(shells, Z, xyz, schwarz_tau=1e-12)
    else:
        shells, Z, xyz, plan = _water()
    rng = np.random.default_rng(3)
    # points around the molecule and close to each nucleus
    g = np.concatenate([rng.uniform(xyz.min(0) - 3, xyz.max(0) + 3, size=(15000, 3)),
                        (xyz[rng.integers(0, len(xyz), 5000)] + rng.normal(0, 0.05, size=(5000, 3)))])
    g = g.astype(np.float32)
    A = plan.ao_values(g).cpu().numpy()
    R = O.ao_values(shells, g.astype(np.float64))
    assert A.shape == R.shape
    assert np.abs(A - R).max() <= 1e-12 * np.abs(R).max()


def test_density_integrates_to_electron_count(torch_cuda):
    """sum_g rho(r_g) h^3 over a box with h = 0.06 bohr gives the 10
    electrons of water (RHF density of the SCF)."""
    torch = torch_cuda
    from dft_amd.grid import box_grid
    from dft_amd.scf import scf_diis
    shells, Z, xyz, plan = _water()
    res = scf_diis(plan, Z, xyz, 5, 5)
    assert res.converged
    P = (res.P_a + res.P_b).contiguous()
    lo, hi = xyz.min(0) - 7.0, xyz.max(0) + 7.0
    g, shape = box_grid(lo, hi, 0.06)
    dv = float(np.prod((hi - lo) / (np.asarray(shape) - 1)))  # linspace spacing per axis
    n = 0.0
    for a in range(0, len(g), 1 << 21):
        ao = plan.ao_values(g[a:a + (1 << 21)])
        n += float(((P @ ao) * ao).sum())
    assert abs(n * dv - 10.0) < 1e-4, n * dv


def test_written_files_follow_the_reference_layout(torch_cuda, tmp_path):
    """write_orbital_densities (shelled_eval.py:61-133) and write_densities
    (main.py:880-924): file names, FP32 dtype, shapes, and values equal to
    the formulas there evaluated with the oracle's AO values."""
    import oracle.md_oracle as O
    from dft_amd.grid import reference_grid, write_densities, write_orbital_densities
    from dft_amd.scf import scf_diis
    shells, Z, xyz, plan = _water()
    res = scf_diis(plan, Z, xyz, 5, 5)
    C = res.C_a.cpu().numpy()
    # the reference pads y by -0.75 bohr (a box fitted to its molecule); water
    # needs a positive padding for a non-empty box
    g, shape = reference_grid(xyz, spacing=0.2, y_padding=1.0)
    d1 = tmp_path / "eval"
    nmo, npts = write_orbital_densities(plan, g, C, C, 5, 5, out_dir=str(d1), chunk=4000)
    assert (nmo, npts) == (5, len(g)) and np.prod(shape) == len(g)
    grid = np.load(d1 / "grid.npy")
    tot = np.load(d1 / "total_density.npy", mmap_mode="r")
    ph = np.load(d1 / "phase_density.npy", mmap_mode="r")
    assert grid.dtype == np.float32 and tot.dtype == np.float32 and ph.dtype == np.float32
    assert tot.shape == (5, len(g)) and ph.shape == (5, len(g))
    ao = O.ao_values(shells, grid.astype(np.float64))
    psi = C[:, :5].astype(np.float32).T.astype(np.float64) @ ao
    assert np.abs(ph - psi).max() <= 1e-6 * np.abs(psi).max()
    assert np.abs(tot - 2 * psi * psi).max() <= 1e-6 * np.abs(2 * psi * psi).max()
    d2 = tmp_path / "main"
    Pa = res.P_a.cpu().numpy()
    write_densities(plan, g, Pa, Pa, C, out_dir=str(d2), chunk=4000)
    rt = np.load(d2 / "total_density.npy")
    rs = np.load(d2 / "spin_density.npy")
    ro = np.load(d2 / "orbital_density.npy")
    rho = np.einsum("mn,mg,ng->g", 2 * Pa, ao, ao)
    assert rt.shape == (len(g),) and rs.shape == (len(g),) and ro.shape == (C.shape[1], len(g))
    assert np.abs(rt - rho).max() <= 1e-6 * np.abs(rho).max() and np.abs(rs).max() <= 1e-6 * np.abs(rho).max()
    assert np.abs(ro - (C.T @ ao) ** 2).max() <= 1e-6 * ((C.T @ ao) ** 2).max()
