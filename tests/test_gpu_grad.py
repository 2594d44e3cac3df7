"""Analytic nuclear gradients (derivative-ERI kernels) against central finite
differences of the SCF energy (h = 1e-4 bohr), of both the GPU SCF and the
CPU oracle SCF (the reference's arithmetic), plus invariances and the
torch.autograd boundary. Bar: 1e-6 Eh/bohr."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-4


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _fd(fun, xyz, coords):
    out = []
    for (a, x) in coords:
        p = xyz.copy()
        p[a, x] += H
        m = xyz.copy()
        m[a, x] -= H
        out.append((fun(p) - fun(m)) / (2 * H))
    return np.array(out)


def test_water_gradient_vs_fd(torch_cuda):
    from dft_amd.basis import load_basis
    from dft_amd.geometry import water
    from dft_amd.grad import energy_and_gradient
    Z, xyz = water()
    xyz = xyz + np.random.default_rng(0).normal(0, 0.05, xyz.shape)
    basis = load_basis()
    E, g, _ = energy_and_gradient(Z, xyz, basis)
    coords = [(a, x) for a in range(3) for x in range(3)]
    fd = _fd(lambda r: energy_and_gradient(Z, r, basis)[0], xyz, coords)
    got = np.array([g[a, x] for a, x in coords])
    assert np.abs(got - fd).max() < 1e-6, np.abs(got - fd).max()
    assert np.abs(g.sum(axis=0)).max() < 1e-8  # translational invariance


def test_water_gradient_vs_oracle_fd(torch_cuda):
    import oracle.md_oracle as O
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.geometry import water
    from dft_amd.grad import energy_and_gradient
    Z, xyz = water()
    xyz = xyz + np.random.default_rng(1).normal(0, 0.05, xyz.shape)
    basis = load_basis()

    def oracle_energy(r):
        sh = build_shells(basis, Z, r)
        T = O.eri_tensor(sh)
        S, Tk, V = O.one_electron(sh, Z, r)
        res = O.scf_plain(Tk + V, S, lambda P, a, b: O.jk_dense(T, P, a, b), O.nuclear_repulsion(Z, r), 5, 5,
                          e_tol=1e-12, p_tol=1e-10, max_count=300)
        return res["E"]
    E, g, _ = energy_and_gradient(Z, xyz, basis)
    assert abs(E - oracle_energy(xyz)) < 1e-8
    coords = [(0, 0), (0, 2), (1, 1), (2, 0), (2, 2)]
    fd = _fd(oracle_energy, xyz, coords)
    got = np.array([g[a, x] for a, x in coords])
    assert np.abs(got - fd).max() < 1e-6, np.abs(got - fd).max()


def test_uhf_cation_gradient_vs_fd(torch_cuda):
    from dft_amd.basis import load_basis
    from dft_amd.geometry import water
    from dft_amd.grad import energy_and_gradient
    Z, xyz = water()
    xyz = xyz + np.random.default_rng(2).normal(0, 0.05, xyz.shape)
    basis = load_basis()
    E, g, res = energy_and_gradient(Z, xyz, basis, charge=1, mult=2)
    assert not np.allclose(res.P_a.cpu().numpy(), res.P_b.cpu().numpy())
    coords = [(0, 1), (1, 0), (2, 2)]
    fd = _fd(lambda r: energy_and_gradient(Z, r, basis, charge=1, mult=2)[0], xyz, coords)
    got = np.array([g[a, x] for a, x in coords])
    assert np.abs(got - fd).max() < 1e-6, np.abs(got - fd).max()


def test_ethane_gradient_vs_fd_subset(torch_cuda):
    from dft_amd.basis import load_basis
    from dft_amd.geometry import alkane_bohr
    from dft_amd.grad import energy_and_gradient
    Z, xyz = alkane_bohr(2, noise_sigma_ang=0.05, seed=3)
    basis = load_basis()
    E, g, _ = energy_and_gradient(Z, xyz, basis)
    coords = [(0, 0), (1, 1), (3, 2), (6, 0)]
    fd = _fd(lambda r: energy_and_gradient(Z, r, basis)[0], xyz, coords)
    got = np.array([g[a, x] for a, x in coords])
    assert np.abs(got - fd).max() < 1e-6, np.abs(got - fd).max()
    assert np.abs(g.sum(axis=0)).max() < 1e-8


def test_autograd_function(torch_cuda):
    torch = torch_cuda
    from dft_amd.basis import load_basis
    from dft_amd.geometry import water
    from dft_amd.grad import energy_and_gradient, scf_energy
    Z, xyz = water()
    basis = load_basis()
    coords = torch.tensor(xyz, requires_grad=True)
    E = scf_energy(coords, Z, basis)
    E.backward()
    E2, g, _ = energy_and_gradient(Z, xyz, basis)
    assert abs(float(E) - E2) < 1e-10
    assert np.allclose(coords.grad.numpy(), g, atol=1e-12)


def test_scf_guess_gives_same_energy_and_gradient(torch_cuda):
    """Seeding the SCF with the densities of a nearby geometry (as the
    steepest-descent driver does) converges to the same state."""
    from dft_amd.basis import load_basis
    from dft_amd.geometry import water
    from dft_amd.grad import energy_and_gradient
    Z, xyz = water()
    basis = load_basis()
    _, _, res0 = energy_and_gradient(Z, xyz + 0.02, basis)
    E1, g1, _ = energy_and_gradient(Z, xyz, basis)
    E2, g2, res2 = energy_and_gradient(Z, xyz, basis, guess=(res0.P_a, res0.P_b))
    assert abs(E1 - E2) < 1e-9 and np.abs(g1 - g2).max() < 1e-7
    assert res2.iterations <= 30


def test_atomic_densities_hold_the_atoms_electrons(torch_cuda):
    """Spherically averaged atomic RHF densities of the SAD guess: Tr(P S)
    equals the nuclear charge, P is symmetric, and hydrogen's single
    contracted 1s function holds one electron (P = 1 / S_11 = 1)."""
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.scf import atomic_density
    basis = load_basis()
    for z in (1, 6, 8):
        sh = build_shells(basis, [z], np.zeros((1, 3)))
        P = atomic_density(z, sh)
        S = Plan(sh, [z], np.zeros((1, 3)), schwarz_tau=0.0).one_electron()[0].cpu().numpy()
        assert abs(np.trace(P @ S) - z) < 1e-10, z
        assert np.abs(P - P.T).max() < 1e-12
        if z == 1:
            assert abs(P[0, 0] * S[0, 0] - 1.0) < 1e-12


def test_sad_and_core_guesses_give_the_same_state(torch_cuda):
    """The superposition-of-atomic-densities start (the default of
    energy_and_gradient) and the core-Hamiltonian start (the reference's)
    converge to the same energy and gradient, in both the graph loop
    (closed shell: ethane, water; there the SAD start takes no more
    iterations) and the eager UHF loop (doublet water cation)."""
    from dft_amd.basis import load_basis
    from dft_amd.geometry import alkane_bohr, water
    from dft_amd.grad import energy_and_gradient
    basis = load_basis()
    Z, xyz = alkane_bohr(2, noise_sigma_ang=0.05, seed=3)
    for Zm, xm, charge, mult in ((Z, xyz, 0, 1), (*water(), 0, 1), (*water(), 1, 2)):
        E1, g1, r1 = energy_and_gradient(Zm, xm, basis, charge, mult)
        E2, g2, r2 = energy_and_gradient(Zm, xm, basis, charge, mult, guess="core")
        assert abs(E1 - E2) < 1e-9 and np.abs(g1 - g2).max() < 1e-7, (charge, mult)
        if mult == 1:
            assert r1.iterations <= r2.iterations, (r1.iterations, r2.iterations)
