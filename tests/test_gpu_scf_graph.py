"""The captured SCF iteration (scf_diis_graph: one hipGraph replay per
iteration, loop state on the device) against the eager damped + DIIS loop
(scf_diis, itself checked against the oracle's restatement of
shelled.py:1103-1292) and the golden energies."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _plan(Z, xyz, tau=0.0):
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    return Plan(build_shells(load_basis(), Z, xyz), Z, xyz, schwarz_tau=tau)


@pytest.mark.parametrize("name,nocc", [("water_sto3g.npz", 5), ("dimer_sto3g.npz", 10)])
def test_graph_scf_follows_the_eager_loop(torch_cuda, name, nocc):
    from dft_amd.scf import scf_diis, scf_diis_graph
    g = golden(name)
    Z, xyz = list(g["Z"]), g["xyz"]
    plan = _plan(Z, xyz)
    eager = scf_diis(plan, Z, xyz, nocc, nocc)
    gr = scf_diis_graph(plan, Z, xyz, nocc)
    assert eager.converged and gr.converged
    n = min(len(eager.energies), len(gr.energies))
    d = np.abs(np.asarray(eager.energies[:n]) - np.asarray(gr.energies[:n]))
    print(f"{name}: eager {eager.iterations} / graph {gr.iterations} iterations, max|dE| first 10 "
          f"{d[:10].max():.2e}, final dE {gr.energy - eager.energy:.2e}", flush=True)
    assert d[:10].max() < 1e-10  # damped iterations: the same schedule
    assert abs(gr.iterations - eager.iterations) <= 2
    assert abs(gr.energy - eager.energy) < 1e-9
    assert abs(gr.energy - float(g["E"])) < 1e-7


def test_graph_scf_water16_converges_to_the_eager_energy(torch_cuda):
    from dft_amd.geometry import water_cluster
    from dft_amd.scf import scf_diis, scf_diis_graph
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    plan = _plan(Z, xyz, tau=1e-10)
    eager = scf_diis(plan, Z, xyz, 80, 80, e_tol=1e-10, p_tol=1e-8)
    gr = scf_diis_graph(plan, Z, xyz, 80, e_tol=1e-10, p_tol=1e-8)
    assert eager.converged and gr.converged
    print(f"(H2O)16: eager {eager.iterations} / graph {gr.iterations} iterations, E {gr.energy:.10f}, "
          f"dE {gr.energy - eager.energy:.2e}, wall/iteration eager {np.median(eager.iter_times) * 1e3:.3f} ms, "
          f"graph {np.median(gr.iter_times) * 1e3:.3f} ms", flush=True)
    assert abs(gr.energy - eager.energy) < 1e-9


def test_graph_scf_with_eigensolve_tolerance_reaches_the_same_energy(torch_cuda):
    """eig_tol = 1e-9 (per-iteration eigensolves stopped at off-diagonal
    entries of that size, exact exit eigensolve): the same converged energy
    as exact eigensolves on the water dimer and on (H2O)16, both loops run to
    1e-10 Eh and 1e-8 in the density (with the default 1e-8 Eh the two runs
    stop at different iterations up to about 2e-8 Eh apart)."""
    from dft_amd.geometry import water_cluster
    from dft_amd.scf import scf_diis_graph
    g = golden("dimer_sto3g.npz")
    cases = [(list(g["Z"]), g["xyz"], 10, 0.0), (*water_cluster((4, 2, 2), 3.1, seed=0), 80, 1e-10)]
    for Z, xyz, nocc, tau in cases:
        plan = _plan(Z, xyz, tau=tau)
        ex = scf_diis_graph(plan, Z, xyz, nocc, e_tol=1e-10, p_tol=1e-8)
        tl = scf_diis_graph(plan, Z, xyz, nocc, e_tol=1e-10, p_tol=1e-8, eig_tol=1e-9)
        assert ex.converged and tl.converged
        print(f"nocc {nocc}: exact {ex.iterations} it {np.median(ex.iter_times) * 1e3:.3f} ms/it, tol 1e-9 "
              f"{tl.iterations} it {np.median(tl.iter_times) * 1e3:.3f} ms/it, dE {tl.energy - ex.energy:.2e}",
              flush=True)
        assert abs(tl.energy - ex.energy) < 1e-9


def test_scf_dense_kernels_match_torch(torch_cuda):
    """The tile kernels of the graph iteration against torch FP64 products on
    random matrices: dft_scf_gemm in its four transposition forms (with a
    shorter inner dimension and the fused copy), dft_scf_fock_residual and
    dft_scf_orbital_matrix, orders 1, 17, 112 and 128 (to rounding: 1e-13
    relative to the largest entry)."""
    from dft_amd.linalg import scf_fock_residual_, scf_gemm_, scf_orbital_matrix_
    torch = torch_cuda
    g = torch.Generator().manual_seed(7)

    def rnd(*shape):
        return torch.randn(*shape, generator=g, dtype=torch.float64).cuda()

    def close(a, b):
        return (a - b).abs().max().item() <= 1e-13 * max(b.abs().max().item(), 1.0)

    for n in (1, 17, 112, 128):
        A, B, C = rnd(n, n), rnd(n, n), torch.empty(n, n, dtype=torch.float64, device="cuda")
        for ta in (False, True):
            for tb in (False, True):
                src, dst = rnd(n, n), torch.zeros(n, n, dtype=torch.float64, device="cuda")
                scf_gemm_(A, B, C, trans_a=ta, trans_b=tb, copy=(src, dst))
                ref = (A.T if ta else A) @ (B.T if tb else B)
                assert close(C, ref), (n, ta, tb)
                assert torch.equal(dst, src)
        k = max(1, n // 3)
        scf_gemm_(A, A, C, trans_b=True, k=k)
        assert close(C, A[:, :k] @ A[:, :k].T), n
        H, J, K, PS = rnd(n, n), rnd(n, n), rnd(n, n), rnd(n, n)
        F, R = torch.empty_like(H), torch.empty_like(H)
        scf_fock_residual_(H, J, K, PS, F, R)
        Fr = (H + J) - K
        assert torch.equal(F, Fr)
        assert close(R, Fr @ PS - PS.T @ Fr), n
        XQ, T, Q = rnd(n, n), rnd(n, n), rnd(n, n)
        Am, V = torch.empty_like(H), torch.empty_like(H)
        scf_orbital_matrix_(XQ, T, Am, Q, V)
        At = XQ.T @ T
        assert close(Am, 0.5 * (At + At.T)), n
        assert torch.equal(Am, Am.T) and torch.equal(V, Q)


def test_graph_scf_stopped_early_reports_the_iterations_it_ran(torch_cuda):
    """The loop queues the next replay before it reads the current state, so
    a run stopped by max_count has run one replay past the last one read: it
    reports that iteration too (energies, count), and its energies are the
    first ones of the converged run."""
    from dft_amd.scf import scf_diis_graph
    g = golden("water_sto3g.npz")
    Z, xyz = list(g["Z"]), g["xyz"]
    plan = _plan(Z, xyz)
    full = scf_diis_graph(plan, Z, xyz, 5)
    part = scf_diis_graph(plan, Z, xyz, 5, max_count=4)
    assert full.converged and not part.converged
    assert part.iterations == len(part.energies), (part.iterations, len(part.energies))
    assert 5 <= len(part.energies) < len(full.energies)
    k = len(part.energies)
    assert np.abs(np.asarray(part.energies) - np.asarray(full.energies[:k])).max() < 1e-12
