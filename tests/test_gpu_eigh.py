"""The batched Jacobi eigensolver of the SCF step (dft_syevj) against the
mathematics of a symmetric eigendecomposition and against LAPACK (NumPy, on
the host): eigenvalues, residuals A V - V diag(w), orthonormality, the
ascending order, odd orders (padded), degenerate spectra, warm starts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _check(A, w, V, tol=1e-12):
    nrm = max(np.abs(A).max(), 1.0)
    wl = np.linalg.eigvalsh(A)
    assert np.abs(w - wl).max() < tol * nrm * A.shape[0], np.abs(w - wl).max()
    assert (np.diff(w) >= 0).all()
    assert np.abs(A @ V - V * w[None, :]).max() < tol * nrm * A.shape[0]
    assert np.abs(V.T @ V - np.eye(A.shape[0])).max() < 1e-12


@pytest.mark.parametrize("n", [1, 2, 7, 58, 96, 97, 112, 128])
def test_syevj_random_batches(torch_cuda, n):
    from dft_amd.linalg import eigh
    torch = torch_cuda
    rng = np.random.default_rng(n)
    M = rng.standard_normal((3, n, n))
    A = M + M.transpose(0, 2, 1)
    w, V = eigh(torch.as_tensor(A, device="cuda"))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    for b in range(3):
        _check(A[b], w[b], V[b])


def test_syevj_degenerate_spectrum(torch_cuda):
    """Identity plus a rank-2 term: n - 2 equal eigenvalues."""
    from dft_amd.linalg import eigh
    torch = torch_cuda
    n = 40
    rng = np.random.default_rng(1)
    u = rng.standard_normal((n, 2))
    A = np.eye(n) * 3.0 + u @ u.T
    w, V = eigh(torch.as_tensor(A, device="cuda"))
    _check(A, w.cpu().numpy(), V.cpu().numpy())


def test_syevj_warm_start_from_nearby_eigenvectors(torch_cuda):
    from dft_amd.linalg import eigh
    torch = torch_cuda
    n = 112
    rng = np.random.default_rng(2)
    M = rng.standard_normal((n, n))
    A = M + M.T
    E = rng.standard_normal((n, n)) * 1e-3
    A2 = A + E + E.T
    _, V0 = eigh(torch.as_tensor(A, device="cuda"))
    w, V = eigh(torch.as_tensor(A2, device="cuda"), V0=V0)
    _check(A2, w.cpu().numpy(), V.cpu().numpy())


def test_syevj_fock_matrix_of_water(torch_cuda):
    """X^T F X of converged water/STO-3G: the orbital energies of LAPACK."""
    from conftest import golden
    import oracle.md_oracle as O
    from dft_amd.linalg import eigh
    torch = torch_cuda
    g = golden("water_sto3g.npz")
    P = g["P_a"]
    J, K, _ = O.jk_dense(g["ERI"], 2 * P, P, P)
    F = g["T"] + g["V"] + J - K
    X = O.orthogonaliser(g["S"])
    Fp = X.T @ F @ X
    w, V = eigh(torch.as_tensor(Fp, device="cuda"))
    _check(Fp, w.cpu().numpy(), V.cpu().numpy(), tol=1e-13)
