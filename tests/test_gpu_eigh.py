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
    # sweeps: from scratch vs from the nearby eigenvectors
    from dft_amd.linalg import syevj_
    sw = torch.zeros(2, dtype=torch.int32, device="cuda")
    for k, (M0, init) in enumerate(((A2, True), (None, False))):
        Ad = torch.as_tensor(A2, device="cuda").reshape(1, n, n).clone()
        Vd = torch.eye(n, dtype=torch.float64, device="cuda").reshape(1, n, n).clone()
        if not init:
            Vd[0] = V0
            Ad[0] = V0.T @ Ad[0] @ V0
            Ad[0] = 0.5 * (Ad[0] + Ad[0].T)
        wd = torch.empty((1, n), dtype=torch.float64, device="cuda")
        syevj_(Ad.contiguous(), wd, Vd, init, sweeps=sw[k:k + 1])
    cold, warm = (int(x) for x in sw.cpu())
    print(f"Jacobi sweeps for n = {n}: {cold} from the identity, {warm} from nearby eigenvectors", flush=True)
    assert warm < cold <= 40


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


def test_syevj_time_per_call_is_reported(torch_cuda):
    """Device time of one eigendecomposition at the (H2O)16 order (112),
    from the identity and from the eigenvectors of a nearby matrix."""
    from dft_amd.linalg import syevj_
    torch = torch_cuda
    n = 112
    rng = np.random.default_rng(3)
    M = rng.standard_normal((n, n))
    A = torch.as_tensor(M + M.T, device="cuda")
    w = torch.empty((1, n), dtype=torch.float64, device="cuda")
    V = torch.empty((1, n, n), dtype=torch.float64, device="cuda")
    sw = torch.zeros(1, dtype=torch.int32, device="cuda")
    for init in (True, False):
        Ad = A.reshape(1, n, n).clone()
        if not init:
            Q = V.clone()
            Ad = (Q.transpose(1, 2) @ (A.reshape(1, n, n) + 1e-4) @ Q).contiguous()
            Ad = (0.5 * (Ad + Ad.transpose(1, 2))).contiguous()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        src = Ad.clone()
        Vs = V.clone()
        e0.record()
        for _ in range(5):
            Ad.copy_(src)
            if not init:
                V.copy_(Vs)
            syevj_(Ad, w, V, init, sweeps=sw)
        e1.record()
        torch.cuda.synchronize()
        print(f"syevj n = {n}, {'identity' if init else 'warm'} start: {e0.elapsed_time(e1) / 5:.3f} ms per call, "
              f"{int(sw.cpu()[0])} sweeps", flush=True)


def test_syevj_tolerance_leaves_only_small_entries(torch_cuda):
    """dft_syevj_tol: from the eigenvectors of a nearby matrix (the SCF warm
    start), a tolerance of 1e-9 on the off-diagonal entries takes fewer
    sweeps than machine precision, the eigenvectors stay orthonormal to
    rounding, and V^T A V differs from diag(w) by no more than the
    tolerance."""
    from dft_amd.linalg import eigh, syevj_
    torch = torch_cuda
    n = 112
    rng = np.random.default_rng(5)
    M = rng.standard_normal((n, n))
    A = M + M.T
    E = rng.standard_normal((n, n)) * 1e-6
    A2 = A + E + E.T
    _, V0 = eigh(torch.as_tensor(A, device="cuda"))
    out = {}
    for tol in (0.0, 1e-9):
        Q = V0.reshape(1, n, n).contiguous().clone()
        Ad = (Q[0].T @ torch.as_tensor(A2, device="cuda") @ Q[0])
        Ad = (0.5 * (Ad + Ad.T)).reshape(1, n, n).contiguous()
        w = torch.empty((1, n), dtype=torch.float64, device="cuda")
        sw = torch.zeros(1, dtype=torch.int32, device="cuda")
        syevj_(Ad, w, Q, False, sweeps=sw, tol=tol)
        out[tol] = (w[0].cpu().numpy(), Q[0].cpu().numpy(), int(sw.cpu()[0]))
    w, V, s1 = out[1e-9]
    s0 = out[0.0][2]
    assert s1 < s0, (s1, s0)
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
    D = V.T @ A2 @ V
    assert np.abs(D - np.diag(w)).max() < 1e-9 + 1e-12 * np.abs(A2).max() * n
    assert np.abs(w - out[0.0][0]).max() < 1e-12 * np.abs(A2).max() * n


@pytest.mark.parametrize("n", [1, 7, 58, 97, 112, 128])
def test_rotation_log_path_equals_the_workgroup_path(torch_cuda, n):
    """dft_syevj_ws (rotations logged, eigenvectors updated by a second
    kernel one row per wave) against dft_syevj_tol (eigenvectors updated in
    the matrix workgroup, in LDS up to n = 114 and in global memory above):
    the same rotations in the same order and arithmetic, so the same
    eigenvalues, sweeps and eigenvectors bit for bit; cold starts and warm
    starts from a nearby matrix's eigenvectors, two matrices per batch."""
    from dft_amd.linalg import syevj_
    torch = torch_cuda
    rng = np.random.default_rng(100 + n)
    M = rng.standard_normal((2, n, n))
    A = M + M.transpose(0, 2, 1)
    Ad = torch.as_tensor(A, device="cuda")
    res = {}
    for log in (False, True):
        A1 = Ad.clone()
        w = torch.empty((2, n), dtype=torch.float64, device="cuda")
        V = torch.empty_like(A1)
        sw = torch.zeros(2, dtype=torch.int32, device="cuda")
        syevj_(A1, w, V, True, sweeps=sw, log=log)
        # warm start: the eigenvectors of A for A + a small symmetric change
        E = torch.as_tensor(rng.standard_normal((2, n, n)) * 1e-4 if not log else res[False][4], device="cuda")
        A2 = Ad + E + E.transpose(1, 2)
        Q = V.clone()
        B = Q.transpose(1, 2) @ A2 @ Q
        B = (0.5 * (B + B.transpose(1, 2))).contiguous()
        w2 = torch.empty_like(w)
        sw2 = torch.zeros_like(sw)
        syevj_(B, w2, Q, False, sweeps=sw2, log=log)
        res[log] = (w, V, sw, (w2, Q, sw2), E.cpu().numpy())
        torch.cuda.synchronize()
    a, b = res[False], res[True]
    assert torch.equal(a[2], b[2]) and torch.equal(a[3][2], b[3][2])
    assert torch.equal(a[0], b[0]) and torch.equal(a[3][0], b[3][0])
    assert torch.equal(a[1], b[1]), (a[1] - b[1]).abs().max().item()
    assert torch.equal(a[3][1], b[3][1]), (a[3][1] - b[3][1]).abs().max().item()
    for k in range(2):
        _check(A[k], b[0][k].cpu().numpy(), b[1][k].cpu().numpy())
