"""Symmetric eigendecomposition of the SCF step on the device.

``eigh`` is ``torch.linalg.eigh`` for the SCF matrices (eigenvalues in
ascending order, eigenvectors as columns) computed by the library's batched
two-sided Jacobi kernel (``dft_syevj``) for orders up to ``JACOBI_NMAX``:
one kernel launch per call, no host synchronisation, so it can run inside a
captured hipGraph. ``V0`` (the previous iteration's orbitals in the same
orthonormal basis) starts the rotations from that basis, which near
convergence leaves a nearly diagonal matrix and a sweep or two of work.
Larger matrices go to torch.linalg.eigh (rocSOLVER).

Reference: xp.linalg.eigh in the SCF loops (shelled.py:1082,1210-1213;
main.py:754,796-799).
"""
from __future__ import annotations

import ctypes as C
import ctypes as _ctypes  # for functions whose output matrix is named C

import torch

from . import _lib

JACOBI_NMAX = 128


# batches whose rotation logs fit this many bytes take the two-kernel path
# (dft_syevj_ws: the eigenvectors updated outside the matrix workgroup)
SYEVJ_LOG_MAX_BYTES = 64 << 20


def syevj_(A: torch.Tensor, w: torch.Tensor, V: torch.Tensor, init_v: bool, stream=None, sweeps=None,
           tol: float = 0.0, log=None) -> None:
    """In place: A (batch, n, n) contiguous float64 is destroyed, w (batch, n)
    and V (batch, n, n) receive the eigenpairs; with init_v False, V holds
    the starting orthonormal basis Q and A must be Q^T F Q. ``sweeps`` (an
    int32 device tensor of batch entries) receives the Jacobi sweeps.
    ``tol`` > 0: off-diagonal entries up to tol are left (dft_syevj_tol).
    ``log``: True / False forces the rotation-log path (dft_syevj_ws, a
    device workspace of batch x dft_syevj_ws_bytes(n) bytes) on or off; None
    takes it when that workspace is at most SYEVJ_LOG_MAX_BYTES."""
    lib = _lib.load()
    n = A.shape[-1]
    batch = A.numel() // (n * n) if n else 0
    for name, t, shape in (("A", A, (batch, n, n)), ("w", w, (batch, n)), ("V", V, (batch, n, n))):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == int(torch.Size(shape).numel())):
            raise ValueError(f"{name} must be a contiguous float64 device tensor with {shape} elements")
    s = torch.cuda.current_stream(A.device) if stream is None else stream
    sw = None
    if sweeps is not None:
        if not (sweeps.is_cuda and sweeps.dtype == torch.int32 and sweeps.numel() >= batch):
            raise ValueError("sweeps must be an int32 device tensor with batch entries")
        sw = sweeps.data_ptr()
    per = int(lib.dft_syevj_ws_bytes(n))
    if log is None:
        log = 0 < batch * per <= SYEVJ_LOG_MAX_BYTES
    work, nbytes = None, 0
    if log and batch > 0:
        nbytes = batch * per
        buf = torch.empty(nbytes, dtype=torch.uint8, device=A.device)
        if stream is not None:
            buf.record_stream(s)  # freed here, still read by the queued kernels
        work = buf.data_ptr()
    _lib.check(lib.dft_syevj_ws(n, batch, A.data_ptr(), w.data_ptr(), V.data_ptr(), int(bool(init_v)), sw,
                                float(tol), work, nbytes, C.c_void_p(s.cuda_stream)))


def eigh(F: torch.Tensor, V0: torch.Tensor | None = None, tol: float = 0.0):
    """(w, V) with F = V diag(w) V^T for one symmetric matrix or a batch;
    ``tol`` > 0 leaves off-diagonal entries up to tol (orbitals to about
    tol / gap, for SCF loops with a coarser convergence test)."""
    n = F.shape[-1]
    if n > JACOBI_NMAX or not F.is_cuda:
        return torch.linalg.eigh(F)
    batched = F.dim() == 3
    Fb = F if batched else F.unsqueeze(0)
    if V0 is None:
        A = Fb.contiguous().clone()
        V = torch.empty_like(A)
        init = True
    else:
        Q = V0 if batched else V0.unsqueeze(0)
        A = (Q.transpose(1, 2) @ Fb @ Q).contiguous()
        A = 0.5 * (A + A.transpose(1, 2))
        A = A.contiguous()
        V = Q.contiguous().clone()
        init = False
    w = torch.empty(A.shape[:2], dtype=A.dtype, device=A.device)
    syevj_(A, w, V, init, tol=tol)
    return (w, V) if batched else (w[0], V[0])


def _check_f64(name, t, rows, cols):
    if not (t.is_cuda and t.dtype == torch.float64 and t.dim() == 2 and t.shape[0] >= rows and t.shape[1] >= cols
            and t.stride(1) == 1):
        raise ValueError(f"{name} must be a float64 device matrix of at least {rows} x {cols} with unit column stride")


def scf_gemm_(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, trans_a: bool = False, trans_b: bool = False,
              k: int | None = None, copy=None, stream=None) -> None:
    """C = op(A) op(B) in place for orders up to 128 (dft_scf_gemm: one 16 x
    16 tile per workgroup); op(A) = A or A^T, op(B) = B or B^T, the inner
    dimension k defaults to all of it. ``copy`` = (src, dst): also dst = src
    (C's shape) in the same launch."""
    lib = _lib.load()
    m, nc = C.shape
    kd = k if k is not None else (A.shape[0] if trans_a else A.shape[1])
    _check_f64("C", C, m, nc)
    _check_f64("A", A, kd if trans_a else m, m if trans_a else kd)
    _check_f64("B", B, nc if trans_b else kd, kd if trans_b else nc)
    src = dst = None
    if copy is not None:
        src, dst = copy
        _check_f64("copy source", src, m, nc)
        _check_f64("copy target", dst, m, nc)
        if src.stride(0) != C.stride(0) or dst.stride(0) != C.stride(0):
            raise ValueError("copy source and target must have C's row stride")
        src, dst = src.data_ptr(), dst.data_ptr()
    s = torch.cuda.current_stream(C.device) if stream is None else stream
    _lib.check(lib.dft_scf_gemm(m, nc, kd, A.data_ptr(), A.stride(0), int(trans_a), B.data_ptr(), B.stride(0),
                                int(trans_b), C.data_ptr(), C.stride(0), src, dst, _ctypes.c_void_p(s.cuda_stream)))


def scf_fock_residual_(H, J, K, PS, F, R, stream=None) -> None:
    """F = (H + J) - K and R = F PS - PS^T F (dft_scf_fock_residual), n x n
    contiguous, n <= 128."""
    lib = _lib.load()
    n = H.shape[0]
    for name, t in (("H", H), ("J", J), ("K", K), ("PS", PS), ("F", F), ("R", R)):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.shape == (n, n)):
            raise ValueError(f"{name} must be a contiguous float64 device {n} x {n} matrix")
    s = torch.cuda.current_stream(H.device) if stream is None else stream
    _lib.check(lib.dft_scf_fock_residual(n, H.data_ptr(), J.data_ptr(), K.data_ptr(), PS.data_ptr(), F.data_ptr(),
                                         R.data_ptr(), _ctypes.c_void_p(s.cuda_stream)))


def scf_orbital_matrix_(XQ, T, A, Q, V, stream=None) -> None:
    """A = (XQ^T T + (XQ^T T)^T) / 2 and V = Q (dft_scf_orbital_matrix), n x n
    contiguous, n <= 128."""
    lib = _lib.load()
    n = XQ.shape[0]
    for name, t in (("XQ", XQ), ("T", T), ("A", A), ("Q", Q), ("V", V)):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.shape == (n, n)):
            raise ValueError(f"{name} must be a contiguous float64 device {n} x {n} matrix")
    s = torch.cuda.current_stream(XQ.device) if stream is None else stream
    _lib.check(lib.dft_scf_orbital_matrix(n, XQ.data_ptr(), T.data_ptr(), A.data_ptr(), Q.data_ptr(), V.data_ptr(),
                                          _ctypes.c_void_p(s.cuda_stream)))
