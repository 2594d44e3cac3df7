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
