"""Drop-in call surface of the reference's integral and Fock-build path.

Same names, argument meaning and in-place behaviour as the reference's
module-level functions and CUDA launches, backed by libdft_hip.so:

* ``nuclear_kernel`` / ``eri_kernel``   reference-layout launches
  (shelled.py:350-532; flat variants main.py:392-563), CuPy-style
  ``kernel(grid, block, args)`` on device tensors.
* ``calc_eri`` / ``process_eri``        the on-disk ERI store
  (shelled.py:596-768), computed by the contracted-quartet kernels.
* ``fill_matrices``                     J, K_a, K_b accumulation
  (shelled.py:770-877), computed by the fused direct J/K build instead of
  reading the store.
* ``normal``, ``combinations``, ``double_factorial``, ``overlap``,
  ``T_raw``, ``nuclear_repulsion``, ``independent``, ``init_UHF_density``,
  ``UHF_density``, ``total_spin``.

Device arrays are PyTorch-ROCm tensors (the reference's are CuPy arrays);
NumPy inputs are accepted and moved to the device. Nothing here computes
integrals on the CPU: every integral goes through the HIP library and fails
loudly if it is not built or no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from .basis import combinations, double_factorial, normal  # noqa: F401  (re-exported names)
from .engine import Plan
from .scf import UHF_density, init_UHF_density, nuclear_repulsion  # noqa: F401

MAX_E = 7


def _dev(x, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device="cuda").contiguous()


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def total_spin(unpaired_elec):
    """Minimum |S| over sign choices of per-atom unpaired electrons (shelled.py:536-546)."""
    u = np.asarray(unpaired_elec, dtype=np.int64).ravel()
    sums = np.zeros(1, dtype=np.int64)
    for x in u:
        sums = np.concatenate([sums + x, sums - x])
    return np.abs(sums).min() / 2


# ------------------------------------------------- reference-layout kernels
class _Launch:
    """CuPy RawKernel-like callable: ``kernel(grid, block, args)``.

    grid and block are accepted for signature compatibility; the library
    chooses its own launch shape and covers every output element.
    """

    def __init__(self, name):
        self.name = name

    def __call__(self, grid, block, args):
        lib = _lib.load()
        if self.name == "nuclear_kernel":
            if len(args) == 20:   # shelled.py:350-371, one centre per side
                (ea, eb, ca, cb, pa, pb, Z, xyz, natom, nfa, nfb, V, Ex, Ey, Ez, p, P, nx, ny, nz) = args
                sa = sb = 0
            elif len(args) == 16:  # main.py:392-409, flat: centre per function
                (ea, ca, pa, Z, xyz, natom, nfa, V, Ex, Ey, Ez, p, P, nx, ny, nz) = args
                eb, cb, pb, nfb, sa, sb = ea, ca, pa, nfa, 3, 3
            else:
                raise TypeError("nuclear_kernel takes 20 (shelled) or 16 (flat) arguments")
            keep = [_dev(ea, torch.float64), _dev(eb, torch.float64), _dev(ca, torch.float64),
                    _dev(cb, torch.float64), _dev(pa, torch.int32), _dev(pb, torch.int32),
                    _dev(Z, torch.int32), _dev(xyz, torch.float64)]
            outs = (V, Ex, Ey, Ez, p, P, nx, ny, nz)
            for o in outs:
                if not (isinstance(o, torch.Tensor) and o.is_cuda and o.is_contiguous()):
                    raise TypeError("outputs must be contiguous device tensors")
            _lib.check(lib.dft_nuclear_prim(int(nfa), int(nfb), *(k.data_ptr() for k in keep[:3]), sa,
                                            keep[3].data_ptr(), sb, keep[4].data_ptr(), keep[5].data_ptr(),
                                            int(natom), keep[6].data_ptr(), keep[7].data_ptr(),
                                            *(o.data_ptr() for o in outs), _stream()))
        elif self.name == "eri_kernel":
            if len(args) == 19:   # shelled.py:452-473
                (Kab, pab, Pab, Exab, Eyab, Ezab, nxab, nyab, nzab,
                 Kcd, pcd, Pcd, Excd, Eycd, Ezcd, nxcd, nycd, nzcd, out) = args
                es, row0 = MAX_E, 0
                nab = int(Kab)
            elif len(args) == 15:  # main.py:490-507: rows [row_start, row_start + num_current/K)
                (row0, K, num, pk, Pk, Ex, Ey, Ez, nx, ny, nz, sx, sy, sz, out) = args
                if not (int(sx) == int(sy) == int(sz)):
                    raise ValueError("E strides must agree")
                Kcd = K
                pab = pcd = pk
                Pab = Pcd = Pk
                Exab = Excd = Ex
                Eyab = Eycd = Ey
                Ezab = Ezcd = Ez
                nxab = nxcd = nx
                nyab = nycd = ny
                nzab = nzcd = nz
                es, row0 = int(sx), int(row0)
                nab = int(num) // int(K)
            else:
                raise TypeError("eri_kernel takes 19 (shelled) or 15 (flat) arguments")
            d = [_dev(x, torch.float64) for x in (pab, Pab, Exab, Eyab, Ezab)]
            di = [_dev(x, torch.int32) for x in (nxab, nyab, nzab)]
            k = [_dev(x, torch.float64) for x in (pcd, Pcd, Excd, Eycd, Ezcd)]
            ki = [_dev(x, torch.int32) for x in (nxcd, nycd, nzcd)]
            if not (isinstance(out, torch.Tensor) and out.is_cuda and out.is_contiguous()):
                raise TypeError("eri_output must be a contiguous device tensor")
            off = [row0, 3 * row0, es * row0, es * row0, es * row0]
            bra = [x.data_ptr() + o * 8 for x, o in zip(d, off)] + [x.data_ptr() + row0 * 4 for x in di]
            _lib.check(lib.dft_eri_prim(nab, *bra, int(Kcd), *(x.data_ptr() for x in k),
                                        *(x.data_ptr() for x in ki), es, out.data_ptr(), _stream()))
        else:
            raise ValueError(self.name)


nuclear_kernel = _Launch("nuclear_kernel")
eri_kernel = _Launch("eri_kernel")


# ---------------------------------------------------- 1-D host helpers
def overlap(alpha, beta, cen_a, cen_b, p1, p2):
    """1-D primitive overlaps [len(alpha), len(beta)] (shelled.py:104-146).

    Evaluated on the GPU as E_0^{p1,p2} sqrt(pi/p) from the pair pass of
    ``dft_nuclear_prim`` (equal to the reference's binomial closed form).
    """
    a = _dev(alpha, torch.float64).reshape(-1)
    b = _dev(beta, torch.float64).reshape(-1)
    na, nb = a.numel(), b.numel()
    pa = torch.zeros((na, 3), dtype=torch.int32, device="cuda")
    pb = torch.zeros((nb, 3), dtype=torch.int32, device="cuda")
    pa[:, 0] = _dev(p1, torch.int32).reshape(-1)
    pb[:, 0] = _dev(p2, torch.int32).reshape(-1)
    ca = torch.zeros(3, dtype=torch.float64, device="cuda")
    cb = torch.zeros(3, dtype=torch.float64, device="cuda")
    ca[0] = float(cen_a)
    cb[0] = float(cen_b)
    K = na * nb
    V = torch.empty(K, dtype=torch.float64, device="cuda")
    E = [torch.empty(K * MAX_E, dtype=torch.float64, device="cuda") for _ in range(3)]
    p = torch.empty(K, dtype=torch.float64, device="cuda")
    P = torch.empty(3 * K, dtype=torch.float64, device="cuda")
    n = [torch.empty(K, dtype=torch.int32, device="cuda") for _ in range(3)]
    nuclear_kernel(None, None, (a, b, ca, cb, pa, pb, torch.zeros(1, dtype=torch.int32),
                                torch.zeros(3, dtype=torch.float64), 0, na, nb, V, E[0], E[1], E[2], p, P,
                                n[0], n[1], n[2]))
    return (E[0].reshape(K, MAX_E)[:, 0] * torch.sqrt(torch.pi / p)).reshape(na, nb)


def T_raw(exp_a, exp_b, cen_a, cen_b, pow_a, pow_b, prev_overlap):
    """Per-axis kinetic piece (shelled.py:149-155), before the -1/2 and the
    products with the other axes' overlaps."""
    eb = _dev(exp_b, torch.float64).reshape(-1)
    pb = _dev(pow_b, torch.int32).reshape(-1)
    prev = _dev(prev_overlap, torch.float64)
    res = -2 * eb[None, :] * (2 * pb[None, :] + 1) * prev
    res = res + 4 * eb[None, :] ** 2 * overlap(exp_a, eb, cen_a, cen_b, pow_a, pb + 2)
    idx = torch.nonzero(pb >= 2).reshape(-1)
    if idx.numel():
        res[:, idx] += (pb[idx] * (pb[idx] - 1))[None, :] * overlap(exp_a, eb[idx], cen_a, cen_b, pow_a,
                                                                     pb[idx] - 2)
    return res


# ------------------------------------------------ Cartesian -> spherical
_S3, _S2, _S5, _S6, _S10, _S30 = (float(np.sqrt(v)) for v in (3, 2, 5, 6, 10, 30))
# rows of the real-solid-harmonic transform per shell, as (spherical row, [(cart col, coef)])
_SPH = {
    2: [[(1, 1.0)], [(3, 1.0)], [(4, 1.0)], [(0, _S3 / 2), (2, -_S3 / 2)], [(5, 1.0), (0, -0.5), (2, -0.5)]],
    3: [[(4, -3 * _S5 / 10), (6, -3 * _S5 / 10), (9, 1.0)],
        [(0, -_S6 / 4), (2, -_S30 / 20), (7, _S30 / 5)],
        [(1, -_S30 / 20), (3, -_S6 / 4), (8, _S30 / 5)],
        [(4, _S3 / 2), (6, -_S3 / 2)],
        [(5, 1.0)],
        [(0, _S10 / 4), (2, -3 * _S2 / 4)],
        [(1, 3 * _S2 / 4), (3, -_S10 / 4)]],
}


def spherical_matrix(shells, sphere_total_ao):
    """A (n_sph x n_cart) such that independent(shells, n, C) == A @ C."""
    ncart = shells[-1]["ao_stop"]
    A = np.zeros((sphere_total_ao, ncart))
    for sh in shells:
        l = sh["angular_momentum"]
        nc = sh["n_components"]
        a, o = sh["ao_start"], sh["sphere_ao_start"]
        for _ in range(sh["n_ao"] // nc):
            if l <= 1:
                A[o:o + nc, a:a + nc] = np.eye(nc)
            elif l in _SPH:
                for r, terms in enumerate(_SPH[l]):
                    for col, coef in terms:
                        A[o + r, a + col] = coef
            else:
                raise ValueError("l > 3 is not in the reference")
            a += nc
            o += 2 * l + 1
    return A


def independent(shells, sphere_total_ao, C):
    """Cartesian -> real-spherical rows for d and f (shelled.py:878-920)."""
    A = spherical_matrix(shells, sphere_total_ao)
    if isinstance(C, torch.Tensor):
        return torch.as_tensor(A, dtype=C.dtype, device=C.device) @ C
    return A @ np.asarray(C)


# ---------------------------------------------------- ERI store + J/K
# Plans of the drop-in calls, most recently used last. A plan is found again
# only for shells of identical content (momenta, AO offsets, centres,
# exponents, coefficients), so a caller that moves atoms in place or reuses a
# list object gets a new plan; at most _PLAN_CACHE plans (and their device
# buffers) are kept.
_PLAN_CACHE = 2
_plans: list = []


def _shell_key(shells, uhf) -> bytes:
    import hashlib
    h = hashlib.sha1()
    for sh in shells:
        h.update(np.asarray([sh["angular_momentum"], sh["ao_start"], sh["n_contractions"],
                             sh["n_exponents"]], dtype=np.int64).tobytes())
        for k in ("atom_pos", "exponents", "coefficients"):
            h.update(np.ascontiguousarray(np.asarray(sh[k], dtype=np.float64)).tobytes())
    h.update(b"uhf" if uhf else b"rhf")
    return h.digest()


def _plan_for(shells, uhf):
    key = _shell_key(shells, bool(uhf))
    for k, (kk, plan) in enumerate(_plans):
        if kk == key:
            _plans.append(_plans.pop(k))
            return plan
    plan = Plan(shells, [], np.zeros((0, 3)), schwarz_tau=0.0, fp32_tau=0.0, uhf=uhf)
    _plans.append((key, plan))
    del _plans[:-_PLAN_CACHE]
    return plan


def calc_eri(pair_list, pair_cache, shells, out_dir="compute_eri"):
    """Write the contracted unique-quartet ERI store (shelled.py:596-700).

    Same files and layout: eri_values.npy (f64), eri_indices.npy ([start,
    stop] int32), eri_quads.npy and eri_quads_shapes.npy (int16), quartets in
    the order ab_idx >= cd_idx over ``pair_list``. ``pair_cache`` (the
    reference's nuclear_kernel records) is not needed and may be None.
    """
    os.makedirs(out_dir, exist_ok=True)
    plan = _plan_for(shells, False)
    quads = [(a, b, c, d) for ab, (a, b) in enumerate(pair_list) for (c, d) in pair_list[:ab + 1]]
    shapes = np.array([[shells[k]["n_ao"] for k in q] for q in quads], dtype=np.int64).reshape(-1, 4)
    sizes = shapes.prod(axis=1)
    N = int(sizes.sum())
    vals = np.lib.format.open_memmap(os.path.join(out_dir, "eri_values.npy"), mode="w+", dtype=np.float64,
                                     shape=(N,))
    idx = np.lib.format.open_memmap(os.path.join(out_dir, "eri_indices.npy"), mode="w+", dtype=np.int32,
                                    shape=(len(quads), 2))
    qd = np.lib.format.open_memmap(os.path.join(out_dir, "eri_quads.npy"), mode="w+", dtype=np.int16,
                                   shape=(len(quads), 4))
    qs = np.lib.format.open_memmap(os.path.join(out_dir, "eri_quads_shapes.npy"), mode="w+", dtype=np.int16,
                                   shape=(len(quads), 4))
    ends = np.cumsum(sizes)
    starts = ends - sizes
    chunk = 20000
    for c0 in range(0, len(quads), chunk):
        blocks = plan.eri_blocks(quads[c0:c0 + chunk])
        flat = np.concatenate([b.ravel() for b in blocks]) if blocks else np.zeros(0)
        vals[starts[c0]:starts[c0] + flat.size] = flat
    idx[:, 0], idx[:, 1] = starts, ends
    qd[:] = np.asarray(quads, dtype=np.int16).reshape(-1, 4)
    qs[:] = shapes
    for m in (vals, idx, qd, qs):
        m.flush()


def process_eri(out_dir="compute_eri"):
    """Sort the store by AO-shape tuple into blocks (shelled.py:702-768)."""
    j = lambda name: os.path.join(out_dir, name)  # noqa: E731
    vals = np.load(j("eri_values.npy"), mmap_mode="r")
    idx = np.load(j("eri_indices.npy"), mmap_mode="r")
    quads = np.load(j("eri_quads.npy"), mmap_mode="r")
    shapes = np.load(j("eri_quads_shapes.npy"), mmap_mode="r")
    order = np.lexsort((shapes[:, 3], shapes[:, 2], shapes[:, 1], shapes[:, 0]))
    pv = np.lib.format.open_memmap(j("processed_eri_values.npy"), mode="w+", dtype=np.float64, shape=vals.shape)
    pi = np.lib.format.open_memmap(j("processed_eri_indices.npy"), mode="w+", dtype=np.int32, shape=idx.shape)
    pq = np.lib.format.open_memmap(j("processed_eri_quads.npy"), mode="w+", dtype=np.int32, shape=quads.shape)
    pq[:] = quads[order]
    write = 0
    for k, (s, e) in enumerate(idx[order]):
        pv[write:write + e - s] = vals[s:e]
        pi[k] = (write, write + e - s)
        write += e - s
    ss = shapes[order]
    change = np.nonzero(np.any(ss[1:] != ss[:-1], axis=1))[0] + 1
    bounds = np.concatenate([[0], change, [len(ss)]])
    np.save(j("block_idxs.npy"), np.stack([bounds[:-1], bounds[1:]], axis=1))
    np.save(j("chunked_quad_shape.npy"), ss[bounds[:-1]])
    for m in (pv, pi, pq):
        m.flush()


def fill_matrices(shells, P, P_a, P_b, J_matrix, K_a, K_b, mode="fp64"):
    """J += J[P], K_a += K[P_a], K_b += K[P_b] in the Cartesian basis, in place
    (shelled.py:770-877).

    Computed by the fused direct J/K build (no ERI store is read). When P_a
    and P_b are equal one exchange matrix is built and added to both.

    The build applies the eight-fold quartet symmetry, which reads each
    density in one orientation; the reference reads all eight
    (shelled.py:791-800), which differs for a non-symmetric density, so such
    a density is rejected with ValueError.
    """
    Pt, Pa, Pb = (_dev(x, torch.float64) for x in (P, P_a, P_b))
    for name, M in (("P", Pt), ("P_a", Pa), ("P_b", Pb)):
        if M.dim() != 2 or M.shape[0] != M.shape[1]:
            raise ValueError(f"{name} must be a square matrix, got {tuple(M.shape)}")
        scale = float(M.abs().max()) if M.numel() else 0.0
        if float((M - M.T).abs().max()) > 1e-12 * max(scale, 1.0):
            raise ValueError(f"{name} is not symmetric: the direct build reads each density in one "
                             "orientation (eight-fold symmetry)")
    rhf = torch.equal(Pa, Pb)
    plan = _plan_for(shells, not rhf)
    J, Ka, Kb = plan.jk(Pt, Pa, None if rhf else Pb, mode=mode)
    Kb = Ka if rhf else Kb
    for dst, src in ((J_matrix, J), (K_a, Ka), (K_b, Kb)):
        if isinstance(dst, torch.Tensor):
            dst += src.to(device=dst.device, dtype=dst.dtype)
        else:
            dst += src.cpu().numpy()
