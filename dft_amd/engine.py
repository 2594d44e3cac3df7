"""Python front end of the HIP integral engine.

``Plan`` wraps one ``dft_plan`` of libdft_hip.so: it flattens the reference's
shell dicts (``shelled.py:58-77,938-952``) into the C-ABI arrays, splits
general contractions into segmented shells (a shell with k coefficient rows
becomes k shells whose AO blocks follow each other, which is exactly the
reference's AO = contraction * ncomp + component order), and exposes the
device entry points on PyTorch-ROCm tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .flops import combo_moments


def _ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def _flags(uhf: bool, split: bool, serial: bool = False, static_workspace: bool = False, mfma: bool = False) -> int:
    return ((_lib.DFT_PLAN_UHF if uhf else 0) | (0 if split else _lib.DFT_PLAN_NOSPLIT)
            | (_lib.DFT_PLAN_SERIAL if serial else 0)
            | (_lib.DFT_PLAN_STATIC_WORKSPACE if static_workspace else 0)
            | (_lib.DFT_PLAN_MFMA if mfma else 0))


class Plan:
    """Integral plan of one molecule (pair data, Schwarz bounds, quartet lists).

    Parameters follow ``dft_plan_create``: ``schwarz_tau`` drops quartets with
    Q_ab Q_cd below it; ``fp32_tau`` marks quartets below it for FP32 in the
    mixed mode; ``uhf`` builds K for two densities; ``split=False`` keeps
    one lane per quartet in every class (no primitive split of classes with
    few quartets), which makes a molecule's summation order independent of
    the other molecules of a batch plan; ``serial=True`` issues all class
    launches on the caller's stream (per-kernel profiling).
    """

    def __init__(self, shells, Z, centers, schwarz_tau=1e-10, fp32_tau=1e-4, uhf=False,
                 device=None, split=True, serial=False, static_workspace=False, mfma=False):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.DFTError("no HIP device visible: the integral engine runs on the GPU only")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.shells = shells
        self.nbf = int(shells[-1]["ao_stop"])
        self.uhf = bool(uhf)
        seg_l, seg_np, seg_ao, seg_xyz, exps, coefs = [], [], [], [], [], []
        self.segments = []  # per reference shell: list of segment ids
        for sh in shells:
            ids = []
            ncomp = sh["n_components"]
            for r in range(sh["n_contractions"]):
                ids.append(len(seg_l))
                seg_l.append(sh["angular_momentum"])
                seg_np.append(sh["n_exponents"])
                seg_ao.append(sh["ao_start"] + r * ncomp)
                seg_xyz.append(np.asarray(sh["atom_pos"], dtype=np.float64))
                exps.append(np.asarray(sh["exponents"], dtype=np.float64))
                coefs.append(np.asarray(sh["coefficients"], dtype=np.float64)[r])
            self.segments.append(ids)
        self.seg_l = np.asarray(seg_l, dtype=np.int32)
        self.nseg = len(seg_l)
        seg_np = np.asarray(seg_np, dtype=np.int32)
        seg_ao = np.asarray(seg_ao, dtype=np.int32)
        seg_xyz = np.ascontiguousarray(np.stack(seg_xyz), dtype=np.float64)
        exps = np.ascontiguousarray(np.concatenate(exps))
        coefs = np.ascontiguousarray(np.concatenate(coefs))
        Zc = np.ascontiguousarray(np.asarray(Z, dtype=np.float64))
        xyz = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
        self.Z, self.centers = Zc, xyz
        handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.dft_plan_create(
                self.nseg, _lib.ptr(self.seg_l, C.c_int32), _lib.ptr(seg_np, C.c_int32),
                _lib.ptr(seg_ao, C.c_int32), _lib.ptr(seg_xyz, C.c_double),
                _lib.ptr(exps, C.c_double), _lib.ptr(coefs, C.c_double), len(Zc),
                _lib.ptr(xyz, C.c_double), _lib.ptr(Zc, C.c_double), self.nbf,
                float(schwarz_tau), float(fp32_tau), _flags(uhf, split, serial, static_workspace, mfma),
                C.byref(handle)))
        self.handle = handle
        self.schwarz_tau, self.fp32_tau = schwarz_tau, fp32_tau

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.dft_plan_destroy(h)
            self.handle = None

    # ------------------------------------------------------------ stats
    def stats(self) -> dict:
        counts = np.zeros(_lib.STATS_COLS * 64, dtype=np.int64)
        nc = C.c_int()
        nun = C.c_int64()
        sb = C.c_int64()
        _lib.check(self.lib.dft_plan_stats(self.handle, _lib.ptr(counts, C.c_int64), 64, C.byref(nc),
                                           C.byref(nun), C.byref(sb)))
        c = counts[:_lib.STATS_COLS * nc.value].reshape(-1, _lib.STATS_COLS)
        # contracted Cartesian AO quartets, counted at shell level (prod n(l))
        ao = [np.prod([_ncart(l) for l in combo_moments(k)]) for k in range(len(c))]
        n_ao = int(sum(int(a) * int(row[0] + row[1]) for a, row in zip(ao, c)))
        return {"per_combo": c, "n_quartets_fp64": int(c[:, 0].sum()), "n_quartets_fp32": int(c[:, 1].sum()),
                "n_quartets": int(c[:, :2].sum()), "n_prim_quartets": int(c[:, 2:4].sum()),
                "n_prim_quartets_unscreened": int(c[:, 4:6].sum()),
                "n_quartets_fp32_level2": int(c[:, 6].sum()),
                "n_ao_quartets": n_ao,
                "n_unscreened": int(nun.value), "workspace_bytes": int(sb.value)}

    def schwarz(self) -> np.ndarray:
        Q = np.zeros(self.nseg * self.nseg)
        _lib.check(self.lib.dft_schwarz(self.handle, _lib.ptr(Q, C.c_double)))
        return Q.reshape(self.nseg, self.nseg)

    # -------------------------------------------------------------- J/K
    def _dev(self, M):
        if isinstance(M, torch.Tensor):
            t = M.to(device=self.device, dtype=torch.float64)
        else:
            t = torch.as_tensor(np.asarray(M, dtype=np.float64), device=self.device)
        t = t.contiguous()
        if t.shape != (self.nbf, self.nbf):
            raise ValueError(f"density must be {self.nbf}x{self.nbf}, got {tuple(t.shape)}")
        return t

    def _nelem(self) -> int:
        return self.nbf * self.nbf

    def check_buffers(self, P, Pa, Pb, J, Ka, Kb, mode="fp64"):
        """Raise ValueError unless the densities and outputs are what the C
        entry points read and write through raw pointers: contiguous float64
        tensors on the plan's device with nbf*nbf (packed batch: nelem)
        elements, Pb/Kb present exactly for a UHF plan. A wrong buffer would
        otherwise be read or written out of bounds on the device."""
        if mode not in _lib.MODES:
            raise ValueError(f"mode must be one of {sorted(_lib.MODES)}, got {mode!r}")
        if self.uhf != (Pb is not None) or self.uhf != (Kb is not None):
            raise ValueError("Pb and Kb must be given exactly for a UHF plan")
        n = self._nelem()
        for name, t in (("P", P), ("Pa", Pa), ("Pb", Pb), ("J", J), ("Ka", Ka), ("Kb", Kb)):
            if t is None:
                continue
            if not isinstance(t, torch.Tensor):
                raise ValueError(f"{name} must be a torch tensor")
            if t.device != self.device:
                raise ValueError(f"{name} is on {t.device}, the plan on {self.device}")
            if t.dtype != torch.float64:
                raise ValueError(f"{name} must be float64, got {t.dtype}")
            if not t.is_contiguous():
                raise ValueError(f"{name} must be contiguous")
            if t.numel() != n:
                raise ValueError(f"{name} has {t.numel()} elements, the plan needs {n}")

    def jk_into(self, P, Pa, Pb, J, Ka, Kb=None, mode="fp64", stream=None):
        """Enqueue J/K on ``stream`` into preallocated device tensors (no sync)."""
        self.check_buffers(P, Pa, Pb, J, Ka, Kb, mode)
        s = torch.cuda.current_stream(self.device) if stream is None else stream
        _lib.check(self.lib.dft_jk(
            self.handle, P.data_ptr(), Pa.data_ptr(), Pb.data_ptr() if Pb is not None else None,
            J.data_ptr(), Ka.data_ptr(), Kb.data_ptr() if Kb is not None else None,
            _lib.MODES[mode], C.c_void_p(s.cuda_stream)))

    def jk(self, P, Pa, Pb=None, mode="fp64"):
        """J[P], K[Pa] (and K[Pb] for a UHF plan) as device tensors."""
        P, Pa = self._dev(P), self._dev(Pa)
        if self.uhf:
            if Pb is None:
                raise ValueError("UHF plan needs Pb")
            Pb = self._dev(Pb)
        else:
            Pb = None
        J = torch.empty_like(P)
        Ka = torch.empty_like(P)
        Kb = torch.empty_like(P) if self.uhf else None
        self.jk_into(P, Pa, Pb, J, Ka, Kb, mode)
        return J, Ka, Kb

    # --------------------------------------------------- 1-e integrals
    def one_electron(self):
        n = self.nbf
        S = torch.empty((n, n), dtype=torch.float64, device=self.device)
        T = torch.empty_like(S)
        V = torch.empty_like(S)
        s = torch.cuda.current_stream(self.device)
        _lib.check(self.lib.dft_oneelec(self.handle, S.data_ptr(), T.data_ptr(), V.data_ptr(),
                                        C.c_void_p(s.cuda_stream)))
        return S, T, V

    def ao_values(self, grid, out=None):
        """Cartesian AO values at grid points: (nbf, npts) FP64 on the device
        (shelled_eval.py:98-113); ``grid`` (npts, 3) bohr, sent as FP32 like
        the reference's grid."""
        g = torch.as_tensor(grid, device=self.device).to(torch.float32).reshape(-1, 3).contiguous()
        npts = g.shape[0]
        if out is None:
            out = torch.empty((self.nbf, npts), dtype=torch.float64, device=self.device)
        if out.shape != (self.nbf, npts) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 (nbf, npts) tensor")
        s = torch.cuda.current_stream(self.device)
        _lib.check(self.lib.dft_ao_grid(self.handle, g.data_ptr(), npts, out.data_ptr(), npts,
                                        C.c_void_p(s.cuda_stream)))
        return out

    # ----------------------------------------------------- raw ERIs
    def eri_segment_blocks(self, quartets) -> np.ndarray:
        """Concatenated blocks for segment quartets (int array [nq, 4])."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        sizes = np.prod(np.array([[_ncart(int(self.seg_l[s])) for s in row] for row in q]), axis=1) \
            if len(q) else np.zeros(0, dtype=np.int64)
        out = np.zeros(int(sizes.sum()))
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.dft_eri_quartets(self.handle, _lib.ptr(q, C.c_int32), len(q),
                                             _lib.ptr(out, C.c_double)))
        return out

    def eri_blocks(self, quartets):
        """Blocks [n_ao_a, n_ao_b, n_ao_c, n_ao_d] for reference shell quartets."""
        seg_q, plan = [], []
        for (a, b, c, d) in quartets:
            sa, sb, sc, sd = (self.segments[k] for k in (a, b, c, d))
            plan.append((a, b, c, d))
            for i in sa:
                for j in sb:
                    for k in sc:
                        for m in sd:
                            seg_q.append((i, j, k, m))
        flat = self.eri_segment_blocks(seg_q)
        out, pos = [], 0
        for (a, b, c, d) in plan:
            sh = [self.shells[k] for k in (a, b, c, d)]
            nc = [s["n_components"] for s in sh]
            blk = np.zeros([s["n_ao"] for s in sh])
            for ri in range(sh[0]["n_contractions"]):
                for rj in range(sh[1]["n_contractions"]):
                    for rk in range(sh[2]["n_contractions"]):
                        for rm in range(sh[3]["n_contractions"]):
                            size = nc[0] * nc[1] * nc[2] * nc[3]
                            sub = flat[pos:pos + size].reshape(nc)
                            pos += size
                            blk[ri * nc[0]:(ri + 1) * nc[0], rj * nc[1]:(rj + 1) * nc[1],
                                rk * nc[2]:(rk + 1) * nc[2], rm * nc[3]:(rm + 1) * nc[3]] = sub
            out.append(blk)
        return out

    def eri_tensor(self) -> np.ndarray:
        """Full (nbf^4) ERI tensor, from the unique shell quartets."""
        n = self.nbf
        nsh = len(self.shells)
        pairs = [(i, j) for i in range(nsh) for j in range(i + 1)]
        quart = [(a, b, c, d) for ab, (a, b) in enumerate(pairs) for (c, d) in pairs[:ab + 1]]
        blocks = self.eri_blocks(quart)
        T = np.zeros((n, n, n, n))
        for (a, b, c, d), blk in zip(quart, blocks):
            s = [slice(self.shells[k]["ao_start"], self.shells[k]["ao_stop"]) for k in (a, b, c, d)]
            for perm in ((0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
                         (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)):
                T[s[perm[0]], s[perm[1]], s[perm[2]], s[perm[3]]] = blk.transpose(perm)
        return T


def boys(T, mmax: int, fp32: bool = False) -> torch.Tensor:
    """F_m(T), m = 0..mmax, from the device evaluator (test hook)."""
    lib = _lib.load()
    T = torch.as_tensor(T, dtype=torch.float64).to("cuda").contiguous()
    out = torch.empty((T.numel(), mmax + 1), dtype=torch.float64, device=T.device)
    s = torch.cuda.current_stream(T.device)
    _lib.check(lib.dft_boys(T.data_ptr(), T.numel(), mmax, out.data_ptr(), int(fp32),
                            C.c_void_p(s.cuda_stream)))
    return out


def profile_start(plan: Plan, nsteps: int) -> None:
    """Record per-launch HIP-event timings for the next ``nsteps`` J/K builds."""
    _lib.check(plan.lib.dft_plan_profile(plan.handle, int(nsteps)))


def profile_read(plan: Plan):
    """(ms summed over recorded steps per slot, number of steps)."""
    ns, st = C.c_int(), C.c_int()
    _lib.check(plan.lib.dft_plan_profile_read(plan.handle, None, 0, C.byref(ns), C.byref(st)))
    ms = np.zeros(ns.value)
    _lib.check(plan.lib.dft_plan_profile_read(plan.handle, _lib.ptr(ms, C.c_double), ns.value, C.byref(ns),
                                              C.byref(st)))
    return ms, st.value


def flatten_shells(shells):
    """Segmented-shell arrays of one molecule: (l, nprim, ao, xyz, exps, coefs, segments)."""
    seg_l, seg_np, seg_ao, seg_xyz, exps, coefs, segments = [], [], [], [], [], [], []
    for sh in shells:
        ids = []
        for r in range(sh["n_contractions"]):
            ids.append(len(seg_l))
            seg_l.append(sh["angular_momentum"])
            seg_np.append(sh["n_exponents"])
            seg_ao.append(sh["ao_start"] + r * sh["n_components"])
            seg_xyz.append(np.asarray(sh["atom_pos"], dtype=np.float64))
            exps.append(np.asarray(sh["exponents"], dtype=np.float64))
            coefs.append(np.asarray(sh["coefficients"], dtype=np.float64)[r])
        segments.append(ids)
    return (np.asarray(seg_l, np.int32), np.asarray(seg_np, np.int32), np.asarray(seg_ao, np.int32),
            np.stack(seg_xyz), np.concatenate(exps), np.concatenate(coefs), segments)


class Graph:
    """An instantiated hipGraph of one J/K build (dft_jk_graph_create)."""

    def __init__(self, plan, P, Pa, Pb, J, Ka, Kb, mode):
        self.plan = plan
        plan.check_buffers(P, Pa, Pb, J, Ka, Kb, mode)
        self.keep = (P, Pa, Pb, J, Ka, Kb)  # the graph binds these buffers
        h = C.c_void_p()
        _lib.check(plan.lib.dft_jk_graph_create(plan.handle, P.data_ptr(), Pa.data_ptr(),
                                                Pb.data_ptr() if Pb is not None else None, J.data_ptr(),
                                                Ka.data_ptr(), Kb.data_ptr() if Kb is not None else None,
                                                _lib.MODES[mode], C.byref(h)))
        self.handle = h

    def launch(self, stream=None):
        s = torch.cuda.current_stream(self.plan.device) if stream is None else stream
        _lib.check(self.plan.lib.dft_graph_launch(self.handle, C.c_void_p(s.cuda_stream)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.plan.lib.dft_graph_destroy(h)
            self.handle = None


def tune(plan, P, Pa, Pb=None, reps=3, log=None):
    """Autotune the plan's launch shapes on these densities, one class
    combination at a time; returns the choices as an int array [ncombo, 6]
    (FP64 {m, block, cap}, FP32 {f, block, cap}); m = 1 where the heavy
    d combination's FP64 part runs on the matrix cores (kernels_mfma.hip),
    f = 1 where the combination stays FP64 in mixed mode."""
    plan.check_buffers(P, Pa, Pb, P, Pa, Pb)  # densities only (the tuner writes no J/K)
    per = plan.stats()["per_combo"]
    ncombo = len(per)
    rep = (C.c_int * (6 * ncombo))()
    s = torch.cuda.current_stream(plan.device)
    for c in range(ncombo):
        if per[c][0] + per[c][1] == 0:
            continue
        _lib.check(plan.lib.dft_plan_tune(plan.handle, P.data_ptr(), Pa.data_ptr(),
                                          Pb.data_ptr() if Pb is not None else None, int(reps), c, rep,
                                          6 * ncombo, C.c_void_p(s.cuda_stream)))
        if log:
            log(f"tuned combination {c}: {list(rep[6 * c:6 * c + 6])}")
    # serial or concurrent class launches, from timed whole builds
    _lib.check(plan.lib.dft_plan_tune(plan.handle, P.data_ptr(), Pa.data_ptr(),
                                      Pb.data_ptr() if Pb is not None else None, int(reps), -2, rep,
                                      6 * ncombo, C.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize(plan.device)
    return np.array(rep[:], dtype=np.int64).reshape(ncombo, 6)


def set_shapes(plan, shapes):
    arr = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).ravel())
    _lib.check(plan.lib.dft_plan_set_shapes(plan.handle, _lib.ptr(arr, C.c_int), arr.size))


class BatchPlan(Plan):
    """Plan over many independent molecules (dft_plan_create_batch).

    Matrices are packed: molecule m's nbf_m x nbf_m block starts at element
    ``offsets[m]`` of a flat device vector; ``pack``/``unpack`` convert.
    """

    def __init__(self, molecules, schwarz_tau=1e-10, fp32_tau=1e-4, uhf=False, device=None, split=True,
                 serial=False):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.DFTError("no HIP device visible: the integral engine runs on the GPU only")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.uhf = bool(uhf)
        self.molecules = molecules
        parts = [flatten_shells(sh) for sh, _, _ in molecules]
        self.nbfs = [int(sh[-1]["ao_stop"]) for sh, _, _ in molecules]
        self.offsets = np.concatenate([[0], np.cumsum([n * n for n in self.nbfs])]).astype(np.int64)
        self.nelem = int(self.offsets[-1])
        mol_nsh = np.array([len(p[0]) for p in parts], np.int32)
        mol_nat = np.array([len(Z) for _, Z, _ in molecules], np.int32)
        mol_nbf = np.array(self.nbfs, np.int32)
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([p[k] for p in parts]), dtype=dt)  # noqa: E731
        l, npr, ao = cat(0, np.int32), cat(1, np.int32), cat(2, np.int32)
        xyz = np.ascontiguousarray(np.concatenate([p[3] for p in parts]).reshape(-1), dtype=np.float64)
        ex, co = cat(4, np.float64), cat(5, np.float64)
        Zc = np.ascontiguousarray(np.concatenate([np.asarray(Z, np.float64) for _, Z, _ in molecules]))
        axyz = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).reshape(-1, 3)
                                                    for _, _, c in molecules]).reshape(-1))
        self.Z, self.centers = Zc, axyz.reshape(-1, 3)
        handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.dft_plan_create_batch(
                len(molecules), _lib.ptr(mol_nsh, C.c_int32), _lib.ptr(mol_nat, C.c_int32),
                _lib.ptr(mol_nbf, C.c_int32), _lib.ptr(l, C.c_int32), _lib.ptr(npr, C.c_int32),
                _lib.ptr(ao, C.c_int32), _lib.ptr(xyz, C.c_double), _lib.ptr(ex, C.c_double),
                _lib.ptr(co, C.c_double), _lib.ptr(axyz, C.c_double), _lib.ptr(Zc, C.c_double),
                float(schwarz_tau), float(fp32_tau), _flags(uhf, split, serial), C.byref(handle)))
        self.handle = handle
        self.nbf = None
        self.schwarz_tau, self.fp32_tau = schwarz_tau, fp32_tau

    def _nelem(self) -> int:
        return self.nelem

    def empty(self):
        return torch.empty(self.nelem, dtype=torch.float64, device=self.device)

    def pack(self, mats):
        return torch.cat([torch.as_tensor(m, dtype=torch.float64, device=self.device).reshape(-1)
                          for m in mats]).contiguous()

    def unpack(self, flat):
        return [flat[self.offsets[m]:self.offsets[m + 1]].reshape(n, n) for m, n in enumerate(self.nbfs)]

    def one_electron(self):
        S, T, V = self.empty(), self.empty(), self.empty()
        s = torch.cuda.current_stream(self.device)
        _lib.check(self.lib.dft_oneelec(self.handle, S.data_ptr(), T.data_ptr(), V.data_ptr(),
                                        C.c_void_p(s.cuda_stream)))
        return S, T, V

    def jk(self, P, Pa, Pb=None, mode="fp64"):
        J, Ka = self.empty(), self.empty()
        Kb = self.empty() if self.uhf else None
        self.jk_into(P, Pa, Pb if self.uhf else None, J, Ka, Kb, mode)
        return J, Ka, Kb

    def schwarz(self):
        raise NotImplementedError("Schwarz matrices are per molecule: build a single-molecule Plan")

    def eri_tensor(self):
        raise NotImplementedError("full ERI tensors are per molecule: build a single-molecule Plan")
