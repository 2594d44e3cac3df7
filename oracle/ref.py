"""ctypes front end of oracle/md_ref.c (the C restatement of the reference).

TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's cpu_baseline leg.
Shells with several contraction rows are split into one segment per row,
which preserves the reference's AO order (contraction-major).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmd_ref.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        lib.ref_eri_blocks.argtypes = [C.c_int, ip, ip, ip, dp, dp, dp, C.c_int, ip, C.c_longlong, dp]
        lib.ref_build_jk.argtypes = [C.c_int, ip, ip, ip, dp, dp, dp, C.c_int, ip, C.c_longlong, dp, dp, dp,
                                     dp, dp, dp, C.c_int]
        lib.ref_boys.argtypes = [C.c_int, C.c_double, dp]
        lib.ref_boys.restype = None
        lib.ref_max_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Segments:
    """Flattened segmented shells of a reference shell list."""

    def __init__(self, shells):
        l, npr, ao, xyz, ex, co = [], [], [], [], [], []
        self.of_shell = []
        for sh in shells:
            ids = []
            for r in range(sh["n_contractions"]):
                ids.append(len(l))
                l.append(sh["angular_momentum"])
                npr.append(sh["n_exponents"])
                ao.append(sh["ao_start"] + r * sh["n_components"])
                xyz.append(np.asarray(sh["atom_pos"], float))
                ex.append(np.asarray(sh["exponents"], float))
                co.append(np.asarray(sh["coefficients"], float)[r])
            self.of_shell.append(ids)
        self.l = np.asarray(l, np.int32)
        self.nprim = np.asarray(npr, np.int32)
        self.ao = np.asarray(ao, np.int32)
        self.xyz = np.ascontiguousarray(np.stack(xyz))
        self.exps = np.ascontiguousarray(np.concatenate(ex))
        self.coefs = np.ascontiguousarray(np.concatenate(co))
        self.nbf = int(shells[-1]["ao_stop"])
        self.n = len(l)

    def args(self):
        return (self.n, _p(self.l, C.c_int), _p(self.nprim, C.c_int), _p(self.ao, C.c_int),
                _p(self.xyz, C.c_double), _p(self.exps, C.c_double), _p(self.coefs, C.c_double), self.nbf)


def unique_quartets(nseg):
    """(a>=b, c>=d, ab>=cd) in calc_eri order (shelled.py:634-640)."""
    pairs = [(i, j) for i in range(nseg) for j in range(i + 1)]
    return np.array([(a, b, c, d) for ab, (a, b) in enumerate(pairs) for (c, d) in pairs[:ab + 1]],
                    dtype=np.int32).reshape(-1, 4)


def _check_order(q):
    """md_ref.c builds pair data for i >= j only: every quartet must list its
    pairs as (a >= b, c >= d)."""
    if len(q) and ((q[:, 0] < q[:, 1]).any() or (q[:, 2] < q[:, 3]).any()):
        raise ValueError("quartets must be ordered a >= b and c >= d")


def build_jk(seg: Segments, quartets, P, Pa, Pb=None, nthreads=0):
    lib = load()
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    _check_order(q)
    n = seg.nbf
    P, Pa = (np.ascontiguousarray(x, dtype=np.float64) for x in (P, Pa))
    J, Ka = np.zeros((n, n)), np.zeros((n, n))
    Kb = np.zeros((n, n)) if Pb is not None else None
    Pbp = _p(np.ascontiguousarray(Pb, dtype=np.float64), C.c_double) if Pb is not None else None
    lib.ref_build_jk(*seg.args(), _p(q, C.c_int), len(q), _p(P, C.c_double), _p(Pa, C.c_double), Pbp,
                     _p(J, C.c_double), _p(Ka, C.c_double), _p(Kb, C.c_double) if Kb is not None else None,
                     int(nthreads))
    return J, Ka, Kb


def eri_blocks(seg: Segments, quartets):
    lib = load()
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    _check_order(q)
    nc = (seg.l + 1) * (seg.l + 2) // 2
    sizes = np.prod(nc[q], axis=1) if len(q) else np.zeros(0, int)
    out = np.zeros(int(sizes.sum()))
    lib.ref_eri_blocks(*seg.args(), _p(q, C.c_int), len(q), _p(out, C.c_double))
    return out, sizes


def boys(max_m, T):
    lib = load()
    F = np.zeros(max_m)
    lib.ref_boys(max_m, float(T), _p(F, C.c_double))
    return F
