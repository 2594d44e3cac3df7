"""Basis-set data and shell construction.

Host-side data contract of the reference's shell list:

* ``combinations(l)`` gives the Cartesian component order
  (reference ``shelled.py:15-23``): p = x, y, z; d = xx, xy, yy, xz, yz, zz.
* ``build_shells`` turns the JSON basis format ``data[Z] = [[l-list],
  [exponents], [[coefficient rows]]]`` (written by ``compile.py:8-23``) into
  the list of shell dicts of ``shelled.py:49-77,934-952``: SP shells are split
  into an s and a p shell that share exponents, AO offsets are laid out in
  shell order, and within a shell AO = contraction * ncomp + component.
* ``normal`` is the primitive normalisation of ``shelled.py:88-101``; only
  primitives are normalised, contracted functions are not renormalised.

Everything here is plain NumPy: it runs once per molecule and never on the
per-iteration path.
"""
from __future__ import annotations

import json
import os
from math import factorial

import numpy as np

_DATA = os.path.join(os.path.dirname(__file__), "data")


def combinations(added: int) -> list[list[int]]:
    """Cartesian exponent triples [lx, ly, lz] with lx+ly+lz == added.

    Loop nest z-outer, y-middle, x-inner, as ``shelled.py:15-23``.
    """
    total = []
    for n in range(added + 1):
        for m in range(added + 1):
            for l in range(added + 1):
                if l + m + n == added:
                    total.append([l, m, n])
    return total


def double_factorial(x):
    """(x)!! for integer x >= -1, with (-1)!! = 1 (``shelled.py:88-92``)."""
    x = np.asarray(x, dtype=np.int64)
    out = np.ones(x.shape, dtype=np.float64)
    for idx in np.ndindex(x.shape):
        v = int(x[idx])
        r = 1
        while v > 1:
            r *= v
            v -= 2
        out[idx] = float(r)
    return out


def normal(alpha, pow):
    """Primitive Cartesian normalisation (``shelled.py:95-101``).

    N = (2a/pi)^(3/4) * sqrt((4a)^(lx+ly+lz) / ((2lx-1)!! (2ly-1)!! (2lz-1)!!))
    ``alpha`` has shape (k,), ``pow`` shape (k, 3).
    """
    alpha = np.asarray(alpha, dtype=np.float64)
    pow = np.asarray(pow, dtype=np.int64)
    term = np.power(4.0 * alpha, pow.sum(axis=1).astype(np.float64))
    for d in range(3):
        term = term / double_factorial(2 * pow[:, d] - 1)
    return np.sqrt(term) * np.power(2.0 * alpha / np.pi, 0.75)


def load_basis(name: str = "sto-3g-hco") -> dict:
    """Load a basis in the reference JSON format from the package data dir."""
    path = name if os.path.isfile(name) else os.path.join(_DATA, name + ".json")
    with open(path) as fh:
        return json.load(fh)


def build_shells(basis: dict, atoms, centers) -> list[dict]:
    """Shell dicts with the keys of ``shelled.py:58-64,938-952``.

    ``atoms`` are atomic numbers (ints or digit strings), ``centers`` an
    (N, 3) array in bohr (the reference applies no unit conversion).
    """
    return build_shells_from_entries([basis[str(int(a))] for a in atoms], centers)


def build_shells_from_entries(entries_per_atom, centers) -> list[dict]:
    """Like ``build_shells`` but with one basis-entry list per atom.

    Used for synthetic bases in which atoms of one element carry different
    shells.
    """
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    shells = []
    for i, entries in enumerate(entries_per_atom):
        for l, exponents, coefficients in entries:
            if len(l) == 1:
                rows = [(l[0], np.asarray(coefficients, dtype=np.float64))]
            elif len(l) == len(coefficients):
                rows = [(m, np.asarray([row], dtype=np.float64))
                        for m, row in zip(l, coefficients)]
            else:
                raise ValueError("basis entry has %d momenta but %d coefficient rows"
                                 % (len(l), len(coefficients)))
            for momentum, coef in rows:
                shells.append({
                    "atom_idx": i,
                    "atom_pos": centers[i].copy(),
                    "angular_momentum": int(momentum),
                    "exponents": np.asarray(exponents, dtype=np.float64),
                    "coefficients": np.atleast_2d(coef),
                })
    total_ao = 0
    sphere_total_ao = 0
    for shell in shells:
        lq = shell["angular_momentum"]
        shell["components"] = np.asarray(combinations(lq), dtype=np.int32)
        shell["n_components"] = len(shell["components"])
        shell["n_contractions"] = shell["coefficients"].shape[0]
        shell["n_ao"] = shell["n_components"] * shell["n_contractions"]
        shell["sphere_n_ao"] = shell["n_contractions"] * (2 * lq + 1)
        shell["n_exponents"] = shell["exponents"].size
        shell["ao_start"] = total_ao
        shell["sphere_ao_start"] = sphere_total_ao
        shell["ao_stop"] = total_ao + shell["n_ao"]
        shell["sphere_ao_stop"] = sphere_total_ao + shell["sphere_n_ao"]
        exp = np.tile(shell["exponents"], shell["n_components"])
        pw = np.repeat(shell["components"], shell["n_exponents"], axis=0).reshape(-1, 3)
        shell["normalization"] = normal(exp, pw)
        total_ao = shell["ao_stop"]
        sphere_total_ao = shell["sphere_ao_stop"]
    return shells


def n_ao(shells) -> int:
    return shells[-1]["ao_stop"] if shells else 0


def binom(n: int, k: int) -> int:
    return factorial(n) // (factorial(k) * factorial(n - k))
