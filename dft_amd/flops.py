"""Algorithmic flop model of the McMurchie-Davidson Fock build.

Per primitive shell quartet of class (la lb|lc ld), L = la+lb+lc+ld,
n(l) = (l+1)(l+2)/2, H(l) = C(l+3, 3) and
NZ(la, lb) = sum over Cartesian a, b of prod_d (a_d + b_d + 1):

  F_prim = 20 (geometry, prefactor) + (20 + 3L) (Boys) + 2 C(L+4, 4) (R_tuv)
         + 2 H(la+lb) NZ(lc, ld) (ket E.R) + 2 NZ(la, lb) n(lc) n(ld) (bra E.(ER))
         + 2 n(la) n(lb) n(lc) n(ld) (contraction)
  F_digest = 12 n(la) n(lb) n(lc) n(ld) per contracted quartet (one K density;
             + 8 prod n per extra exchange density)

This is the model the roofline fractions are quoted against; it counts the
work of MD with shell-level shared R and dense per-component E, not the
instructions any particular kernel issues.
"""
from math import comb

from .basis import combinations


def ncart(l):
    return (l + 1) * (l + 2) // 2


def nz(la, lb):
    return sum((a[0] + b[0] + 1) * (a[1] + b[1] + 1) * (a[2] + b[2] + 1)
               for a in combinations(la) for b in combinations(lb))


def f_prim(la, lb, lc, ld):
    L = la + lb + lc + ld
    return (20 + (20 + 3 * L) + 2 * comb(L + 4, 4) + 2 * comb(la + lb + 3, 3) * nz(lc, ld)
            + 2 * nz(la, lb) * ncart(lc) * ncart(ld)
            + 2 * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld))


def f_digest(la, lb, lc, ld, nk=1):
    n4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    return 12 * n4 + 8 * n4 * (nk - 1)


def pair_class_moments(pc):
    la = 0
    while (la + 1) * (la + 2) // 2 <= pc:
        la += 1
    return la, pc - la * (la + 1) // 2


def combo_moments(combo):
    pcb = 0
    while (pcb + 1) * (pcb + 2) // 2 <= combo:
        pcb += 1
    pck = combo - pcb * (pcb + 1) // 2
    return pair_class_moments(pcb) + pair_class_moments(pck)


def combo_flops(combo, n_quartets, n_prim_quartets, nk=1):
    """Model flops of one class-combination launch."""
    la, lb, lc, ld = combo_moments(combo)
    return n_prim_quartets * f_prim(la, lb, lc, ld) + n_quartets * f_digest(la, lb, lc, ld, nk)


def build_flops(per_combo, nk=1, model=True, mode="mixed"):
    """(FP64-part flops, FP32-eligible-part flops) of one Fock build from
    Plan.stats()['per_combo'].

    ``model=True`` counts K_a K_b K_c K_d primitive quartets per surviving
    contracted quartet (the algorithmic model); ``model=False`` counts the
    primitive quartets left by primitive-pair screening (the kernels' own
    run-time primitive-quartet cut removes more), with the FP32 records in
    ``mode="mixed"`` and the FP64 ones for the FP32-eligible part in
    ``mode="fp64"``.
    """
    f64 = f32 = 0
    for c, row in enumerate(per_combo):
        if model:
            p64, p32 = row[4], row[5]
        else:
            p64, p32 = row[2], (row[3] if mode == "mixed" else row[7])
        f64 += combo_flops(c, int(row[0]), int(p64), nk)
        f32 += combo_flops(c, int(row[1]), int(p32), nk)
    return f64, f32
