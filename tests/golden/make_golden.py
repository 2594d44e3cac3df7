"""Write the golden fixtures under tests/golden/.

* boys_mpmath.npz  F_m(T), m = 0..16, at edge and interior T values, from
  mpmath at 40 digits: F_m(T) = gamma_lower(m + 1/2, T) / (2 T^(m+1/2)),
  F_m(0) = 1/(2m+1). Independent of every implementation in this repo.
* water_sto3g.npz  config-0 water (O at the origin, H at (+-0.757, 0.586, 0) A,
  STO-3G): full 7^4 ERI tensor, S, T, V, E_NN and the converged plain-SCF
  total energy from the NumPy oracle (oracle/md_oracle.py), which is itself
  pinned by the mpmath and textbook checks in tests/test_oracle.py.
* dimer_sto3g.npz  (H2O)_2 from the first two sites of the config-1 lattice
  (seed 0): geometry, ERI tensor and plain-SCF energy from the oracle.

Run from the repo root: python tests/golden/make_golden.py
"""
import os
import sys

import mpmath as mp
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dft_amd.basis import build_shells, load_basis  # noqa: E402
from dft_amd.geometry import water, water_cluster  # noqa: E402
import oracle.md_oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def boys_mp(m, T):
    mp.mp.dps = 40
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    T = mp.mpf(T)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def boys_fixture():
    Ts = [0.0, 1e-16, 1e-14, 1e-10, 1e-6, 1e-3, 0.04999, 0.05, 0.05001, 0.1, 0.25, 0.5, 1.0, 1.23456, 2.0, 3.3,
          5.0, 5.95, 5.99999, 6.0, 6.00001, 7.5, 10.0, 12.3, 17.77, 25.0, 33.3, 39.95, 39.99999, 40.0, 40.00001,
          45.0, 60.0, 100.0, 250.0, 1000.0, 1e4]
    rng = np.random.default_rng(7)
    Ts += list(rng.uniform(0, 45, 40))
    Ts = np.array(Ts)
    F = np.array([[float(boys_mp(m, T)) for m in range(17)] for T in Ts])
    np.savez(os.path.join(OUT, "boys_mpmath.npz"), T=Ts, F=F)


def molecule_fixture(name, Z, xyz):
    shells = build_shells(load_basis(), Z, xyz)
    blocks = O.eri_blocks(shells)
    ERI = O.eri_tensor(shells, blocks)
    S, T, V = O.one_electron(shells, Z, xyz)
    E_NN = O.nuclear_repulsion(Z, xyz)
    ne = int(sum(Z)) // 2
    res = O.scf_plain(T + V, S, lambda P, Pa, Pb: O.jk_dense(ERI, P, Pa, Pb), E_NN, ne, ne)
    np.savez(os.path.join(OUT, name), Z=np.array(Z), xyz=xyz, ERI=ERI, S=S, T=T, V=V, E_NN=E_NN,
             E=res["E"], P_a=res["P_a"], count=res["count"])
    print(name, res["E"], res["count"])


if __name__ == "__main__" and len(sys.argv) == 1:
    boys_fixture()
    molecule_fixture("water_sto3g.npz", *water())
    Z, xyz = water_cluster()
    molecule_fixture("dimer_sto3g.npz", Z[:6], xyz[:6])


def methane_spd_fixture():
    """CH4 with the synthetic s/p/d basis rules of the C12H26 workload (seed 1).

    Reference values from the C restatement (oracle/md_ref.c): full ERI tensor
    (Cartesian), and J/K for two random symmetric densities.
    """
    import oracle.ref as R
    from dft_amd.basis import build_shells_from_entries
    from dft_amd.geometry import methane, synthetic_spd_basis
    Z, xyz = methane()
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    seg = R.Segments(shells)
    q = R.unique_quartets(seg.n)
    flat, sizes = R.eri_blocks(seg, q)
    n = seg.nbf
    ERI = np.zeros((n, n, n, n))
    pos = 0
    for (a, b, c, d), sz in zip(q, sizes):
        nc = [(l + 1) * (l + 2) // 2 for l in seg.l[[a, b, c, d]]]
        blk = flat[pos:pos + sz].reshape(nc)
        pos += sz
        s = [slice(seg.ao[k], seg.ao[k] + m) for k, m in zip((a, b, c, d), nc)]
        for perm in ((0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
                     (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)):
            ERI[s[perm[0]], s[perm[1]], s[perm[2]], s[perm[3]]] = blk.transpose(perm)
    S, T, V = O.one_electron(shells, Z, xyz)
    np.savez_compressed(os.path.join(OUT, "methane_spd.npz"), Z=np.array(Z), xyz=xyz, ERI=ERI, S=S, T=T, V=V)
    print("methane_spd", n, "bf", len(q), "quartets")


if __name__ == "__main__" and "methane" in sys.argv:
    methane_spd_fixture()
