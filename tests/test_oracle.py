"""Pins of the CPU oracle against mathematics and textbook values.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index) fails: mpmath Boys values, closed forms of E and
of the (ss|ss) integral, finite-difference identities for R and for the
centre derivative that generates p functions, the Szabo-Ostlund H2/STO-3G
integrals and energy, ERI permutation symmetry and the Schwarz inequality,
and SCF invariants.
"""
import math

import numpy as np
import pytest

import oracle.md_oracle as O
from conftest import golden


def test_boys_oracle_vs_mpmath():
    g = golden("boys_mpmath.npz")
    T, F = g["T"], g["F"]
    for mmax in (1, 5, 13):
        got = O.boys(mmax, T)
        ref = F[:, :mmax]
        rel = np.abs(got - ref) / np.abs(ref)
        # the reference's 40-term alternating Taylor series loses a few digits near T = 6
        assert rel.max() < 1e-11, rel.max()


def test_boys_c_oracle_matches_numpy_oracle():
    import oracle.ref as R
    for T in (0.0, 1e-15, 0.3, 5.999, 6.0, 17.0, 80.0):
        assert np.allclose(R.boys(9, T), O.boys(9, T), rtol=1e-15, atol=0)


def test_calc_E_closed_forms():
    a, b, A, B = 1.3, 0.7, 0.2, -0.5
    p, mu, Q = a + b, a * b / (a + b), A - B
    K = math.exp(-mu * Q * Q)
    E00 = O.calc_E(a, b, A, B, 0, 0)
    E10 = O.calc_E(a, b, A, B, 1, 0)
    E01 = O.calc_E(a, b, A, B, 0, 1)
    assert np.isclose(E00[0], K, rtol=1e-15)
    assert np.isclose(E10[0], -b * Q / p * K, rtol=1e-14)  # X_PA K
    assert np.isclose(E10[1], K / (2 * p), rtol=1e-14)
    assert np.isclose(E01[0], a * Q / p * K, rtol=1e-14)    # X_PB K
    assert np.allclose(E00[1:], 0) and np.allclose(E10[2:], 0)


@pytest.mark.parametrize("la,lb", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 1), (2, 2), (3, 2)])
def test_overlap_from_E_equals_binomial_form(la, lb):
    # S_ij = E_0^{ij} sqrt(pi/p): independent of the binomial closed form
    a, b, A, B = 0.9, 2.1, 0.3, -0.4
    E = O.calc_E(a, b, A, B, la, lb)
    S = O.overlap_1d(np.array([a]), np.array([b]), A, B, np.array([la]), np.array([lb]))[0, 0]
    assert np.isclose(E[0] * math.sqrt(math.pi / (a + b)), S, rtol=1e-12, atol=1e-16)


def test_calc_R_base_and_finite_differences():
    X, Y, Z, al = 0.3, -0.7, 0.45, 0.8
    R = O.calc_R(3, 3, 3, 7, X, Y, Z, al)
    F0 = O.boys(1, al * (X * X + Y * Y + Z * Z))[0]
    assert np.isclose(R[0, 0, 0, 0], F0, rtol=1e-15)
    h = 1e-5

    def r(tuv, dx=0, dy=0, dz=0):
        return O.calc_R(3, 3, 3, 7, X + dx, Y + dy, Z + dz, al)[tuv]

    # R_{t+1,u,v} = d/dX R_tuv, likewise for Y and Z
    for (t, u, v) in [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 0)]:
        fd_x = (r((t, u, v, 0), dx=h) - r((t, u, v, 0), dx=-h)) / (2 * h)
        fd_y = (r((t, u, v, 0), dy=h) - r((t, u, v, 0), dy=-h)) / (2 * h)
        fd_z = (r((t, u, v, 0), dz=h) - r((t, u, v, 0), dz=-h)) / (2 * h)
        assert np.isclose(R[t + 1, u, v, 0], fd_x, rtol=1e-7, atol=1e-10)
        assert np.isclose(R[t, u + 1, v, 0], fd_y, rtol=1e-7, atol=1e-10)
        assert np.isclose(R[t, u, v + 1, 0], fd_z, rtol=1e-7, atol=1e-10)


def _prim_shell(l, a, center):
    from dft_amd.basis import build_shells_from_entries
    return build_shells_from_entries([[[[l], [a], [[1.0]]]]], [center])[0]


def test_ss_ss_closed_form():
    A, B, Cc, D = [0, 0, 0], [0.5, 0.2, -0.1], [1.0, -0.4, 0.3], [0.2, 0.9, 0.6]
    a, b, c, d = 0.8, 1.7, 0.5, 2.3
    sh = [_prim_shell(0, e, r) for e, r in ((a, A), (b, B), (c, Cc), (d, D))]
    val = O.eri_prim(O.pair_data(sh[0], sh[1]), O.pair_data(sh[2], sh[3]))[0, 0]
    p, q = a + b, c + d
    P = (a * np.array(A) + b * np.array(B)) / p
    Q = (c * np.array(Cc) + d * np.array(D)) / q
    Kab = math.exp(-a * b / p * np.sum((np.array(A) - np.array(B)) ** 2))
    Kcd = math.exp(-c * d / q * np.sum((np.array(Cc) - np.array(D)) ** 2))
    rho = p * q / (p + q)
    F0 = float(mpmath_boys0(rho * np.sum((P - Q) ** 2)))
    ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0
    assert np.isclose(val, ref, rtol=1e-13)


def mpmath_boys0(T):
    import mpmath as mp
    mp.mp.dps = 30
    return mp.sqrt(mp.pi) / 2 * mp.erf(mp.sqrt(T)) / mp.sqrt(T)


def test_centre_derivative_generates_p():
    # d/dA_x (s_A s_B|s_C s_D) = 2a (px_A s_B|s_C s_D) for unnormalised primitives
    a, b, c, d = 0.9, 1.4, 0.6, 1.1
    A, B, Cc, D = np.array([0.1, 0.2, -0.3]), np.array([0.7, -0.2, 0.1]), np.array([-0.5, 0.4, 0.8]), \
        np.array([0.3, 0.3, 0.3])
    h = 1e-5

    def ssss(Ax):
        s = [_prim_shell(0, e, r) for e, r in ((a, Ax), (b, B), (c, Cc), (d, D))]
        return O.eri_prim(O.pair_data(s[0], s[1]), O.pair_data(s[2], s[3]))[0, 0]
    fd = [(ssss(A + h * e) - ssss(A - h * e)) / (2 * h) for e in np.eye(3)]
    pa = _prim_shell(1, a, A)
    sh = [_prim_shell(0, e, r) for e, r in ((b, B), (c, Cc), (d, D))]
    p_ss = O.eri_prim(O.pair_data(pa, sh[0]), O.pair_data(sh[1], sh[2]))[:, 0]
    assert np.allclose(fd, 2 * a * p_ss, rtol=1e-7, atol=1e-12)


def test_szabo_ostlund_h2(sto3g):
    from dft_amd.basis import build_shells
    xyz = np.array([[0, 0, 0], [1.4, 0, 0.0]])
    sh = build_shells(sto3g, [1, 1], xyz)
    T = O.eri_tensor(sh)
    S, Tk, V = O.one_electron(sh, [1, 1], xyz)
    assert abs(S[0, 1] - 0.6593) < 1e-4
    assert abs(T[0, 0, 0, 0] - 0.7746) < 1e-4
    assert abs(T[0, 0, 1, 1] - 0.5697) < 1e-4
    assert abs(T[0, 1, 0, 1] - 0.2970) < 1e-4
    assert abs(T[0, 0, 0, 1] - 0.4441) < 1e-4
    assert abs(Tk[0, 0] - 0.7600) < 1e-4 and abs(Tk[0, 1] - 0.2365) < 1e-4
    res = O.scf_plain(Tk + V, S, lambda P, a, b: O.jk_dense(T, P, a, b), O.nuclear_repulsion([1, 1], xyz), 1, 1)
    assert abs(res["E"] - (-1.1167)) < 1e-4


def test_water_tensor_symmetry_schwarz_energy(water_shells):
    Z, xyz, sh = water_shells
    g = golden("water_sto3g.npz")
    T = O.eri_tensor(sh)
    assert np.abs(T - g["ERI"]).max() < 1e-13
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        assert np.allclose(T, T.transpose(perm), atol=1e-14)
    d = np.einsum("ijij->ij", T)
    assert (d >= 0).all()
    assert (T ** 2 <= d[:, :, None, None] * d[None, None, :, :] + 1e-14).all()
    assert abs(float(g["E"]) - (-74.96)) < 1e-2  # literature STO-3G water, sanity only


def test_fill_matrices_equals_dense(water_shells):
    Z, xyz, sh = water_shells
    blocks = O.eri_blocks(sh)
    T = O.eri_tensor(sh, blocks)
    rng = np.random.default_rng(5)
    Pa = rng.standard_normal((7, 7))
    Pa = Pa + Pa.T
    Pb = rng.standard_normal((7, 7))
    Pb = Pb + Pb.T
    got = O.fill_matrices(sh, blocks, Pa + Pb, Pa, Pb)
    ref = O.jk_dense(T, Pa + Pb, Pa, Pb)
    for x, y in zip(got, ref):
        assert np.abs(x - y).max() < 1e-12


def test_scf_invariants(water_shells):
    Z, xyz, sh = water_shells
    g = golden("water_sto3g.npz")
    S, P = g["S"], g["P_a"]
    assert np.isclose(np.trace(P @ S), 5.0, atol=1e-10)
    assert np.allclose(P @ S @ P, P, atol=1e-10)
    F = g["T"] + g["V"] + O.jk_dense(g["ERI"], 2 * P, P, P)[0] - O.jk_dense(g["ERI"], 2 * P, P, P)[1]
    assert np.allclose(F @ P @ S, S @ P @ F, atol=1e-5)


def test_c_oracle_blocks_match_numpy_oracle():
    import oracle.ref as R
    from dft_amd.basis import build_shells, load_basis
    g = golden("dimer_sto3g.npz")
    sh = build_shells(load_basis(), list(g["Z"]), g["xyz"])
    seg = R.Segments(sh)
    q = R.unique_quartets(seg.n)
    rng = np.random.default_rng(2)
    Pa = rng.standard_normal((14, 14))
    Pa = Pa + Pa.T
    J, K, _ = R.build_jk(seg, q, 2 * Pa, Pa)
    Jr, Kr, _ = O.jk_dense(g["ERI"], 2 * Pa, Pa, Pa)
    assert np.abs(J - Jr).max() < 1e-12 and np.abs(K - Kr).max() < 1e-12


def test_spherical_transform_orthonormal_d():
    # one normalised primitive d shell: A S_cart A^T = I_5 (real solid harmonics)
    sh = [_prim_shell(2, 0.8, [0.0, 0.0, 0.0])]
    S, _, _ = O.one_electron(sh, [1], np.zeros((1, 3)))
    A = O.independent(sh, 5, np.eye(6))
    assert np.allclose(A @ S @ A.T, np.eye(5), atol=1e-12)


def test_spherical_transform_orthonormal_f():
    sh = [_prim_shell(3, 1.1, [0.0, 0.0, 0.0])]
    S, _, _ = O.one_electron(sh, [1], np.zeros((1, 3)))
    A = O.independent(sh, 7, np.eye(10))
    assert np.allclose(A @ S @ A.T, np.eye(7), atol=1e-12)


def test_centre_derivative_generates_d():
    # d/dA_x (px_A s_B|s_C s_D) = 2a (dxx_A s_B|..) - (s_A s_B|..), unnormalised primitives
    a, b, c, d = 0.7, 1.2, 0.9, 0.5
    A, B, Cc, D = np.array([0.2, -0.1, 0.4]), np.array([0.6, 0.3, -0.2]), np.array([-0.4, 0.5, 0.7]), \
        np.array([0.1, -0.6, 0.2])
    h = 1e-5

    def raw(l, Ax):
        s = [_prim_shell(l, a, Ax)] + [_prim_shell(0, e, r) for e, r in ((b, B), (c, Cc), (d, D))]
        return O.eri_prim(O.pair_data(s[0], s[1]), O.pair_data(s[2], s[3]))[:, 0]
    ex = np.array([1.0, 0, 0])
    fd = (raw(1, A + h * ex)[0] - raw(1, A - h * ex)[0]) / (2 * h)   # component px
    dxx = raw(2, A)[0]                                             # component xx
    ss = raw(0, A)[0]
    assert np.isclose(fd, 2 * a * dxx - ss, rtol=1e-7, atol=1e-12)


def test_c_oracle_matches_numpy_oracle_d():
    import oracle.ref as R
    from dft_amd.basis import build_shells_from_entries
    ent = [[[[2], [0.9, 3.0], [[0.6, 0.8]]], [[1], [1.3], [[1.0]]]], [[[0], [0.5, 2.0], [[0.7, 0.7]]]]]
    sh = build_shells_from_entries(ent, [[0, 0, 0], [0.5, 0.9, -1.1]])
    seg = R.Segments(sh)
    q = R.unique_quartets(seg.n)
    flat, sizes = R.eri_blocks(seg, q)
    pairs = O.all_pairs(sh)
    pos = 0
    for (a, b, c, d), sz in zip(q, sizes):
        ref = O.shell_quartet(sh, a, b, c, d, pairs).ravel()
        assert np.allclose(flat[pos:pos + sz], ref, rtol=1e-12, atol=1e-14)
        pos += sz


def _grid_overlap(shells, half, h, block):
    """sum_g phi_m(r_g) phi_n(r_g) h^3 on a cube of side 2*half (spacing h),
    evaluated in slabs; an exact-in-the-limit quadrature of S_mn for
    functions much wider than h."""
    ax = np.arange(-half, half + 1e-9, h)
    S = np.zeros((len(block), len(block)))
    for x in ax:
        g = np.stack(np.meshgrid([x], ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
        v = O.ao_values(shells, g)[block]
        S += v @ v.T
    return S * h ** 3


def test_ao_values_integrate_to_the_overlap_matrix():
    """The grid AO values (shelled_eval.py:98-113) squared and integrated
    give the overlap matrix of the independent closed form: O 2p block of
    STO-3G (a wrong monomial or angular factor fails) and one Cartesian d
    primitive with every component normalised (shelled.py:88-101):
    <xx|xx> = <xy|xy> = 1, <xx|yy> = 1/3, <xx|xy> = 0."""
    from dft_amd.basis import build_shells, build_shells_from_entries, load_basis
    sh = build_shells(load_basis(), [8], np.zeros((1, 3)))
    p = [k for k, s in enumerate(sh) if s["angular_momentum"] == 1][0]
    blk = list(range(sh[p]["ao_start"], sh[p]["ao_stop"]))
    S, _, _ = O.one_electron(sh, [8], np.zeros((1, 3)))
    Sg = _grid_overlap(sh, 7.0, 0.1, blk)
    assert np.abs(Sg - S[np.ix_(blk, blk)]).max() < 1e-6
    d = build_shells_from_entries([[[[2], [0.8], [[1.0]]]]], np.zeros((1, 3)))
    Sd = _grid_overlap(d, 6.0, 0.1, list(range(6)))
    # components xx, xy, yy, xz, yz, zz (combinations order)
    assert abs(Sd[0, 0] - 1.0) < 1e-6 and abs(Sd[0, 2] - 1.0 / 3.0) < 1e-6
    assert abs(Sd[1, 1] - 1.0) < 1e-6 and abs(Sd[0, 1]) < 1e-9
