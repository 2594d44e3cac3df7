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


def test_centre_derivative_generates_f():
    # d/dA_x (dxx_A s_B|s_C s_D) = 2a (fxxx_A s_B|..) - 2 (px_A s_B|..), unnormalised primitives
    a, b, c, d = 0.8, 1.1, 0.6, 1.4
    A, B, Cc, D = np.array([0.1, 0.3, -0.2]), np.array([-0.5, 0.2, 0.4]), np.array([0.3, -0.6, 0.1]), \
        np.array([0.7, 0.1, -0.5])
    h = 1e-5

    def raw(l, Ax):
        s = [_prim_shell(l, a, Ax)] + [_prim_shell(0, e, r) for e, r in ((b, B), (c, Cc), (d, D))]
        return O.eri_prim(O.pair_data(s[0], s[1]), O.pair_data(s[2], s[3]))[:, 0]
    ex = np.array([1.0, 0, 0])
    fd = (raw(2, A + h * ex)[0] - raw(2, A - h * ex)[0]) / (2 * h)   # component xx
    fxxx = raw(3, A)[0]                                            # component xxx
    px = raw(1, A)[0]                                              # component x
    assert np.isclose(fd, 2 * a * fxxx - 2 * px, rtol=1e-7, atol=1e-12)


def test_c_oracle_matches_numpy_oracle_f():
    """The C restatement against the NumPy one for classes with an f shell
    (the GPU f-shell tests use the C oracle as the checker)."""
    import oracle.ref as R
    from dft_amd.basis import build_shells_from_entries
    ent = [[[[3], [1.3], [[1.0]]], [[1], [0.8], [[1.0]]]], [[[0], [0.5, 2.0], [[0.7, 0.7]]]]]
    sh = build_shells_from_entries(ent, [[0, 0, 0], [0.4, -0.7, 0.9]])
    seg = R.Segments(sh)
    q = R.unique_quartets(seg.n)
    keep = [k for k, x in enumerate(q) if sum(int(seg.l[i]) for i in x) <= 5]  # up to (fp|ps)-size classes
    q = q[keep]
    flat, sizes = R.eri_blocks(seg, q)
    pairs = O.all_pairs(sh)
    pos = 0
    for (a, b, c, d), sz in zip(q, sizes):
        ref = O.shell_quartet(sh, a, b, c, d, pairs).ravel()
        assert np.allclose(flat[pos:pos + sz], ref, rtol=1e-12, atol=1e-14)
        pos += sz


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


# ------------------------------------------------------------------ 1-e pins
def _prim_shells(specs):
    """Single-primitive shells (coefficient 1): specs = [(l, alpha, centre)]."""
    from dft_amd.basis import build_shells_from_entries
    entries = [[[[l], [alpha], [[1.0]]]] for l, alpha, _ in specs]
    return build_shells_from_entries(entries, np.array([c for _, _, c in specs], float))


def _kin_axis(l, a):
    """-1/2 <x^l e^{-a x^2} | d2/dx2 | x^l e^{-a x^2}> / <x^l | x^l> of one axis."""
    return a * (2 * l + 1) / 2 - (2 * a * l * (l - 1) / (2 * l - 1) if l > 1 else 0.0)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_kinetic_diagonal_closed_forms(l):
    """T of a normalised primitive is a sum of per-axis closed forms
    (s 3a/2, p_x 5a/2, d_xx 13a/6, d_xy 7a/2, f_xxx 21a/10, ...); the d and f
    rows take the pow_b >= 2 branch of T_raw (shelled.py:149-155)."""
    from dft_amd.basis import combinations
    a = 0.73
    sh = _prim_shells([(l, a, (0.1, -0.2, 0.3))])
    _, T, _ = O.one_electron(sh, [], np.zeros((0, 3)))
    for k, c in enumerate(combinations(l)):
        want = sum(_kin_axis(int(x), a) for x in c)
        assert abs(T[k, k] - want) < 1e-12 * max(1.0, want), (l, c, T[k, k], want)
    assert abs(_kin_axis(0, a) * 3 - 1.5 * a) < 1e-15
    assert abs(_kin_axis(2, a) + 2 * _kin_axis(0, a) - 13 * a / 6) < 1e-15


def _quad_1d(la, a, A, lb, b, B):
    """Overlap and -1/2 <g_a|d2/dx2|g_b> = +1/2 <g_a'|g_b'> of two unnormalised
    1-D Gaussians by the trapezoidal rule on a uniform grid (spectrally
    accurate for these integrands): independent of the binomial overlap and
    of the T_raw recurrence."""
    w = 1.0 / math.sqrt(min(a, b))
    x = np.linspace(min(A, B) - 14 * w, max(A, B) + 14 * w, 40001)
    h = x[1] - x[0]
    xa, xb = x - A, x - B
    ga = xa ** la * np.exp(-a * xa ** 2)
    gb = xb ** lb * np.exp(-b * xb ** 2)
    da = ((la * xa ** (la - 1) if la else 0.0) - 2 * a * xa ** (la + 1)) * np.exp(-a * xa ** 2)
    db = ((lb * xb ** (lb - 1) if lb else 0.0) - 2 * b * xb ** (lb + 1)) * np.exp(-b * xb ** 2)
    return h * np.sum(ga * gb), 0.5 * h * np.sum(da * db)


@pytest.mark.parametrize("la,lb", [(0, 0), (1, 0), (2, 1), (2, 2), (3, 2), (1, 3)])
def test_kinetic_and_overlap_two_centres_vs_quadrature(la, lb):
    """Two-centre S and T blocks of the oracle against 1-D quadrature of
    S_x S_y S_z and T = T_x S_y S_z + S_x T_y S_z + S_x S_y T_z."""
    from dft_amd.basis import combinations, normal
    a, b = 0.9, 0.45
    A, B = np.array([0.2, -0.4, 0.5]), np.array([-0.6, 0.3, 1.1])
    sh = _prim_shells([(la, a, A), (lb, b, B)])
    S, T, _ = O.one_electron(sh, [], np.zeros((0, 3)))
    na = len(combinations(la))
    for i, ca in enumerate(combinations(la)):
        for j, cb in enumerate(combinations(lb)):
            s1 = [_quad_1d(ca[d], a, A[d], cb[d], b, B[d]) for d in range(3)]
            ov = np.prod([q[0] for q in s1])
            kin = sum(s1[d][1] * np.prod([s1[e][0] for e in range(3) if e != d]) for d in range(3))
            nrm = normal([a], [ca])[0] * normal([b], [cb])[0]
            assert abs(S[i, na + j] - nrm * ov) < 1e-11, (ca, cb)
            assert abs(T[i, na + j] - nrm * kin) < 1e-11, (ca, cb, T[i, na + j], nrm * kin)


def _df2(n):
    return float(np.prod(np.arange(n, 0, -2))) if n > 0 else 1.0


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_nuclear_attraction_one_centre_closed_form(l):
    """<g_i|-Z/r|g_j> for normalised primitives of one shell centred on the
    nucleus: radial integral Gamma((L_i+L_j)/2+1) / (2 (2a)^(...)) times the
    angular integral 4 pi prod (n_d - 1)!! / (n+1)!! of x^n over the sphere
    (zero for any odd power); the diagonal is -Z Gamma(l+1) sqrt(2a) / Gamma(l+3/2)."""
    from dft_amd.basis import combinations, normal
    a, Zc, C = 1.3, 2.0, np.array([0.3, 0.1, -0.7])
    sh = _prim_shells([(l, a, C)])
    _, _, V = O.one_electron(sh, [Zc], C[None, :])
    comps = combinations(l)
    for i, ci in enumerate(comps):
        for j, cj in enumerate(comps):
            n = [ci[d] + cj[d] for d in range(3)]
            if any(x % 2 for x in n):
                want = 0.0
            else:
                ang = 4 * math.pi * np.prod([_df2(x - 1) for x in n]) / _df2(sum(n) + 1)
                rad = math.gamma(sum(n) / 2 + 1) / (2 * (2 * a) ** (sum(n) / 2 + 1))
                want = -Zc * normal([a], [ci])[0] * normal([a], [cj])[0] * ang * rad
            assert abs(V[i, j] - want) < 1e-12 * max(1.0, abs(want)), (ci, cj, V[i, j], want)
        diag = -Zc * math.gamma(l + 1) * math.sqrt(2 * a) / math.gamma(l + 1.5)
        assert abs(V[i, i] - diag) < 1e-12 * abs(diag)


def test_nuclear_attraction_off_centre_from_s_closed_form():
    """p and d attraction integrals with the nucleus off both centres from the
    s closed form V_ss = -Z (2 pi / p) exp(-ab/p R_AB^2) F_0(p |P - C|^2),
    F_0 = sqrt(pi/T) erf(sqrt T) / 2, by the centre-derivative identities
    x_A e^{-a r_A^2} = (1/2a) d/dA_x e^{-a r_A^2} and
    x_A^2 e^{-a r_A^2} = (d2/dA_x^2 + 2a) e^{-a r_A^2} / (4 a^2)."""
    from scipy.special import erf
    from dft_amd.basis import combinations, normal
    a, b, Zc = 0.8, 0.55, 1.7
    A, B, C = np.array([0.1, 0.2, -0.3]), np.array([-0.7, 0.4, 0.6]), np.array([0.5, -0.9, 0.2])

    def vss(Ap):
        p = a + b
        P = (a * Ap + b * B) / p
        T = p * np.sum((P - C) ** 2)
        f0 = 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))
        return -Zc * 2 * math.pi / p * math.exp(-a * b / p * np.sum((Ap - B) ** 2)) * f0

    h = 1e-4
    e = np.eye(3)

    def d1(k):
        return (vss(A + h * e[k]) - vss(A - h * e[k])) / (2 * h)

    def d2(k, m):
        return (vss(A + h * e[k] + h * e[m]) - vss(A + h * e[k] - h * e[m]) - vss(A - h * e[k] + h * e[m])
                + vss(A - h * e[k] - h * e[m])) / (4 * h * h)
    for l in (1, 2):
        sh = _prim_shells([(l, a, A), (0, b, B)])
        _, _, V = O.one_electron(sh, [Zc], C[None, :])
        nb = normal([b], [[0, 0, 0]])[0]
        for i, c in enumerate(combinations(l)):
            axes = [d for d in range(3) for _ in range(c[d])]
            if l == 1:
                raw = d1(axes[0]) / (2 * a)
            elif axes[0] != axes[1]:
                raw = d2(axes[0], axes[1]) / (4 * a * a)
            else:
                raw = (d2(axes[0], axes[0]) + 2 * a * vss(A)) / (4 * a * a)
            want = normal([a], [c])[0] * nb * raw
            got = V[i, len(combinations(l))]
            assert abs(got - want) < 2e-6 * abs(want) + 1e-9, (l, c, got, want)


# ---------------------------------------------------------------- SCF pins
def test_oracle_scf_diis_converges_to_the_roothaan_energy(water_shells):
    """The damped + DIIS loop of shelled.py:1103-1292 and the plain loop of
    main.py:753-828 have the same fixed point: their converged energies agree
    (the schedule changes the path, not the solution); the DIIS loop switches
    to DIIS and needs more iterations than the undamped one."""
    Z, xyz, sh = water_shells
    S, T, V = O.one_electron(sh, Z, xyz)
    ERI = O.eri_tensor(sh)
    jk = lambda P, a, b: O.jk_dense(ERI, P, a, b)  # noqa: E731
    E_NN = O.nuclear_repulsion(Z, xyz)
    plain = O.scf_plain(T + V, S, jk, E_NN, 5, 5, e_tol=1e-12, p_tol=1e-10)
    diis = O.scf_diis(T + V, S, jk, E_NN, 5, 5, e_tol=1e-12, p_tol=1e-10)
    assert plain["converged"] and diis["converged"]
    assert abs(plain["E"] - diis["E"]) < 1e-10
    # at the reference's own tolerances the damped loop stops earlier on its
    # path: its exit energy uses the damped (non-idempotent) density, whose
    # energy error is first order in the remaining density change
    d8 = O.scf_diis(T + V, S, jk, E_NN, 5, 5)
    assert d8["converged"] and abs(d8["E"] - plain["E"]) < 1e-7
    assert diis["count"] > plain["count"]
    assert abs(diis["E"] - float(golden("water_sto3g.npz")["E"])) < 1e-8
    # damping keeps the energy sequence of the first iterations monotone here
    e = diis["energies"]
    assert e[1] < e[0]


def test_oracle_process_eri_invariants(tmp_path):
    """Restated process_eri on a synthetic store: processed values are the
    quartets' values in the stable lexsort order of their shapes, each block
    holds one shape, consecutive blocks differ, and the blocks partition the
    list."""
    rng = np.random.default_rng(4)
    nq = 300
    shapes = rng.choice([1, 3, 6], size=(nq, 4)).astype(np.int16)
    sizes = shapes.astype(np.int64).prod(1)
    ends = np.cumsum(sizes)
    idx = np.stack([ends - sizes, ends], 1).astype(np.int32)
    vals = rng.standard_normal(int(ends[-1]))
    quads = rng.integers(0, 50, size=(nq, 4)).astype(np.int16)
    for name, arr in (("eri_values", vals), ("eri_indices", idx), ("eri_quads", quads), ("eri_quads_shapes", shapes)):
        np.save(tmp_path / f"{name}.npy", arr)
    O.process_eri(str(tmp_path))
    pv = np.load(tmp_path / "processed_eri_values.npy")
    pi = np.load(tmp_path / "processed_eri_indices.npy")
    pq = np.load(tmp_path / "processed_eri_quads.npy")
    bl = np.load(tmp_path / "block_idxs.npy")
    cs = np.load(tmp_path / "chunked_quad_shape.npy")
    order = sorted(range(nq), key=lambda k: tuple(shapes[k]))  # stable: ties keep input order
    assert pq.dtype == np.int32 and np.array_equal(pq, quads[order])
    assert np.array_equal(np.sort(pv), np.sort(vals))
    for k, q in enumerate(order):
        assert np.array_equal(pv[pi[k, 0]:pi[k, 1]], vals[idx[q, 0]:idx[q, 1]])
    assert bl[0, 0] == 0 and bl[-1, 1] == nq and np.array_equal(bl[1:, 0], bl[:-1, 1])
    ss = shapes[order]
    for (s0, s1), shp in zip(bl, cs):
        assert (ss[s0:s1] == shp).all()
    assert all((cs[k] != cs[k + 1]).any() for k in range(len(cs) - 1))
