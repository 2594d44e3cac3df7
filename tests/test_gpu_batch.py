"""Batched plans (many molecules, one J/K build), hipGraph replay and the
launch-shape autotuner, against single-molecule plans."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _mols():
    from dft_amd.basis import build_shells, build_shells_from_entries, load_basis
    from dft_amd.geometry import methane, synthetic_spd_basis, water, water_cluster
    b = load_basis()
    out = []
    Z, xyz = water()
    out.append((build_shells(b, Z, xyz), Z, xyz))
    Z, xyz = water_cluster((2, 2, 1), 3.1, seed=5)
    out.append((build_shells(b, Z, xyz), Z, xyz))
    Z, xyz = methane()
    out.append((build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz), Z, xyz))
    g = golden("dimer_sto3g.npz")
    Z = list(g["Z"])
    out.append((build_shells(b, Z, g["xyz"]), Z, g["xyz"]))
    return out


def _dens(n, rng):
    A = rng.standard_normal((n, n)) * 0.1
    return A + A.T


@pytest.mark.parametrize("uhf", [False, True])
def test_batch_jk_bitwise_equal_to_single(torch_cuda, uhf):
    torch = torch_cuda
    from dft_amd.engine import BatchPlan, Plan
    mols = _mols()
    # no primitive split: the split of a class depends on its quartet count,
    # which differs between the batch and the single plans
    bp = BatchPlan(mols, schwarz_tau=1e-12, fp32_tau=0.0, uhf=uhf, split=False)
    rng = np.random.default_rng(0)
    Pa = [_dens(n, rng) for n in bp.nbfs]
    Pb = [_dens(n, rng) for n in bp.nbfs]
    Pt = bp.pack([a + b for a, b in zip(Pa, Pb)])
    J, Ka, Kb = bp.jk(Pt, bp.pack(Pa), bp.pack(Pb) if uhf else None)
    for m, (sh, Z, xyz) in enumerate(mols):
        p = Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=0.0, uhf=uhf, split=False)
        j, ka, kb = p.jk(Pa[m] + Pb[m], Pa[m], Pb[m] if uhf else None)
        assert torch.equal(bp.unpack(J)[m], j)
        assert torch.equal(bp.unpack(Ka)[m], ka)
        if uhf:
            assert torch.equal(bp.unpack(Kb)[m], kb)


def test_batch_one_electron_equals_single(torch_cuda):
    torch = torch_cuda
    from dft_amd.engine import BatchPlan, Plan
    mols = _mols()
    bp = BatchPlan(mols)
    S, T, V = bp.one_electron()
    for m, (sh, Z, xyz) in enumerate(mols):
        s, t, v = Plan(sh, Z, xyz).one_electron()
        for a, b in ((S, s), (T, t), (V, v)):
            assert torch.allclose(bp.unpack(a)[m], b, rtol=0, atol=1e-14)


def test_graph_replays_match_eager(torch_cuda):
    torch = torch_cuda
    from dft_amd.engine import BatchPlan, Graph
    mols = _mols()
    bp = BatchPlan(mols, schwarz_tau=1e-12)
    rng = np.random.default_rng(1)
    Pa = bp.pack([_dens(n, rng) for n in bp.nbfs])
    P = (2 * Pa).contiguous()
    J, K = bp.empty(), bp.empty()
    g = Graph(bp, P, Pa, None, J, K, None, "mixed")
    for rep in range(4):
        new = bp.pack([_dens(n, rng) for n in bp.nbfs])
        Pa.copy_(new)
        torch.mul(Pa, 2.0, out=P)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        Je, Ke, _ = bp.jk(P, Pa, None, mode="mixed")
        assert torch.equal(J, Je) and torch.equal(K, Ke), rep


def test_tuner_keeps_results(torch_cuda):
    torch = torch_cuda
    from dft_amd.engine import BatchPlan, tune
    mols = _mols()
    bp = BatchPlan(mols, schwarz_tau=1e-12, fp32_tau=1e-4)
    rng = np.random.default_rng(2)
    Pa = bp.pack([_dens(n, rng) for n in bp.nbfs])
    P = (2 * Pa).contiguous()
    J0, K0, _ = bp.jk(P, Pa, None, mode="fp64")
    rep = tune(bp, P, Pa, None, reps=2)
    # 10 pair classes (s, p, d, f) -> 55 class combinations
    assert rep.shape == (55, 6) and (rep[:, 2] > 0).all()
    J1, K1, _ = bp.jk(P, Pa, None, mode="fp64")
    scale = max(1.0, J0.abs().max().item())
    assert (J1 - J0).abs().max().item() < 1e-12 * scale
    assert (K1 - K0).abs().max().item() < 1e-12 * scale


def test_scf_batch_matches_single(torch_cuda):
    from dft_amd.batch import scf_batch
    from dft_amd.engine import BatchPlan, Plan
    from dft_amd.scf import scf_plain
    mols = [m for m in _mols() if max(s["angular_momentum"] for s in m[0]) <= 1]
    bp = BatchPlan(mols, schwarz_tau=1e-12)
    nocc = [int(sum(Z)) // 2 for _, Z, _ in mols]
    res = scf_batch(bp, nocc, e_tol=1e-10, p_tol=1e-7, max_iter=200)
    assert res.converged.all()
    for m, (sh, Z, xyz) in enumerate(mols):
        single = scf_plain(Plan(sh, Z, xyz, schwarz_tau=1e-12), Z, xyz, nocc[m], nocc[m], e_tol=1e-10, p_tol=1e-7,
                           max_count=200)
        assert abs(res.energies[m] - single.energy) < 1e-8, (m, res.energies[m], single.energy)


@pytest.mark.parametrize("uhf", [False, True])
def test_primitive_split_matches_unsplit(torch_cuda, uhf):
    """Classes with few quartets are split over primitive ranges (partial J/K
    blocks summed by the reduction); the result equals the unsplit build to
    rounding."""
    torch = torch_cuda
    from dft_amd.engine import Plan
    for sh, Z, xyz in _mols():
        a = Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=0.0, uhf=uhf)
        b = Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=0.0, uhf=uhf, split=False)
        rng = np.random.default_rng(1)
        n = a.nbf
        Pa, Pb = _dens(n, rng), _dens(n, rng)
        ra = a.jk(Pa + Pb, Pa, Pb if uhf else None)
        rb = b.jk(Pa + Pb, Pa, Pb if uhf else None)
        for x, y in zip(ra, rb):
            if x is None:
                continue
            assert float((x - y).abs().max()) < 1e-12


@pytest.mark.parametrize("uhf", [False, True])
def test_lane_table_split_everywhere_matches_unsplit(torch_cuda, monkeypatch, uhf):
    """A split threshold above every class's quartet count puts each s/p
    class (and the light d classes) through lane tables: bra chunks in
    neighbouring lanes, waves closed at molecule and FP64 / FP32 part
    boundaries, ket chunks as slot rows. FP64 and mixed builds of single and
    batch plans equal the unsplit builds to rounding (mixed: the FP32 partial
    integrals of a chunk are rounded before the chunks are added)."""
    torch = torch_cuda
    from dft_amd.engine import BatchPlan, Plan
    mols = _mols()
    monkeypatch.setenv("DFT_SPLIT_LANES", str(1 << 26))
    split = [Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=1e-4, uhf=uhf) for sh, Z, xyz in mols]
    bsplit = BatchPlan(mols, schwarz_tau=1e-12, fp32_tau=1e-4, uhf=uhf)
    monkeypatch.delenv("DFT_SPLIT_LANES")
    plain = [Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=1e-4, uhf=uhf, split=False) for sh, Z, xyz in mols]
    rng = np.random.default_rng(7)
    Pa = [_dens(p.nbf, rng) for p in plain]
    Pb = [_dens(p.nbf, rng) for p in plain]
    for mode, tol in (("fp64", 1e-12), ("mixed", 1e-10)):
        Jb, Kab, Kbb = bsplit.jk(bsplit.pack([a + b for a, b in zip(Pa, Pb)]), bsplit.pack(Pa),
                                  bsplit.pack(Pb) if uhf else None, mode=mode)
        for m in range(len(mols)):
            rs = split[m].jk(Pa[m] + Pb[m], Pa[m], Pb[m] if uhf else None, mode=mode)
            rp = plain[m].jk(Pa[m] + Pb[m], Pa[m], Pb[m] if uhf else None, mode=mode)
            rb = (bsplit.unpack(Jb)[m], bsplit.unpack(Kab)[m], bsplit.unpack(Kbb)[m] if uhf else None)
            for x, y, z in zip(rs, rp, rb):
                if x is None:
                    continue
                assert float((x - y).abs().max()) < tol, (mode, m)
                assert float((z - y).abs().max()) < tol, (mode, m)


def test_serial_and_concurrent_launches_bitwise_equal(torch_cuda):
    """The class launches of a build write disjoint staging slots and the
    reduction order is fixed, so running them on one stream or spread over
    four gives bitwise-identical J/K (s/p and the heavy d classes)."""
    torch = torch_cuda
    from dft_amd.engine import Plan
    for sh, Z, xyz in _mols():
        a = Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=1e-4)
        b = Plan(sh, Z, xyz, schwarz_tau=1e-12, fp32_tau=1e-4, serial=True)
        rng = np.random.default_rng(3)
        Pa = _dens(a.nbf, rng)
        for mode in ("fp64", "mixed"):
            ja, ka, _ = a.jk(2 * Pa, Pa, mode=mode)
            jb, kb, _ = b.jk(2 * Pa, Pa, mode=mode)
            assert torch.equal(ja, jb) and torch.equal(ka, kb), mode
