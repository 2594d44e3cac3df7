"""f shells through the integral and J/K kernels (the class combinations
with an f shell run the run-time-momentum kernels) against the oracle:
contracted ERI blocks of every unique shell quartet and direct J/K (RHF and
UHF densities) against the C restatement of the reference's
per-primitive-Cartesian McMurchie-Davidson path (itself checked against the
NumPy restatement for f classes in test_oracle.py); S, T, V against the
NumPy oracle; the damped + DIIS SCF energy against the oracle's SCF loop fed
by the C oracle's J/K. Reference capacity: L <= 12, f transform
(shelled.py:168-171,901-919)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _molecule():
    """HF-like pair: heavy centre with s, p, d and two-primitive f shells,
    light centre with s and f; 31 Cartesian functions, 10 electrons."""
    from dft_amd.basis import build_shells_from_entries
    heavy = [[[0], [74.0, 11.5, 2.4], [[0.15, 0.55, 0.45]]], [[0], [0.65], [[1.0]]],
             [[1], [4.9, 1.05], [[0.45, 0.65]]], [[2], [1.3], [[1.0]]], [[3], [1.7, 0.55], [[0.55, 0.6]]]]
    light = [[[0], [3.4, 0.62], [[0.45, 0.7]]], [[3], [0.9], [[1.0]]]]
    Z = [9, 1]
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.73]])
    return Z, xyz, build_shells_from_entries([heavy, light], xyz)


@pytest.fixture(scope="module")
def fmol(torch_cuda):
    from dft_amd.engine import Plan
    Z, xyz, shells = _molecule()
    return Z, xyz, shells, Plan(shells, Z, xyz, schwarz_tau=0.0, fp32_tau=1e-4)


def _all_quartets(n):
    ii, jj = np.tril_indices(n)
    bra, ket = np.tril_indices(len(ii))
    return np.stack([ii[bra], jj[bra], ii[ket], jj[ket]], axis=1).astype(np.int32)


def test_f_eri_blocks_vs_c_oracle(fmol):
    import oracle.ref as ref
    Z, xyz, shells, plan = fmol
    seg = ref.Segments(shells)
    q = _all_quartets(seg.n)
    got = plan.eri_segment_blocks(q)
    want, _ = ref.eri_blocks(seg, q)
    err = float(np.abs(got - want).max())
    lmax = seg.l[q].max(axis=1)
    print(f"{len(q)} quartets ({int((lmax == 3).sum())} with an f shell), max abs err {err:.2e}, "
          f"max |ERI| {float(np.abs(want).max()):.2e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("uhf", [False, True])
def test_f_jk_vs_c_oracle(fmol, torch_cuda, uhf):
    import oracle.ref as ref
    from dft_amd.engine import Plan
    torch = torch_cuda
    Z, xyz, shells, plan = fmol
    if uhf:
        plan = Plan(shells, Z, xyz, schwarz_tau=0.0, uhf=True)
    n = plan.nbf
    rng = np.random.default_rng(7)
    dens = []
    for _ in range(2):
        A = rng.standard_normal((n, n)) * 0.1
        dens.append((A + A.T) / 2)
    Pa = dens[0]
    Pb = dens[1] if uhf else dens[0]
    P = Pa + Pb
    seg = ref.Segments(shells)
    q = _all_quartets(seg.n)
    Jr, Kar, Kbr = ref.build_jk(seg, q, P, Pa, Pb if uhf else None)
    t = lambda x: torch.as_tensor(x, device="cuda")
    for mode in ("fp64", "mixed"):
        J, Ka, Kb = plan.jk(t(P), t(Pa), t(Pb) if uhf else None, mode=mode)
        dj = float(np.abs(J.cpu().numpy() - Jr).max())
        dk = float(np.abs(Ka.cpu().numpy() - Kar).max())
        if uhf:
            dk = max(dk, float(np.abs(Kb.cpu().numpy() - Kbr).max()))
        print(f"uhf={uhf} {mode}: max|dJ| {dj:.2e}, max|dK| {dk:.2e}", flush=True)
        # mixed: the f classes run in FP64, the rest may run FP32 quartets
        tol = 1e-10 if mode == "fp64" else 1e-7
        assert dj < tol and dk < tol, (mode, dj, dk)


def test_f_one_electron_vs_numpy_oracle(fmol):
    import oracle.md_oracle as O
    Z, xyz, shells, plan = fmol
    S, T, V = (m.cpu().numpy() for m in plan.one_electron())
    So, To, Vo = O.one_electron(shells, Z, xyz)
    for name, a, b in (("S", S, So), ("T", T, To), ("V", V, Vo)):
        err = float(np.abs(a - b).max())
        print(f"{name}: max abs err {err:.2e}", flush=True)
        assert err < 1e-10, name
        assert np.array_equal(a, a.T), name


def test_f_scf_energy_vs_oracle(fmol):
    import oracle.md_oracle as O
    import oracle.ref as ref
    from dft_amd.scf import scf_diis
    Z, xyz, shells, plan = fmol
    res = scf_diis(plan, Z, xyz, 5, 5, e_tol=1e-10, p_tol=1e-8)
    seg = ref.Segments(shells)
    q = _all_quartets(seg.n)
    S, T, V = O.one_electron(shells, Z, xyz)

    def jk(P, Pa, Pb):
        return ref.build_jk(seg, q, P, Pa, Pb)
    want = O.scf_diis(T + V, S, jk, O.nuclear_repulsion(Z, xyz), 5, 5, e_tol=1e-10, p_tol=1e-8)
    e_or = want["E"] if isinstance(want, dict) else want[0]
    print(f"SCF with f shells: E {res.energy:.10f} ({res.iterations} iterations), oracle {e_or:.10f}", flush=True)
    assert res.converged
    assert abs(res.energy - e_or) < 1e-8
