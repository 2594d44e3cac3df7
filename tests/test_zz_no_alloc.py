"""dft_jk_static, the allocation-free variant of dft_jk, through the C ABI.

Its J/K must be bitwise equal to dft_jk's on the same plan and densities
(same kernels, workspace in the library's static scratch); two calls in a
row, the second with a smaller molecule, so that the second runs on scratch
the first left filled; plans built without a workspace of their own
(DFT_PLAN_STATIC_WORKSPACE) work only through it; a plan whose workspace
exceeds the scratch is rejected with an error.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _molecule(name):
    from dft_amd.basis import build_shells, load_basis
    g = golden(name)
    Z = list(g["Z"])
    return Z, g["xyz"], build_shells(load_basis(), Z, g["xyz"])


def _density(torch, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    Pa = torch.as_tensor((A + A.T) / 4, device="cuda")
    return (2 * Pa).contiguous(), Pa.contiguous()


def _static(torch, plan, P, Pa, Pb=None, mode="fp64"):
    from dft_amd import _lib
    J, Ka = torch.empty_like(P), torch.empty_like(P)
    Kb = torch.empty_like(P) if Pb is not None else None
    s = torch.cuda.current_stream()
    rc = plan.lib.dft_jk_static(plan.handle, P.data_ptr(), Pa.data_ptr(), Pb.data_ptr() if Pb is not None else None,
                                J.data_ptr(), Ka.data_ptr(), Kb.data_ptr() if Kb is not None else None,
                                _lib.MODES[mode], C.c_void_p(s.cuda_stream))
    return rc, J, Ka, Kb


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_static_matches_dft_jk_two_calls_second_smaller(torch_cuda, mode):
    from dft_amd import _lib
    from dft_amd.engine import Plan
    torch = torch_cuda
    for k, name in enumerate(("dimer_sto3g.npz", "water_sto3g.npz")):  # second call: smaller input
        Z, xyz, shells = _molecule(name)
        plan = Plan(shells, Z, xyz, schwarz_tau=1e-12, fp32_tau=1e-4)
        assert plan.stats()["workspace_bytes"] <= _lib.DFT_STATIC_SCRATCH_BYTES
        P, Pa = _density(torch, plan.nbf, k)
        J0, K0, _ = plan.jk(P, Pa, mode=mode)
        rc, J1, K1, _ = _static(torch, plan, P, Pa, mode=mode)
        _lib.check(rc)
        torch.cuda.synchronize()
        assert torch.equal(J0, J1) and torch.equal(K0, K1), name


def test_static_workspace_plan_uhf_and_repeat(torch_cuda):
    """A plan built without its own workspace (UHF, two exchange densities):
    dft_jk refuses it, dft_jk_static gives the J/K of a regular plan, twice
    in a row bitwise the same."""
    from dft_amd import _lib
    from dft_amd.engine import Plan
    torch = torch_cuda
    Z, xyz, shells = _molecule("water_sto3g.npz")
    ref = Plan(shells, Z, xyz, schwarz_tau=0.0, uhf=True)
    bare = Plan(shells, Z, xyz, schwarz_tau=0.0, uhf=True, static_workspace=True)
    P, Pa = _density(torch, ref.nbf, 5)
    Pb = (P - Pa).contiguous() * 0.5
    J0, Ka0, Kb0 = ref.jk(P, Pa, Pb)
    with pytest.raises(_lib.DFTError, match="DFT_PLAN_STATIC_WORKSPACE"):
        bare.jk(P, Pa, Pb)
    outs = []
    for _ in range(2):
        rc, J1, Ka1, Kb1 = _static(torch, bare, P, Pa, Pb)
        _lib.check(rc)
        outs.append((J1, Ka1, Kb1))
    torch.cuda.synchronize()
    for J1, Ka1, Kb1 in outs:
        assert torch.equal(J0, J1) and torch.equal(Ka0, Ka1) and torch.equal(Kb0, Kb1)


def test_static_rejects_a_plan_beyond_the_scratch(torch_cuda):
    from dft_amd import _lib
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Plan
    from dft_amd.geometry import water_cluster
    torch = torch_cuda
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    plan = Plan(build_shells(load_basis(), Z, xyz), Z, xyz, schwarz_tau=1e-10, static_workspace=True)
    assert plan.stats()["workspace_bytes"] > _lib.DFT_STATIC_SCRATCH_BYTES
    P, Pa = _density(torch, plan.nbf, 1)
    rc, *_ = _static(torch, plan, P, Pa)
    assert rc == _lib.DFT_EINVAL
    assert b"static scratch holds" in plan.lib.dft_last_error()
