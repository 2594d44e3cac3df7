"""The benchmark workloads: water (plain RHF), (H2O)16 direct J/K, the
C12H26 s/p/d mixed-precision build, the batch of 64 water clusters, the
C8H18 analytic gradient and the (H2O)16 grid density. Each function runs one workload on the GPU and
returns a dict of measured numbers (device times from HIP events).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .basis import build_shells, build_shells_from_entries, load_basis
from .engine import BatchPlan, Graph, Plan, tune
from .geometry import alkane_bohr, synthetic_spd_basis, water
from .scf import electron_counts, scf_diis, scf_plain


def log(msg):
    import sys
    print(f"[workload] {msg}", file=sys.stderr, flush=True)


def _events():
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_builds(fn, steps, warmup):
    """Median-free total device time of ``steps`` calls of fn() (after warmup), ms per call."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def water_rhf():
    """Plain RHF water / STO-3G (SCF to 1e-8 Eh, max|dP| 1e-6 and RMS dP 1e-6, no DIIS)."""
    Z, xyz = water()
    shells = build_shells(load_basis(), Z, xyz)
    t0 = time.perf_counter()
    plan = Plan(shells, Z, xyz, schwarz_tau=0.0)
    res = scf_plain(plan, Z, xyz, 5, 5, p_rms_tol=1e-6)
    torch.cuda.synchronize()
    return {"energy_Eh": res.energy, "converged": res.converged, "iterations": res.iterations,
            "s_total": time.perf_counter() - t0, "s_per_fock_build": float(np.median(res.fock_times))}


def alkane_spd_mixed(steps=3, warmup=1):
    """C12H26 with the synthetic s/p/d basis (seed 1): FP64 vs mixed direct J/K.

    SCF in the spherical basis with damping + DIIS (shelled.py conventions).
    max|dF| is the Fock-matrix difference of the two modes at the converged
    FP64 density; |dE| the difference of the converged SCF energies.
    """
    from .shelled import spherical_matrix
    Z, xyz = alkane_bohr(12)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    t0 = time.perf_counter()
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    st = plan.stats()
    n_sph = shells[-1]["sphere_ao_stop"]
    A = torch.as_tensor(spherical_matrix(shells, n_sph), device=plan.device)
    S, _, _ = plan.one_electron()
    smin = float(torch.linalg.eigvalsh(A @ S @ A.T)[0])
    N_a, N_b = electron_counts(Z)
    log(f"C12H26 spd plan {t_plan:.1f} s: {st['n_quartets']} quartets ({st['n_quartets_fp32']} FP32), "
        f"min eig S {smin:.2e}")
    r64 = scf_diis(plan, Z, xyz, N_a, N_b, mode="fp64", A=A, max_count=200)
    log(f"SCF fp64: {r64.iterations} iterations, E {r64.energy:.10f}")
    r32 = scf_diis(plan, Z, xyz, N_a, N_b, mode="mixed", A=A, max_count=200)
    log(f"SCF mixed: {r32.iterations} iterations, E {r32.energy:.10f}")
    Pc = (A.T @ r64.P_a @ A).contiguous()
    P2 = (2 * Pc).contiguous()
    J64, K64, _ = plan.jk(P2, Pc, mode="fp64")
    J32, K32, _ = plan.jk(P2, Pc, mode="mixed")
    dF = float((A @ ((J64 - K64) - (J32 - K32)) @ A.T).abs().max())
    J, K = torch.empty_like(P2), torch.empty_like(P2)
    ms64 = time_builds(lambda: plan.jk_into(P2, Pc, None, J, K, None, "fp64"), steps, warmup)
    ms32 = time_builds(lambda: plan.jk_into(P2, Pc, None, J, K, None, "mixed"), steps, warmup)
    # the same builds after the autotuner (launch shapes, FP32 choice per class)
    t0 = time.perf_counter()
    tune(plan, P2, Pc, None, reps=2)
    t_tune = time.perf_counter() - t0
    ms64t = time_builds(lambda: plan.jk_into(P2, Pc, None, J, K, None, "fp64"), steps, warmup)
    ms32t = time_builds(lambda: plan.jk_into(P2, Pc, None, J, K, None, "mixed"), steps, warmup)
    log(f"builds untuned fp64 {ms64:.2f} / mixed {ms32:.2f} ms; tuned ({t_tune:.1f} s) {ms64t:.2f} / {ms32t:.2f} ms")
    return {"nbf_cart": plan.nbf, "nbf_sph": n_sph, "n_shells": len(shells), "min_eig_S": smin,
            "n_quartets": st["n_quartets"], "n_quartets_fp32": st["n_quartets_fp32"],
            "n_unscreened": st["n_unscreened"], "n_prim_quartets": st["n_prim_quartets"], "plan_s": t_plan,
            "energy_fp64_Eh": r64.energy, "energy_mixed_Eh": r32.energy, "converged": [r64.converged, r32.converged],
            "abs_dE_Eh": abs(r64.energy - r32.energy), "max_abs_dF_Eh": dF,
            "ms_per_build_fp64": ms64, "ms_per_build_mixed": ms32, "mixed_speedup": ms64 / ms32,
            "tune_s": t_tune, "ms_per_build_fp64_tuned": ms64t, "ms_per_build_mixed_tuned": ms32t,
            "mixed_speedup_tuned": ms64t / ms32t}


def water_batch_scf(steps=5, warmup=2, tune_reps=2, scf_check=True):
    """64 (H2O)_n clusters (n ~ U{1..24}, seed 2), STO-3G.

    Plain FP64 path: one plan per molecule, one dft_jk per molecule per
    iteration. Batched path: one plan for all molecules, one hipGraph launch
    per iteration, autotuned launch shapes; FP64 and mixed.
    """
    from .batch import scf_batch, water_batch
    mols = water_batch(64, 2)
    log(f"batch: {len(mols)} molecules, {sum(sh[-1]['ao_stop'] for sh, _, _ in mols)} bf")
    t0 = time.perf_counter()
    bp = BatchPlan(mols, schwarz_tau=1e-10, fp32_tau=1e-4)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    st = bp.stats()
    log(f"batch plan {t_plan:.1f} s: {st['n_quartets']} quartets of {st['n_unscreened']}")
    nocc = [int(sum(Z)) // 2 for _, Z, _ in mols]
    warm = scf_batch(bp, nocc, fixed_iters=3, use_graph=False)
    log("warm-up iterations done")
    P, Pa = warm.P.contiguous(), warm.Pa.contiguous()
    # plain FP64 HIP path: per-molecule plans and launches
    t0 = time.perf_counter()
    singles = [Plan(sh, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4) for sh, Z, xyz in mols]
    log(f"64 single-molecule plans {time.perf_counter() - t0:.1f} s")
    Pm = [x.contiguous() for x in bp.unpack(P)]
    Pam = [x.contiguous() for x in bp.unpack(Pa)]
    Jm = [torch.empty_like(x) for x in Pm]
    Km = [torch.empty_like(x) for x in Pm]

    def per_molecule(mode):
        for k, pl in enumerate(singles):
            pl.jk_into(Pm[k], Pam[k], None, Jm[k], Km[k], None, mode)
    ms_plain64 = time_builds(lambda: per_molecule("fp64"), steps, warmup)
    log(f"plain fp64 per molecule: {ms_plain64:.2f} ms/iter")
    del singles, Jm, Km
    J, K = bp.empty(), bp.empty()
    ms_batch64_eager = time_builds(lambda: bp.jk_into(P, Pa, None, J, K, None, "fp64"), steps, warmup)
    log(f"batched fp64 eager untuned: {ms_batch64_eager:.2f} ms/iter; tuning ...")
    t0 = time.perf_counter()
    shapes = tune(bp, P, Pa, None, reps=tune_reps, log=log)
    t_tune = time.perf_counter() - t0
    log(f"tuned in {t_tune:.1f} s")
    g64 = Graph(bp, P, Pa, None, J, K, None, "fp64")
    ms_batch64 = time_builds(g64.launch, steps, warmup)
    g32 = Graph(bp, P, Pa, None, J, K, None, "mixed")
    ms_batch32 = time_builds(g32.launch, steps, warmup)
    log(f"batched graph tuned: fp64 {ms_batch64:.2f} ms, mixed {ms_batch32:.2f} ms")
    out = {"n_molecules": len(mols), "n_quartets": st["n_quartets"], "n_quartets_fp32": st["n_quartets_fp32"],
           "n_unscreened": st["n_unscreened"], "n_prim_quartets": st["n_prim_quartets"], "plan_s": t_plan,
           "fock_workspace_bytes": st["workspace_bytes"], "tune_s": t_tune, "tuned_shapes": shapes.tolist(),
           "ms_per_iter_plain_fp64_per_molecule": ms_plain64, "ms_per_iter_batched_fp64_eager_untuned": ms_batch64_eager,
           "ms_per_iter_batched_fp64_graph_tuned": ms_batch64, "ms_per_iter_batched_mixed_graph_tuned": ms_batch32,
           "speedup_batched_mixed_vs_plain_fp64": ms_plain64 / ms_batch32,
           "speedup_batched_fp64_vs_plain_fp64": ms_plain64 / ms_batch64}
    if scf_check:
        r64 = scf_batch(bp, nocc, mode="fp64", max_iter=60)
        log(f"batched SCF fp64: {r64.iterations} iterations, {int(r64.converged.sum())} converged")
        r32 = scf_batch(bp, nocc, mode="mixed", max_iter=60)
        log(f"batched SCF mixed: {r32.iterations} iterations, {int(r32.converged.sum())} converged")
        out.update({"scf_converged_fp64": int(r64.converged.sum()), "scf_converged_mixed": int(r32.converged.sum()),
                    "scf_iterations": [r64.iterations, r32.iterations],
                    "max_abs_dE_mixed_Eh": float(np.abs(r64.energies - r32.energies).max()),
                    "s_per_iter_wall_fp64": float(np.median(r64.iter_times[1:]))})
    return out


def alkane_gradient(fd_coords=12, sd_steps=10, h=1e-4):
    """C8H18 (+ N(0, 0.05 A) noise, seed 3), STO-3G: analytic gradient, FD check
    on a subset of coordinates, and timed steepest-descent steps."""
    from .grad import electronic_gradient, energy_and_gradient, steepest_descent
    Z, xyz = alkane_bohr(8, noise_sigma_ang=0.05, seed=3)
    basis = load_basis()
    t0 = time.perf_counter()
    E, g, res = energy_and_gradient(Z, xyz, basis)
    t_call = time.perf_counter() - t0
    # the same call again: plan, SCF from the superposition of atomic
    # densities (cached per element) and gradient, with the process's
    # one-time library and kernel loading already done
    t0 = time.perf_counter()
    energy_and_gradient(Z, xyz, basis)
    t_call_warm = time.perf_counter() - t0
    # and from the core Hamiltonian guess (the reference's start)
    t0 = time.perf_counter()
    E_core, g_core, res_core = energy_and_gradient(Z, xyz, basis, guess="core")
    t_call_core = time.perf_counter() - t0
    # device time of the gradient kernels alone, on the same plan and orbitals
    shells = build_shells(basis, Z, xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-12, fp32_tau=0.0)
    N_a, N_b = electron_counts(Z)
    electronic_gradient(plan, res.C_a, res.eps_a, N_a, res.C_b, res.eps_b, N_b)  # builds the gradient data
    ms_grad = time_builds(lambda: electronic_gradient(plan, res.C_a, res.eps_a, N_a, res.C_b, res.eps_b, N_b), 3, 1)
    rng = np.random.default_rng(0)
    coords = [(int(a), int(x)) for a, x in zip(rng.choice(len(Z), fd_coords, replace=False),
                                                 rng.integers(0, 3, fd_coords))]
    fd_err = []
    t0 = time.perf_counter()
    for a, x in coords:
        p = xyz.copy()
        p[a, x] += h
        m = xyz.copy()
        m[a, x] -= h
        guess = (res.P_a, res.P_b)  # the reference geometry's densities seed both SCFs
        fd = (energy_and_gradient(Z, p, basis, guess=guess)[0]
              - energy_and_gradient(Z, m, basis, guess=guess)[0]) / (2 * h)
        fd_err.append(abs(fd - g[a, x]))
    t_fd = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, hist = steepest_descent(Z, xyz, basis, steps=sd_steps, step_size=0.3)
    t_sd = time.perf_counter() - t0
    return {"n_atoms": len(Z), "energy_Eh": E, "max_abs_grad": float(np.abs(g).max()),
            "s_per_gradient_call_total": t_call, "s_per_gradient_call_warm": t_call_warm,
            "scf_iterations_sad_guess": res.iterations, "s_per_gradient_call_core_guess": t_call_core,
            "scf_iterations_core_guess": res_core.iterations,
            "core_vs_sad_abs_dE_Eh": abs(E_core - E), "core_vs_sad_max_abs_dgrad": float(np.abs(g_core - g).max()),
            "ms_gradient_kernels": ms_grad,
            "fd_coords_checked": len(coords), "fd_max_abs_err_Eh_per_bohr": float(max(fd_err)),
            "fd_seconds": t_fd, "sd_steps": sd_steps, "sd_seconds": t_sd,
            "sd_energies": [hh[0] for hh in hist], "sd_s_per_step": [hh[2] for hh in hist],
            "sd_scf_iterations": [hh[3] for hh in hist]}


def water16_grid(spacing=0.05, chunk=1 << 22, steps=5):
    """(H2O)16 STO-3G (112 bf): AO values on the box grid around the cluster
    (3 bohr padding, the given spacing), in chunks; device time per chunk of
    the AO kernel against the HBM bytes it writes (8 B per value), and the
    whole density pass (AO values + rho = sum (P phi) phi) per point."""
    from .geometry import water_cluster
    from .grid import box_grid
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    xyz = np.asarray(xyz)
    shells = build_shells(load_basis(), Z, xyz)
    plan = Plan(shells, Z, xyz, schwarz_tau=1e-10)
    N_a, N_b = electron_counts(Z)
    res = scf_diis(plan, Z, xyz, N_a, N_b)
    P = (res.P_a + res.P_b).contiguous()
    lo, hi = xyz.min(0) - 3.0, xyz.max(0) + 3.0
    g, shape = box_grid(lo, hi, spacing)
    dv = float(np.prod((hi - lo) / (np.asarray(shape) - 1)))
    gd = torch.as_tensor(g[:chunk], device=plan.device)
    out = torch.empty((plan.nbf, gd.shape[0]), dtype=torch.float64, device=plan.device)
    ms_ao = time_builds(lambda: plan.ao_values(gd, out=out), steps, 1)
    gbps = plan.nbf * gd.shape[0] * 8 / (ms_ao * 1e-3) / 1e9
    # whole grid: electron count and time of the density pass
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    n = torch.zeros((), dtype=torch.float64, device=plan.device)
    for a in range(0, len(g), chunk):
        ao = plan.ao_values(g[a:a + chunk])
        n += ((P @ ao) * ao).sum()
    e1.record()
    torch.cuda.synchronize()
    ms_all = e0.elapsed_time(e1)
    log(f"grid {shape}: AO kernel {ms_ao:.3f} ms per {gd.shape[0]} points ({gbps:.0f} GB/s written), "
        f"electrons {float(n) * dv:.6f}")
    return {"grid_shape": list(shape), "n_points": int(len(g)), "nbf": plan.nbf, "chunk_points": int(gd.shape[0]),
            "ms_ao_kernel_per_chunk": ms_ao, "ao_kernel_write_GBps": gbps, "hbm_peak_GBps": 8000.0,
            "ao_kernel_hbm_frac": gbps / 8000.0, "ms_density_pass_whole_grid": ms_all,
            "ns_per_point_density_pass": ms_all * 1e6 / len(g), "electrons_integrated": float(n) * dv,
            "electrons": int(N_a + N_b)}


def alkane_spd_class_roofline(steps=3):
    """C12H26 with the synthetic s/p/d basis (seed 1), Schwarz 1e-10: every
    class combination's serial FP64 launch time (integrals + J/K digest, HIP
    events recorded by the library on the launch stream) on two plans, the
    heavy d classes on the column kernel and on the matrix-core variant,
    with the algorithmic model flops of each class and the fraction of the
    FP64 peak (78.6 TFLOP/s) of the faster variant."""
    from .engine import profile_read, profile_start
    from .flops import combo_flops, combo_moments
    Z, xyz = alkane_bohr(12)
    shells = build_shells_from_entries(synthetic_spd_basis(Z, seed=1), xyz)
    n = shells[-1]["ao_stop"]
    rng = np.random.default_rng(0)
    A = rng.standard_normal((n, n)) * 0.02
    Pa = torch.as_tensor((A + A.T) / 2, device="cuda").contiguous()
    P = (2 * Pa).contiguous()
    ms, st = {}, None
    for mf in (False, True):
        plan = Plan(shells, Z, xyz, schwarz_tau=1e-10, fp32_tau=1e-4, serial=True, mfma=mf)
        J = torch.empty((n, n), dtype=torch.float64, device="cuda")
        K = torch.empty_like(J)
        for _ in range(2):
            plan.jk_into(P, Pa, None, J, K, None, "fp64")
        profile_start(plan, steps)
        for _ in range(steps):
            plan.jk_into(P, Pa, None, J, K, None, "fp64")
        torch.cuda.synchronize()
        t, nrec = profile_read(plan)
        profile_start(plan, 0)
        ms[mf] = np.asarray(t) / nrec
        st = plan.stats()
        del plan
    per = st["per_combo"]
    rows = []
    for c in range(len(per)):
        nq = int(per[c][0] + per[c][1])
        if nq == 0:
            continue
        f = combo_flops(c, nq, int(per[c][4] + per[c][5]))
        t0 = float(ms[False][2 * c] + ms[False][2 * c + 1])
        t1 = float(ms[True][2 * c] + ms[True][2 * c + 1])
        rows.append({"class": "(%d%d|%d%d)" % combo_moments(c), "quartets": nq, "model_flops": f,
                     "ms_column": t0, "ms_mfma": t1, "frac_best": f / (min(t0, t1) * 1e-3) / 78.6e12})
    return {"serial_build_ms_column": float(ms[False].sum()), "serial_build_ms_mfma": float(ms[True].sum()),
            "per_class": rows}


def water16_scf_graph(eig_tol=0.0):
    """(H2O)16 / STO-3G closed-shell damped + DIIS SCF with one captured
    hipGraph replay per iteration (scf_diis_graph: Fock build, tile products,
    DIIS stage, Jacobi eigensolver with its rotation log, damping/control),
    from the core guess; the same loop with eig_tol = 1e-9 beside it."""
    from .geometry import water_cluster
    from .scf import scf_diis_graph
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    plan = Plan(build_shells(load_basis(), Z, xyz), Z, xyz, schwarz_tau=1e-10)
    nocc = int(sum(Z)) // 2
    out = {}
    for tol, key in ((eig_tol, "exact"), (1e-9, "eig_tol_1e-9")):
        r = scf_diis_graph(plan, Z, xyz, nocc, eig_tol=tol)
        out[key] = {"energy_Eh": r.energy, "converged": r.converged, "iterations": r.iterations,
                    "ms_per_iteration_median": float(np.median(r.iter_times[1:]) * 1e3)}
    return out


WORKLOADS = ("water_rhf", "alkane_spd_mixed", "water_batch_scf", "alkane_gradient", "water16_grid",
             "alkane_spd_class_roofline", "water16_scf_graph")


def main(argv=None):
    """``python -m dft_amd.workloads NAME [NAME ...]``: run the named
    workloads (default: all) and print one JSON line per workload."""
    import json
    import sys
    names = list(argv if argv is not None else sys.argv[1:]) or list(WORKLOADS)
    for name in names:
        if name not in WORKLOADS:
            raise SystemExit(f"unknown workload {name!r}; choose from {', '.join(WORKLOADS)}")
        t0 = time.perf_counter()
        log(f"== {name}")
        r = globals()[name]()
        r["wall_s"] = time.perf_counter() - t0
        print(json.dumps({name: r}, default=str), flush=True)


if __name__ == "__main__":
    main()
