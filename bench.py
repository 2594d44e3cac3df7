"""Benchmark: direct FP64 Fock build of (H2O)16 / STO-3G on one MI355X.

Workload: 16 waters on a 4x2x2 lattice (3.1 A spacing), each rotated by a
uniform random quaternion (seed 0); STO-3G, 112 basis functions, 80 shells;
Schwarz threshold 1e-10. A few plain-SCF iterations give a realistic
density; then ``--steps`` direct J/K builds are timed back to back on the
device (one step = one Fock build = every screened shell quartet evaluated
and digested into J and K), each one replay of the captured hipGraph of the
build, as the device SCF iteration runs it (the same builds issued eagerly,
with an event fork/join per build, are timed beside it). The metric value is unique (8-fold) contracted
Cartesian AO quartets surviving screening per second, counted at shell level
(prod n(l) per screened unique shell quartet); shell and primitive shell
quartets/s are reported beside it; ms_per_step is the Fock-build time per
SCF iteration.

Also reported: the mixed-precision build (FP32 for Q_ab Q_cd < 1e-4) and
its SCF energy difference to FP64, per-launch HIP-event timings with the
roofline of the dominant kernel (FP64 VALU), and the CPU baseline: the C
restatement of the reference's per-primitive-Cartesian McMurchie-Davidson
path (oracle/md_ref.c, OpenMP) on a uniform random sample of the same
screened quartet list, extrapolated to a full build.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6e12  # MI355X dense FP64 (vector and MFMA), FLOP/s
FP32_PEAK = 157.3e12  # MI355X FP32 vector, FLOP/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def screened_quartets(Q, tau):
    """Unique (a>=b, c>=d, ab>=cd) segment quartets with Q_ab Q_cd >= tau."""
    n = Q.shape[0]
    ii, jj = np.tril_indices(n)
    qp = Q[ii, jj]
    bra, ket = np.tril_indices(len(ii))
    keep = qp[bra] * qp[ket] >= tau
    bra, ket = bra[keep], ket[keep]
    return np.stack([ii[bra], jj[bra], ii[ket], jj[ket]], axis=1).astype(np.int32)


TRAFFIC = os.path.join(ROOT, "profiles", "eri_jk_h2o16_sto3g_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this benchmark, each
    scaled by the same counter's reading for a device copy of known size),
    or None when that summary has no entry for it."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def traffic_source():
    """Where ``traffic`` comes from: the committed counter summary and the
    commit of the code it was measured on."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
        return {"file": "profiles/eri_jk_h2o16_sto3g_traffic.json", "measured_at_commit": t.get("commit"),
                "method": t.get("method")}
    except (OSError, ValueError):
        return None


def cpu_baseline(shells, Q, tau, P, seconds, n_quartets_gpu, jk_gpu=None):
    """C restatement timed on the host; when the timed sample is the whole
    screened list (it is for (H2O)16 at the default time budget), its J/K are
    also compared with the HIP FP64 J/K ``jk_gpu`` of the same density."""
    import oracle.ref as ref
    seg = ref.Segments(shells)
    quart = screened_quartets(Q, tau)
    threads = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(quart))
    n_try = min(len(quart), 2000)
    t0 = time.perf_counter()
    ref.build_jk(seg, quart[perm[:n_try]], P, P / 2, None, nthreads=threads)
    t_try = time.perf_counter() - t0
    n = min(len(quart), max(n_try, int(n_try * seconds / max(t_try, 1e-6))))
    t0 = time.perf_counter()
    Jc, Kc, _ = ref.build_jk(seg, quart[perm[:n]], P, P / 2, None, nthreads=threads)
    t = time.perf_counter() - t0
    check = None
    if n == len(quart) and jk_gpu is not None:
        check = {"quartets": int(n), "max_abs_dJ": float(np.abs(jk_gpu[0] - Jc).max()),
                 "max_abs_dK": float(np.abs(jk_gpu[1] - Kc).max()),
                 "note": "HIP FP64 J/K vs the C restatement over the identical screened list"}
    nc = (seg.l + 1) * (seg.l + 2) // 2
    s = quart[perm[:n]]
    n_ao = int((nc[s[:, 0]] * nc[s[:, 1]] * nc[s[:, 2]] * nc[s[:, 3]]).sum())
    rate = n / t
    # one thread on 1/threads of the sample (about the threaded run's time)
    n1 = max(1, n // max(threads, 1))
    t1 = time.perf_counter()
    ref.build_jk(seg, quart[perm[:n1]], P, P / 2, None, nthreads=1)
    t1 = time.perf_counter() - t1
    s1 = quart[perm[:n1]]
    n_ao1 = int((nc[s1[:, 0]] * nc[s1[:, 1]] * nc[s1[:, 2]] * nc[s1[:, 3]]).sum())
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n_ao / t, "unit": "AO quartets/s", "cores": threads, "kind": "port",
            "cpu_model": model, "value_1_thread": n_ao1 / t1,
            "sample_1_thread": f"{n1} quartets of the same sample, {t1:.1f} s",
            "sample": f"{n} of {len(quart)} screened unique shell quartets ({n_ao} AO quartets), uniform random "
                      f"(seed 0), {t:.1f} s; C restatement of the reference per-primitive-Cartesian MD path "
                      f"(oracle/md_ref.c, OpenMP {threads} threads)",
            "shell_quartets_per_s": rate,
            "s_per_fock_build_extrapolated": len(quart) / rate,
            "n_screened_list": int(len(quart)), "gpu_list_matches": int(len(quart)) == n_quartets_gpu,
            "jk_parity": check}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other workloads (C12H26 s/p/d mixed, the 64-molecule batch, the C8H18 "
                         "gradient) reported beside the headline")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--fp64-only", action="store_true",
                    help="only the FP64 timed region (for profiler runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the Fock build shards by quartet blocks with one "
              "all-reduce of J/K per iteration (design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from dft_amd.basis import build_shells, load_basis
    from dft_amd.engine import Graph, Plan, profile_read, profile_start, tune
    from dft_amd.flops import build_flops, combo_flops
    from dft_amd.geometry import water_cluster
    from dft_amd.scf import scf_diis, scf_diis_graph, scf_plain

    tau = 1e-10
    Z, xyz = water_cluster((4, 2, 2), 3.1, seed=0)
    shells = build_shells(load_basis(), Z, xyz)
    log(f"(H2O)16 STO-3G: {shells[-1]['ao_stop']} bf, {len(shells)} shells")
    t0 = time.perf_counter()
    plan = Plan(shells, Z, xyz, schwarz_tau=tau, fp32_tau=1e-4)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    st = plan.stats()
    log(f"plan {t_plan:.2f} s: {st['n_quartets']} screened quartets of {st['n_unscreened']}")

    # realistic density from a short plain SCF (untimed)
    nocc = int(sum(Z)) // 2
    scf = scf_plain(plan, Z, xyz, nocc, nocc, max_count=max(args.warmup, 1) - 1, e_tol=-1, p_tol=-1)
    P_a = scf.P_a.contiguous()
    P = (2 * P_a).contiguous()
    n = plan.nbf
    J = torch.empty((n, n), dtype=torch.float64, device=plan.device)
    K = torch.empty_like(J)

    def run(k, mode):
        for _ in range(k):
            plan.jk_into(P, P_a, None, J, K, None, mode)

    graphs = {}

    def run_graph(k, mode):
        """The build as the captured SCF iteration runs it: one replay of the
        hipGraph of the build (class launches on four streams, reduction,
        symmetrisation; the same kernels and launch shapes as jk_into)."""
        g = graphs.get(mode)
        if g is None:
            g = graphs[mode] = Graph(plan, P, P_a, None, J, K, None, mode)
        for _ in range(k):
            g.launch()

    def time_builds(fn, mode, k):
        fn(args.warmup, mode)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn(k, mode)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    # launch shapes per class combination from the autotuner (untimed)
    shapes, t_tune = None, 0.0
    if not args.no_tune:
        t0 = time.perf_counter()
        shapes = tune(plan, P, P_a, None, reps=3)
        t_tune = time.perf_counter() - t0
        log(f"tuned in {t_tune:.1f} s")

    # warmup, then the timed region (no per-launch events inside it): K
    # replays of the captured build, as the device SCF iteration runs it
    run_graph(args.warmup, "fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    run_graph(args.steps, "fp64")
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / args.steps
    # the same builds issued eagerly (jk_into: class launches forked to the
    # auxiliary streams and joined back by events, per build)
    ms_eager = time_builds(run, "fp64", args.steps)
    # per-build times (outside the timed region: one event pair per build),
    # median over builds 2..K
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    ev[0].record()
    for i in range(args.steps):
        run_graph(1, "fp64")
        ev[i + 1].record()
    torch.cuda.synchronize()
    per_build = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    ms_median = float(np.median(per_build[1:])) if args.steps > 1 else per_build[0]
    quartets = st["n_quartets"]
    value = st["n_ao_quartets"] / (ms_per_step * 1e-3)
    log(f"fp64: {ms_per_step:.3f} ms/build, {value:.3e} AO quartets/s, "
        f"{quartets / (ms_per_step * 1e-3):.3e} shell quartets/s (wall {wall:.3f} s)")

    def roofline():
        """Per-launch HIP-event timings (recorded by the library on the
        launch stream) of args.steps FP64 builds on a second plan of the same
        molecule whose class launches run serially (same kernels and launch
        shapes; in the concurrent timed builds a launch shares the chip with
        its neighbours, so its duration is not its own). Run last, so the
        last args.steps dispatches of each kernel in a trace are these."""
        sp = Plan(shells, Z, xyz, schwarz_tau=tau, fp32_tau=1e-4, serial=True)
        if shapes is not None:
            from dft_amd.engine import set_shapes
            set_shapes(sp, shapes)
        Js, Ks = torch.empty_like(J), torch.empty_like(K)
        for _ in range(2):
            sp.jk_into(P, P_a, None, Js, Ks, None, "fp64")
        profile_start(sp, args.steps)
        for _ in range(args.steps):
            sp.jk_into(P, P_a, None, Js, Ks, None, "fp64")
        torch.cuda.synchronize()
        ms_slots, nrec = profile_read(sp)
        profile_start(sp, 0)
        # dominant kernel: the longest class launch; flops from the
        # algorithmic model, K_a K_b K_c K_d primitive quartets per
        # surviving contracted quartet
        per = st["per_combo"]
        ncombo = len(per)
        eri_ms = ms_slots[:2 * ncombo]
        dom = int(np.argmax(eri_ms))
        combo = dom // 2
        nq_dom = int(per[combo][0] + per[combo][1])
        f_dom = combo_flops(combo, nq_dom, int(per[combo][4] + per[combo][5]))
        f_dom_eval = combo_flops(combo, nq_dom, int(per[combo][2] + per[combo][7]))
        t_dom = eri_ms[dom] / nrec * 1e-3
        achieved = f_dom / t_dom / 1e12
        f_all = sum(build_flops(per))          # whole list in FP64 in this mode
        f_all_eval = sum(build_flops(per, model=False, mode="fp64"))
        from dft_amd.flops import combo_moments
        kname = "eri_jk_kernel<%d, %d, %d, %d, double, 1>" % combo_moments(combo)
        # every class launch: model flops over its serial launch time
        per_class = []
        for c in range(ncombo):
            ms = eri_ms[2 * c] / nrec
            nq_c = int(per[c][0] + per[c][1])
            if nq_c == 0 or ms <= 0:
                continue
            fc = combo_flops(c, nq_c, int(per[c][4] + per[c][5]))
            per_class.append({"class": "(%d%d|%d%d)" % combo_moments(c), "quartets": nq_c,
                              "model_flops": fc, "ms": ms, "tflops": fc / (ms * 1e-3) / 1e12,
                              "frac": fc / (ms * 1e-3) / FP64_PEAK})
        return {"bound": "fp64-valu", "pipe": "compute-bound on the FP64 vector ALUs (no MFMA instruction is "
                                              "issued by this kernel); MI355X's FP64 VALU and FP64 MFMA peaks are "
                                              "both 78.6 TFLOP/s",
                "kernel": kname, "achieved": achieved, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved * 1e12 / FP64_PEAK, "traffic": measured_traffic(kname),
                "traffic_source": traffic_source(),
                "per_class": per_class,
                "avg_launch_ms": eri_ms[dom] / nrec, "launch_timing": "serial class launches",
                "flops_per_launch": f_dom, "flops_per_launch_evaluated": f_dom_eval,
                "whole_build_tflops": f_all / (ms_per_step * 1e-3) / 1e12,
                "whole_build_frac": f_all / (ms_per_step * 1e-3) / FP64_PEAK,
                "whole_build_frac_evaluated": f_all_eval / (ms_per_step * 1e-3) / FP64_PEAK,
                "serial_build_ms": float(sum(ms_slots)) / nrec,
                "slot_ms_per_step": [float(x) / nrec for x in ms_slots]}

    per = st["per_combo"]
    out = {
        "metric": "Fock-build s/iter + ERI quartets/s, FP64 vs mixed at 1 µEh; gradient s/call",
        "value": value, "unit": "AO quartets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "(H2O)16 STO-3G direct J/K: 4x2x2 lattice 3.1 A, seed-0 rotations, "
                               "112 bf, Schwarz 1e-10", "n_ao_quartets": st["n_ao_quartets"],
                   "n_quartets": quartets,
                   "n_unscreened": st["n_unscreened"], "n_prim_quartets": st["n_prim_quartets"],
                   "parallelism": "single GPU"},
        "per_combo": {"columns": ["fp64 quartets", "fp32-eligible quartets", "fp64 prim quartets",
                                  "fp32 prim quartets", "fp64 K-products", "fp32 K-products",
                                  "fp32 level-2 quartets", "fp64 prim quartets of the fp32-eligible part"],
                      "rows": [list(map(int, r)) for r in per]},
        "s_per_iter_fock_fp64": ms_per_step * 1e-3,
        "s_per_iter_fock_fp64_median_2_to_K": ms_median * 1e-3,
        "build_launch": "hipGraph replay of the captured build (the device SCF iteration's path)",
        "s_per_iter_fock_fp64_eager": ms_eager * 1e-3,
        "shell_quartets_per_s": quartets / (ms_per_step * 1e-3),
        "prim_shell_quartets_per_s": st["n_prim_quartets"] / (ms_per_step * 1e-3),
        "plan_seconds": t_plan,
        # device workspace of one Fock build: half matrices, J/K staging slots
        "fock_workspace_bytes": st["workspace_bytes"],
        "tune_seconds": t_tune,
        "tuned_shapes": None if shapes is None else [list(map(int, r)) for r in shapes if r.any()],
        "roofline": None,
        "cpu_baseline": None,
    }
    if args.fp64_only:
        out["roofline"] = roofline()
        print(json.dumps(out), flush=True)
        return

    # mixed precision (graph replays, as the FP64 line; eager beside it)
    ms_mixed = time_builds(run_graph, "mixed", args.steps)
    ms_mixed_eager = time_builds(run, "mixed", args.steps)
    # converged SCF energies in both modes (damped + DIIS loop; the plain
    # Roothaan loop oscillates on this cluster), on an untuned plan of the
    # same molecule so that every FP32-eligible quartet runs in FP32 (the
    # tuned plan may run some classes all-FP64 in mixed mode)
    acc = Plan(shells, Z, xyz, schwarz_tau=tau, fp32_tau=1e-4)
    scf64 = scf_diis(acc, Z, xyz, nocc, nocc, mode="fp64")
    scf32 = scf_diis(acc, Z, xyz, nocc, nocc, mode="mixed")
    # the same loop with each iteration one replay of a captured hipGraph
    scfg = scf_diis_graph(acc, Z, xyz, nocc, mode="fp64")
    # and with the per-iteration eigensolves stopped at off-diagonal entries
    # of 1e-9 (1e-3 of the loop's density tolerance; exact exit eigensolve)
    scfg_tol = scf_diis_graph(acc, Z, xyz, nocc, mode="fp64", eig_tol=1e-9)
    del acc
    f64_part, f32_part = build_flops(per)
    log(f"mixed: {ms_mixed:.3f} ms/build; E64 {scf64.energy:.10f} E_mixed {scf32.energy:.10f}")

    out.update({
        "s_per_iter_fock_mixed": ms_mixed * 1e-3,
        "s_per_iter_fock_mixed_eager": ms_mixed_eager * 1e-3,
        "mixed_speedup": ms_per_step / ms_mixed,
        "mixed_fp32_quartets": st["n_quartets_fp32"],
        # the tuner keeps the FP32 launch of a class only where it pays
        "mixed_fp32_quartets_run": (int(sum(int(per[i][1]) for i in range(len(per)) if shapes[i][3] == 0))
                                    if shapes is not None else st["n_quartets_fp32"]),
        # model flops of the FP64 part at the FP64 peak plus the FP32 part at
        # the FP32 vector peak, over the mixed build time
        "mixed_whole_build_frac": (f64_part / FP64_PEAK + f32_part / FP32_PEAK) / (ms_mixed * 1e-3),
        "scf_energy_fp64": scf64.energy, "scf_energy_mixed": scf32.energy,
        "abs_dE_mixed_Eh": abs(scf64.energy - scf32.energy),
        "scf_iterations_fp64": scf64.iterations,
        "scf_converged": [bool(scf64.converged), bool(scf32.converged)],
        "scf_s_per_iter_wall": float(np.median(scf64.iter_times[1:])) if len(scf64.iter_times) > 1 else None,
        "scf_graph": {"s_per_iter_wall": float(np.median(scfg.iter_times[1:])), "iterations": scfg.iterations,
                      "energy": scfg.energy, "abs_dE_vs_eager_Eh": abs(scfg.energy - scf64.energy),
                      "converged": bool(scfg.converged),
                      "note": "damped + DIIS loop, one hipGraph replay per iteration (J/K, DIIS stage, Jacobi "
                              "eigensolver, density, control on the device; one 32-double read per iteration)"},
        "scf_graph_eig_tol_1e-9": {"s_per_iter_wall": float(np.median(scfg_tol.iter_times[1:])),
                                   "iterations": scfg_tol.iterations, "energy": scfg_tol.energy,
                                   "abs_dE_vs_exact_eigensolves_Eh": abs(scfg_tol.energy - scfg.energy),
                                   "converged": bool(scfg_tol.converged),
                                   "note": "the same loop with the per-iteration Jacobi eigensolves stopped at "
                                           "off-diagonal entries of 1e-9 (dft_syevj_tol)"},
    })
    if not args.no_cpu:
        log("cpu baseline ...")
        plan.jk_into(P, P_a, None, J, K, None, "fp64")
        jk_gpu = (J.cpu().numpy(), K.cpu().numpy())
        out["cpu_baseline"] = cpu_baseline(shells, plan.schwarz(), tau, P.cpu().numpy(), args.cpu_seconds,
                                           quartets, jk_gpu)
        log(f"cpu: {out['cpu_baseline']['value']:.3e} quartets/s on {out['cpu_baseline']['cores']} threads")
    if not args.no_extras:
        out.update(extras())
    out["roofline"] = roofline()
    log(f"dominant kernel {out['roofline']['kernel']}: {out['roofline']['achieved']:.2f} TFLOP/s "
        f"({100 * out['roofline']['frac']:.1f}% of FP64 peak)")
    print(json.dumps(out), flush=True)


def extras():
    """The other benchmark workloads, each measured by its own function of
    dft_amd.workloads (device times from HIP events); a failure is recorded
    in the line instead of ending the run."""
    from dft_amd import workloads as W
    res = {}
    for key, fn in (("config2_c12h26_spd_mixed", lambda: W.alkane_spd_mixed(steps=5, warmup=2)),
                    ("config3_batch64", lambda: W.water_batch_scf(steps=5, warmup=2, scf_check=False)),
                    ("config4_c8h18_gradient", lambda: W.alkane_gradient(fd_coords=4, sd_steps=10))):
        t0 = time.perf_counter()
        log(f"extra workload {key} ...")
        try:
            r = fn()
        except Exception as exc:  # recorded, not fatal: the headline line still prints
            r = {"error": f"{type(exc).__name__}: {exc}"}
        r["wall_s"] = time.perf_counter() - t0
        res[key] = r
        log(f"{key} done in {r['wall_s']:.1f} s")
    b = res.get("config3_batch64", {})
    if "ms_per_iter_batched_mixed_graph_tuned" in b:
        res["north_star_batched_mixed_vs_plain_fp64"] = {
            "ratio": b["speedup_batched_mixed_vs_plain_fp64"], "target": 3.0,
            "ms_plain_fp64": b["ms_per_iter_plain_fp64_per_molecule"],
            "ms_batched_mixed": b["ms_per_iter_batched_mixed_graph_tuned"]}
    return res


if __name__ == "__main__":
    main()
