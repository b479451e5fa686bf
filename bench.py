#!/usr/bin/env python3
"""Benchmark: voxel-iterations/s of the LS-EVPFFT fixed-point loop on one MI355X.

Workload (default): the 128^3 / 512-grain FCC creep RVE (sigma33 = 300 MPa,
n = 10, Voce hardening, Delta-t = 1 s), FP64 form, synthetic seeded
microstructure.  One "step" is one accepted increment: all of its fixed-point
passes (FFT + Green operator, correction, constitutive Newton solve,
reductions, boundary-condition update) plus the accepted-increment updates.
value = voxel-iterations/s = (passes in the timed increments) * V / T.

Also reported: the mixed form (FFT + Green stage in FP32) on the same RVE, the
roofline of the dominant kernel (HIP-event times of the library's own stream),
and the CPU oracle (a C/OpenMP FP64 port of the reference loop) timed on a
bounded sample on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Hardware counters of the timed region of `bench.py --steps 20 --warmup 5` (rocprofv3 --pmc runs
# of the FP64 form; per-pass kernels averaged over their last dispatches, which are the timed
# increments'): HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (scaled by what the same
# counters read for a 1 GiB device copy) and FP64 instruction counts; the file records the
# commands and raw counts.
COUNTER_FILE = ROOT / "profiles" / "counters_fcc128_timed.json"


def measured(stage, key, args):
    """A counter-derived value of `stage` for this workload and timed region, or None."""
    try:
        d = json.loads(COUNTER_FILE.read_text())
    except (OSError, ValueError):
        return None
    e = d.get("stages", {}).get(stage)
    w = d.get("workload", {})
    if not e or (w.get("config"), w.get("n"), w.get("grains"), w.get("steps"), w.get("warmup")) != \
            (args.config, args.n, args.grains, args.steps, args.warmup):
        return None
    return e.get(key)


# algorithmic bytes per voxel-iteration of the FFT + Green stage: read the 9-component
# FP64 PK1 field once, write the 9-component FP64 correction once, plus 2P-2 = 4
# half-spectrum passes of 9 complex values (P = 3 when one component's complex xy slab fits
# the LDS; r = (n/2+1)/n).
def fft_alg_bytes(n, sc=16):
    r = (n // 2 + 1) / n
    return 144.0 + 4 * 9 * sc * r


def flops_constitutive(stats, ns):
    """FP64 flop model of the constitutive stage (evaluations, Newton steps, extra
    line-search evaluations, hardening calls)."""
    ev = stats.evals
    nr = stats.nr_steps
    outer = stats.outer
    per_eval = 100 * ns
    return nr * 932 + ev * per_eval + (ev - nr) * 198 + outer * (2 * ns * ns + 10 * ns + 20)


STAGE_STEPS = 3  # increments after the timed region that time every stage


def run_gpu(problem, precision, steps, warmup, graphs=False):
    """warm-up increments, then `steps` timed increments (wall clock; the constitutive kernel's own
    launch events, cg_set_timing mode 2, which adds nothing between kernels), then STAGE_STEPS
    increments with every stage timed by stream-marker events (mode 1) for the stage breakdown.
    Returns (timed rows, timed stats, wall seconds, stage stats)."""
    import torch

    from creep_modeling_amd import driver

    torch.cuda.init()
    s = driver.Solver(problem, precision=precision, scan_cond=True, graphs=graphs)
    lib = s.L
    for i in range(warmup):
        r = s.step()
        print("warm-up increment %d: %d passes" % (i + 1, r["passes"]), file=sys.stderr, flush=True)
    lib.cg_set_timing(s.h, 2)
    s.ctx.stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = []
    for i in range(steps):
        rows.append(s.step())
        print("timed increment %d: %d passes" % (i + 1, rows[-1]["passes"]), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = s.ctx.stats()
    lib.cg_set_timing(s.h, 1)
    s.ctx.stats(reset=True)
    for i in range(STAGE_STEPS):
        r = s.step()
        print("stage-timing increment %d: %d passes" % (i + 1, r["passes"]), file=sys.stderr, flush=True)
    st_stage = s.ctx.stats()
    s.close()
    return rows, st, t1 - t0, st_stage


def pass_costs(problem, steps, warmup):
    """Per-pass cost of the constitutive kernel against its work: after `warmup` increments, every
    pass of `steps` increments is timed with the kernel's own launch events and paired with the
    device counters of that pass (Newton steps and outer hardening iterations per voxel); prints the
    mean time per work class and the least-squares split t = a + b * NR + c * outer."""
    import numpy as np
    import torch

    from creep_modeling_amd import driver

    torch.cuda.init()
    s = driver.Solver(problem, precision=0, scan_cond=True)
    for i in range(warmup):
        s.step()
    s.L.cg_set_timing(s.h, 2)
    rows = []
    for i in range(steps):
        assert s.begin()
        s.ctx.stats(reset=True)
        while True:
            s._chk(s.L.cg_iterate(s.h, C.byref(s.P), C.byref(s.st), C.byref(s.out)), "cg_iterate")
            st = s.ctx.stats(reset=True)
            rows.append((st.t_ms[1], st.nr_steps / st.voxel_iters, st.outer / st.voxel_iters))
            if s.pass_done():
                break
        s.finish()
        print("increment %d: %d passes" % (warmup + i + 1, s.passes), file=sys.stderr, flush=True)
    s.close()
    R = np.array(rows)
    classes = {}
    for t, nr, ou in R:
        classes.setdefault((round(nr, 2), round(ou, 2)), []).append(t)
    fit = np.linalg.lstsq(np.c_[np.ones(len(R)), R[:, 1], R[:, 2]], R[:, 0], rcond=None)[0]
    out = {"passes": len(R), "fit_ms": {"a": fit[0], "per_newton_step": fit[1], "per_outer_iteration": fit[2]},
           "classes": [{"nr_per_voxel": k[0], "outer_per_voxel": k[1], "passes": len(v), "solve_ms": float(np.mean(v))}
                       for k, v in sorted(classes.items(), key=lambda kv: -len(kv[1]))[:12]]}
    print(json.dumps(out))


def host_cpu():
    """The host's processor as lscpu reports it (model, sockets, cores per socket, threads per core)."""
    info = {}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "cpus_online": info.get("CPU(s)"), "cpus_usable": affinity}


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(problem, max_passes):
    """The C/OpenMP oracle (a port of the reference loop) on the first increment of the same RVE,
    capped at max_passes fixed-point passes (accepted at itmax)."""
    from oracle import bind

    p = problem.processes[0]
    saved = (p.itmax, p.tdotmin)
    p.itmax = max_passes
    p.tdotmin = p.tdot  # accept at itmax
    threads = omp_threads()
    o = bind.Oracle(problem, nthreads=threads)
    t0 = time.perf_counter()
    r = o.step()
    t1 = time.perf_counter()
    o.close()
    p.itmax, p.tdotmin = saved
    return dict(value=r.passes * problem.nvox / (t1 - t0), unit="voxel-iters/s", cores=threads, kind="port",
                sample="first increment of the %dx%dx%d RVE capped at %d fixed-point passes (incl. the increment's "
                       "setup and accepted-step updates), %.1f s on %d OpenMP threads" % (*problem.n, r.passes, t1 - t0,
                                                                                      threads),
                host=host_cpu())


def cpu_config0(args):
    """The CPU baseline plan's full run: the oracle on the 32^3 / 64-grain FCC creep RVE, all 50
    increments of 1 s (sigma33 = 300 MPa), on OMP_NUM_THREADS threads."""
    from creep_modeling_amd import inputs, parse
    from oracle import bind

    ms, mat, procs = inputs.config_case(0)
    problem = parse.problem_from_case(ms, mat, procs)
    threads = omp_threads()
    o = bind.Oracle(problem, nthreads=threads)
    t0 = time.perf_counter()
    passes = incr = 0
    while True:
        r = o.step()
        if r is None:
            break
        passes += r.passes
        incr += 1
        print("increment %d: %d passes, %.1f s" % (incr, r.passes, time.perf_counter() - t0), file=sys.stderr,
              flush=True)
    T = time.perf_counter() - t0
    o.close()
    print(json.dumps({"metric": "voxel-iters/s, CPU oracle (C/OpenMP port of the reference loop), 32^3 / 64 grains "
                                "FCC creep, 50 increments", "value": passes * problem.nvox / T, "unit": "voxel-iters/s",
                      "rve_increments_per_s": incr / T, "passes": passes, "increments": incr, "seconds": T,
                      "threads": threads, "host": host_cpu()}))


def checksums():
    """sha256 of the microstructure files (the reference's .dat text) of every workload."""
    import hashlib

    from creep_modeling_amd import inputs

    def digest(ms):
        n1, n2, n3 = ms.n
        h = hashlib.sha256()
        V = n1 * n2 * n3
        for k0 in range(0, n3):  # one z-plane at a time
            lines = []
            for v in range(k0 * n1 * n2, (k0 + 1) * n1 * n2):
                e = ms.euler[v]
                lines.append("%.6f %.6f %.6f %d %d %d %d %d\n" % (e[0], e[1], e[2], v % n1 + 1, (v // n1) % n2 + 1,
                                                                   v // (n1 * n2) + 1, ms.grain[v], ms.phase[v]))
            h.update("".join(lines).encode())
        assert V > 0
        return h.hexdigest()

    out = {}
    for cfg, kw in ((0, {}), (1, {}), (2, {"n": 256, "ngrains": 4096}), (3, {})):
        ms, _, _ = inputs.config_case(cfg, **kw)
        out["config%d %dx%dx%d/%d grains" % (cfg, *ms.n, int(ms.grain.max()))] = digest(ms)
        print(cfg, file=sys.stderr, flush=True)
    sweep = [digest(c[0]) for _, c in inputs.sweep_cases(64, nsteps=1)]
    out["config4 64 RVEs (sha256 of the 64 digests in order)"] = hashlib.sha256("".join(sweep).encode()).hexdigest()
    print(json.dumps(out, indent=1))


def bench_sweep(args):
    """The batched sweep: `rves` independent FCC RVEs, `warmup + steps` increments each, all
    contexts created (and their microstructures uploaded) before the timed region; the timed
    region runs every RVE's increments with `workers` RVEs in flight."""
    import torch

    from creep_modeling_amd import batch, driver, inputs, parse

    cases = inputs.sweep_cases(args.rves, nsteps=args.warmup + args.steps)
    problems = [parse.problem_from_case(*c) for _, c in cases]
    torch.cuda.init()
    for prec in ((0,) if args.no_mixed else (0, 1)):
        solvers = [driver.Solver(p, precision=prec, scan_cond=False) for p in problems]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rows = batch.run_batch(problems, workers=args.workers, solvers=solvers, engine=args.engine)
        torch.cuda.synchronize()
        T = time.perf_counter() - t0
        vi = sum(sum(r["passes"] for r in rr) * p.nvox for rr, p in zip(rows, problems))
        print("sweep precision %d: %.2f s" % (prec, T), file=sys.stderr, flush=True)
        if prec == 0:
            out = {
                "metric": "voxel-iters/s, batched sweep of independent RVEs (aggregate)",
                "value": vi / T, "unit": "voxel-iters/s", "n_gpus": 1,
                # the whole sweep is timed: every increment of every RVE is a timed step
                "steps": args.warmup + args.steps, "warmup": 0,
                "ms_per_step": T / (args.warmup + args.steps) * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64",
                "data": "synthetic (seeded periodic Voronoi RVEs, uniform SO(3) orientations)",
                "config": {"workload": "%d FCC RVEs, edges in {32,48,64,96}, 32-512 grains, sigma33 200-400 MPa, "
                                       "n 8-14, Arrhenius-scaled gamma0 (475/525/575 C); %d increments each; %s; "
                                       "step = one increment of every RVE"
                                       % (args.rves, args.warmup + args.steps,
                                          "one constitutive launch per round over all RVEs" if args.engine == "shared"
                                          else "%d RVEs in flight on host threads" % args.workers),
                           "voxels": sum(p.nvox for p in problems),
                           "passes": sum(sum(r["passes"] for r in rr) for rr in rows)},
            }
        else:
            out["mixed"] = {"value": vi / T, "ratio_vs_fp64": (vi / T) / out["value"]}
    print(json.dumps(out))


def copy_calibration(gib=1.0, reps=5):
    """A plain device copy of known size (torch's device-to-device copy, which runs the HIP
    runtime's copy kernel), to scale the FETCH_SIZE / WRITE_SIZE counters: each launch
    reads and writes exactly `gib` GiB."""
    import torch

    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": reps,
                      "gbs_read_plus_write": 2 * n * 8 / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3, 4),
                    help="1: FCC n=10 (the headline workload); 2: FCC n=15; 3: BCC 48 systems, biaxial, dt=0.5 s; "
                         "4: batched sweep of independent FCC RVEs (32^3-96^3)")
    ap.add_argument("--rves", type=int, default=64, help="config 4: number of RVEs in the sweep")
    ap.add_argument("--workers", type=int, default=8, help="config 4, thread engine: RVEs in flight at a time")
    ap.add_argument("--engine", default="shared", choices=("shared", "threads"),
                    help="config 4: shared constitutive launch over all RVEs (default) or a host-thread pool")
    ap.add_argument("--n", type=int, default=None, help="grid edge (default 128 for configs 1-2, 256 for 3)")
    ap.add_argument("--grains", type=int, default=None, help="Voronoi grains (default 512, 2048 for config 3)")
    ap.add_argument("--s33", type=float, default=None, help="creep stress sigma33 in MPa (config default 300)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--graphs", action="store_true", help="replay each pass' fixed device work as HIP graphs")
    ap.add_argument("--cpu-passes", type=int, default=8,
                    help="fixed-point passes of the CPU oracle's bounded sample (first increment)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--cpu-config0", action="store_true",
                    help="run only the CPU oracle on the 32^3 / 64-grain RVE, all 50 increments, and exit")
    ap.add_argument("--checksums", action="store_true", help="print the sha256 of every workload's microstructure")
    ap.add_argument("--pass-costs", action="store_true",
                    help="time every pass of the constitutive kernel against its Newton and hardening work and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_config0:
        cpu_config0(args)
        return
    if args.checksums:
        checksums()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the multi-GPU layout (independent RVEs per GPU, or z-slab "
              "decomposition with two all-to-alls per pass) is a design, not part of this build", file=sys.stderr)
        sys.exit(2)

    from creep_modeling_amd import inputs, parse

    if args.config == 4:
        bench_sweep(args)
        return
    if args.n is None:
        args.n = 256 if args.config == 3 else 128
    if args.grains is None:
        args.grains = 2048 if args.config == 3 else 512
    ms, mat, procs = inputs.config_case(args.config, n=args.n, ngrains=args.grains,
                                        nsteps=args.warmup + args.steps + STAGE_STEPS + 1, s33=args.s33)
    problem = parse.problem_from_case(ms, mat, procs)
    pr0 = procs[0]
    V = problem.nvox
    ns = problem.phases[0].nsyst

    if args.pass_costs:
        pass_costs(problem, args.steps, args.warmup)
        return
    rows, st, T, sst = run_gpu(problem, 0, args.steps, args.warmup, args.graphs)
    passes = sum(r["passes"] for r in rows)
    value = passes * V / T
    ms_per_step = T / args.steps * 1e3
    t_fft = sst.t_ms[0] / max(sst.passes, 1)
    t_solve = st.t_ms[1] / max(st.passes, 1)
    fft_bytes = fft_alg_bytes(args.n) * V
    fft_gbs = fft_bytes / (t_fft * 1e-3) / 1e9
    flops = flops_constitutive(st, ns)
    solve_tflops = flops / (st.t_ms[1] * 1e-3) / 1e12 if st.t_ms[1] > 0 else 0.0
    dominant_solve = t_solve >= t_fft
    tr = measured("k_correct_solve", "bytes_per_launch", args)
    if dominant_solve:
        # the constitutive kernel runs on the FP64 vector ALU (no matrix-core instruction): its peak is
        # the FP64 vector rate; achieved = the flop model (counted Newton steps, evaluations and
        # hardening calls) / the kernel's average HIP-event duration
        roof = dict(bound="fp64_valu", kernel="k_correct_solve (constitutive Newton, FP64 vector ALU)",
                    achieved=solve_tflops, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=solve_tflops / FP64_PEAK_TFLOPS,
                    traffic=measured("k_correct_solve", "bytes_per_launch", args), avg_ms=t_solve)
        cf = measured("k_correct_solve", "fp64_flops_per_launch", args)
        # FP64 instructions counted by the hardware (64 lanes per issued instruction, FMA = 2) over
        # the same launches, against the live average duration
        roof["counter_fp64_tflops"] = cf / (t_solve * 1e-3) / 1e12 if cf else None
    else:
        roof = dict(bound="hbm", kernel="FFT + Green stage", achieved=fft_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=fft_gbs / HBM_PEAK_GBS, traffic=measured("fft_green", "bytes_per_launch", args),
                    avg_ms=t_fft)
    out = {
        "metric": "voxel-iters/s at %d^3 creep RVE (mixed vs FP64 GPU) + %% HBM roofline" % args.n,
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded periodic Voronoi RVE, uniform SO(3) orientations)",
        "config": {"workload": "%d^3 RVE, %d grains, %s %d systems, n=%g, %screep, FP64 form; step = one "
                               "accepted increment" % (args.n, args.grains, "BCC" if args.config == 3 else "FCC", ns,
                                                       problem.phases[0].nrs[0],
                                                       ("sigma11=%g " % pr0.s11 if pr0.s11 else "") +
                                                       "sigma33=%g MPa " % pr0.s33),
                   "passes": passes, "voxels": V, "rve_increments_per_s": args.steps / T},
        "roofline": roof,
        "stages_ms_per_pass": {"fft_green": t_fft, "correct_solve": t_solve,
                               "sums_copy": sst.t_ms[2] / max(sst.passes, 1),
                               "per_increment_updates_ms": st.t_ms[3] / max(args.steps, 1),
                               "timing": "correct_solve and per-increment updates: the timed increments "
                                         "(events attached to the solve's launches); fft_green: events attached "
                                         "to the stage's first and last kernels, sums_copy: stream markers, over "
                                         "the %d increments after them (%d passes)" % (STAGE_STEPS, sst.passes)},
        "fft_green_roofline": {"bound": "hbm", "achieved": fft_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": fft_gbs / HBM_PEAK_GBS, "alg_bytes_per_voxel": fft_alg_bytes(args.n),
                               "traffic": measured("fft_green", "bytes_per_launch", args)},
        "constitutive": {"tflops": solve_tflops, "frac_fp64_peak": solve_tflops / FP64_PEAK_TFLOPS,
                         # the same kernel seen from HBM: the counted traffic of its timed launches
                         # (FETCH_SIZE / WRITE_SIZE, scaled) over the live average duration
                         "hbm_gbs_counted": (tr / (t_solve * 1e-3) / 1e9) if tr else None,
                         "hbm_frac_counted": (tr / (t_solve * 1e-3) / 1e9 / HBM_PEAK_GBS) if tr else None,
                         "nr_steps_per_voxel_iter": st.nr_steps / max(st.voxel_iters, 1),
                         "evals_per_voxel_iter": st.evals / max(st.voxel_iters, 1),
                         "pow_per_voxel_iter": st.evals * ns / max(st.voxel_iters, 1)},
    }
    if not args.no_mixed:
        rows_m, st_m, T_m, sst_m = run_gpu(problem, 1, args.steps, args.warmup, args.graphs)
        passes_m = sum(r["passes"] for r in rows_m)
        vm = passes_m * V / T_m
        t_fft_m = sst_m.t_ms[0] / max(sst_m.passes, 1)
        out["mixed"] = {"value": vm, "ratio_vs_fp64": vm / value, "passes": passes_m,
                        "fft_green_ms_per_pass": t_fft_m,
                        "fft_green_gbs_alg": fft_alg_bytes(args.n, 8) * V / (t_fft_m * 1e-3) / 1e9,
                        "creep_rate_rel_err_vs_fp64": abs(rows_m[-1]["edotpavg"][2] - rows[-1]["edotpavg"][2])
                        / abs(rows[-1]["edotpavg"][2])}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(problem, args.cpu_passes)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
