"""Benchmark: force + induction evaluations per second of the 36,864-site water box.

Workload: 12,288 rigid 3-site waters (36,864 sites) in a 4.95 nm cubic box, PME with a 1.0 nm
cutoff, a 64^3 grid and B-splines of order 5 (the Ewald option: real space within the cutoff,
smooth PME in reciprocal space, Ewald parameter from the error tolerance 1e-4), Mutual polarization
solved by PCG to an RMS dipole update of 1e-6 D (2.08e-8 e*nm).  One step = one complete
evaluation (neighbour tiles, permanent field with the stored real-space blocks of T, reciprocal
space, induced-dipole solve, energy, forces) with positions and forces resident in HBM
(tdp_execute_device).

The reported value is the mixed-precision path (real-space pairs beyond the Thole range in FP32,
near pairs, reciprocal space, sums and PCG norms in FP64); the same evaluation fully in FP64 is
timed beside it.  The roofline entry covers the T*mu kernel of the solve (k_ew_spmv, the stored
real-space blocks applied to the dipoles), timed with HIP events on the context's stream over a
timed region of its own; it is memory bound (ceiling: MI355X's 8 TB/s).  The CPU baseline is the
oracle's Ewald evaluation (C, FP64, cell lists, SPME) of the same system on the host's threads.

Also measured: the same box with the reference's PME semantics (minimum image over all pairs,
no cutoff), the 96,000-site saline box in both models, a warm-started trajectory, the batched
cluster solve and the dense cluster responses.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# All-pairs (reference semantics) T*mu pass: per unordered pair (N(N-1)/2 of them) the minimum-image
# separation and the damped rr3/rr5 once (3 sub + 15 min-image + 5 r^2 + 1 rsqrt + 5 powers + 7
# damping-test = 36 flops) and both directions accumulated (2 x 18 = 36 flops): 72 flops.
FLOPS_PER_PAIR_MATVEC = 72.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
DEBYE = 0.020819434  # e*nm
# HBM bytes per launch of k_sym_matvec_mixed (all-pairs model) on this box, from rocprofv3 FETCH_SIZE
# and WRITE_SIZE counters scaled by a 512 MiB copy probe (fetch x2, write x1), default schedule: 98.8 MB
# read + 166.8 MB written (the FP32 tile partial slab, 128 MB, and the FP32 work-item partials).
# Counters cannot be read inside this process.
MATVEC_HBM_BYTES = 265.6e6
# HBM bytes per launch of k_ew_spmv<true> on this box (64^3 grid) from the same counters and scaling:
# 546.7 MB read (the stored lists' jagged rows rounded to whole lines, 520.0 MB, the lists' slot tables
# and slot data of the LDS copies that miss L2) + 9.4 MB written.
SPMV_HBM_BYTES = 556.1e6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-clusters", action="store_true", help="skip the batched cluster-solve measurement")
    p.add_argument("--clusters-only", action="store_true",
                   help="run only the cluster-solve measurement and print its object (for profiling)")
    p.add_argument("--no-dense", action="store_true", help="skip the dense cluster-response measurement")
    p.add_argument("--no-saline", action="store_true", help="skip the 96,000-site saline box measurement")
    p.add_argument("--no-md", action="store_true", help="skip the warm-started trajectory measurement")
    p.add_argument("--no-allpairs", action="store_true",
                   help="skip the all-pairs (reference PME semantics) measurement of the 36,864-site box")
    p.add_argument("--pmode-only", action="store_true",
                   help="run only the Ewald evaluation of the 36,864-site box (mixed and FP64) and print its line")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep each context's default schedule (no tdp_autotune before the warm-up)")
    p.add_argument("--dense-only", action="store_true",
                   help="run only the dense cluster-response measurement and print its object (for profiling)")
    p.add_argument("--cpu-seconds", type=float, default=4.0,
                   help="approximate CPU time budget per pass sample of the all-pairs CPU baseline")
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="after the run, stream this many MiB through a plain device copy (y = x * 1 in float32) "
                        "three times: the reference kernel for scaling HBM byte counters")
    return p.parse_args()


def build_system():
    from tholedipoleplugin_amd import systems
    return systems.config1()


def make_handle(s, precision, ewald=False):
    from tholedipoleplugin_amd import _lib
    return _lib.HipContextHandle(s.force, s.box, precision=precision, solver="pcg", device=0,
                                 flags=_lib.FLAG_EWALD if ewald else 0)


def time_path(torch, s, precision, steps, warmup, timing=False, ewald=False):
    h = make_handle(s, precision, ewald)
    # schedule chosen on this configuration, untimed (the all-pairs passes' work-item length)
    tiles_per_item = 0 if (NO_AUTOTUNE or ewald) else h.autotune(s.positions, s.box)
    dev = torch.device("cuda:0")
    pos = torch.tensor(s.positions.reshape(-1), dtype=torch.float64, device=dev)
    frc = torch.zeros(pos.numel(), dtype=torch.float64, device=dev)
    en = torch.zeros(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    for _ in range(warmup):
        h.execute_device(pos.data_ptr(), s.box, frc.data_ptr(), en.data_ptr(), stream)
    torch.cuda.synchronize()
    if timing:
        h.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.execute_device(pos.data_ptr(), s.box, frc.data_ptr(), en.data_ptr(), stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = h.solver_stats()
    if timing:
        h.set_timing(False)
    # one extra evaluation through the host API for the accuracy comparison
    from tholedipoleplugin_amd import _lib
    e, f = h.execute(s.positions, s.box)
    mu = h.dipoles(s.positions, s.box, _lib.DIPOLES_INDUCED)
    info = h.ewald_info() if ewald else None
    h.close()
    stats["tiles_per_item"] = tiles_per_item
    return dict(seconds=dt, ms_per_step=1000.0 * dt / steps, stats=stats, energy=e, forces=f, mu=mu, info=info)


def cpu_baseline(s, mu, iterations, budget_s):
    """Oracle (C restatement of the reference loops) on bounded slices of each pass; the per-eval
    time is extrapolated as t_field + iterations * t_matvec + t_electrostatic."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n = s.n
    total_pairs = n * (n - 1) / 2.0
    threads = cpu_threads()
    # calibrate the slice size on the first pass, then spread slices over the row range
    rates = {}
    for which in (0, 1, 2):
        secs, pairs = oracle.time_pass_slices(s.force, s.positions, mu, s.box, which, [(0, 64)], threads)
        rate = pairs / max(secs, 1e-9)
        target_pairs = rate * budget_s / 3.0
        rows_needed = min(n, max(64, int(target_pairs / n)))   # at most one full pass
        starts = np.linspace(0, n - rows_needed // 8 - 1, 8).astype(int)
        slices = [(int(a), int(a) + rows_needed // 8) for a in starts]
        secs, pairs = oracle.time_pass_slices(s.force, s.positions, mu, s.box, which, slices, threads)
        rates[which] = pairs / secs
    t_field = total_pairs / rates[0]
    t_matvec = total_pairs / rates[1]
    t_elec = total_pairs / rates[2]
    t_eval = t_field + iterations * t_matvec + t_elec
    return dict(value=1.0 / t_eval, unit="evals/s", cores=threads, kind="port", host_cpus=os.cpu_count(),
                sample=("oracle pair passes of the 36,864-site box (all pairs, minimum image), each timed on 8 "
                        "slices of rows (about %.0f s of work, at most one full pass): per-pass times %.1f s field, "
                        "%.1f s T*mu, "
                        "%.1f s energy/force, per eval t_field + %d t_T*mu + t_energy (the GPU PCG pass count; the "
                        "reference's Jacobi iteration diverges on this system)"
                        % (budget_s / 3.0, t_field, t_matvec, t_elec, iterations)))


def cpu_threads():
    """Host threads for the CPU baselines: every CPU the process may use, at most 16 (the share of one
    GPU on the GPU hosts)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_ewald(s, alpha, grid):
    """The oracle's Ewald evaluation (FP64 C: cell-list real space, SPME of order 5, PCG to the same
    tolerance) of the whole system, once, on the host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = cpu_threads()
    t0 = time.perf_counter()
    r = oracle.ewald_evaluate(s.force, s.positions, s.box, s.force.getCutoffDistance(), alpha, grid=grid, order=5,
                              threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=1.0 / dt, unit="evals/s", cores=threads, kind="port", host_cpus=os.cpu_count(),
                sample="one full evaluation of the 36,864-site box by the oracle (FP64; cell-list real space, SPME "
                       "%dx%dx%d order 5, PCG %d iterations to 1e-6 D): %.1f s on %d threads"
                       % (grid[0], grid[1], grid[2], r["iterations"], dt, threads),
                energy=r["energy"], forces=r["forces"], induced=r["induced"])


def cluster_systems():
    """512 non-periodic water clusters, 20-400 molecules (60-1,200 sites) drawn log-uniformly, liquid
    density, Mutual induction to 1e-7 D (2.08e-9 e*nm)."""
    from tholedipoleplugin_amd import systems
    sizes = systems.cluster_sizes_config3()
    out = []
    for k, m in enumerate(sizes):
        c = systems.water_cluster(int(m), seed=20261206 + 1 + k)
        c.force.setMutualInducedTargetEpsilon(1e-7 * DEBYE)
        c.force.setMutualInducedMaxIterations(1000)
        out.append(c)
    return out


def cluster_bench(torch, steps, warmup):
    """Cluster solves per second: the whole set in one batched evaluation (mixed and FP64) against
    one FP64 context per cluster evaluated in turn.  A solve is a full evaluation of a cluster
    (field, induced dipoles to 1e-7 D, energy, forces)."""
    from tholedipoleplugin_amd import _lib
    cl = cluster_systems()
    sizes = [c.n for c in cl]
    dev = torch.device("cuda:0")
    pos_h = np.concatenate([c.positions for c in cl])
    pos = torch.tensor(pos_h.reshape(-1), dtype=torch.float64, device=dev)
    frc = torch.zeros(pos.numel(), dtype=torch.float64, device=dev)
    en = torch.zeros(len(cl), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    res = {}
    for prec in ("mixed", "double"):
        b = _lib.HipBatchHandle([c.force for c in cl], precision=prec, solver="pcg")
        if not NO_AUTOTUNE:
            b.autotune(pos_h)
        tiles_per_item = b.work_schedule()[0]   # per size bucket: 1-2, 3-4, 5-8, 9-16, 17+ tiles
        for _ in range(warmup):
            b.execute_device(pos.data_ptr(), frc.data_ptr(), en.data_ptr(), stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.execute_device(pos.data_ptr(), frc.data_ptr(), en.data_ptr(), stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        iters = [b.system_stats(g)["iterations"] for g in range(len(cl))]
        e, f = b.execute(pos_h)
        mu = b.dipoles(pos_h, None, _lib.DIPOLES_INDUCED)
        b.close()
        res[prec] = dict(seconds=dt, iters=iters, energy=e, forces=f, mu=mu, tiles_per_item=tiles_per_item)
        print("# clusters batched %s: %.2f ms per set of %d, PCG iterations %d..%d, tiles per work item by size "
              "bucket %s" % (
            prec, dt * 1e3, len(cl), min(iters), max(iters), tiles_per_item), file=sys.stderr, flush=True)
    # one FP64 context per cluster, evaluated in turn on device buffers
    handles = [_lib.HipContextHandle(c.force, np.zeros(9), precision="double", solver="pcg") for c in cl]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    def loop():
        for k, h in enumerate(handles):
            h.execute_device(pos.data_ptr() + 24 * int(offs[k]), None, frc.data_ptr() + 24 * int(offs[k]),
                             en.data_ptr() + 8 * k, stream)
    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop()
    torch.cuda.synchronize()
    dt_loop = time.perf_counter() - t0
    for h in handles:
        h.close()
    print("# clusters one FP64 context each: %.2f ms per set" % (dt_loop * 1e3), file=sys.stderr, flush=True)
    m, d = res["mixed"], res["double"]
    ncl = len(cl)
    return {
        "workload": "512 water clusters, 20-400 molecules (60-1,200 sites, %d in all), NoCutoff all pairs, "
                    "Mutual PCG to 1e-7 D" % int(offs[-1]),
        "unit": "cluster solves/s",
        "batched_mixed": ncl / m["seconds"],
        "batched_fp64": ncl / d["seconds"],
        "per_cluster_fp64_loop": ncl / dt_loop,
        "batched_mixed_vs_loop": dt_loop / m["seconds"],
        "batched_fp64_vs_loop": dt_loop / d["seconds"],
        "mixed_vs_fp64_batched": d["seconds"] / m["seconds"],
        "mixed_force_rel_rms_err_vs_fp64": float(np.sqrt(np.mean((m["forces"] - d["forces"]) ** 2)) /
                                                 np.sqrt(np.mean(d["forces"] ** 2))),
        "mixed_mu_max_err_vs_fp64_debye": float(np.max(np.abs(m["mu"] - d["mu"]))) / DEBYE,
        "pcg_iterations_max": int(max(d["iters"])),
        "tiles_per_work_item_by_size_bucket": {"buckets_tiles": ["1-2", "3-4", "5-8", "9-16", "17+"],
                                               "mixed": res["mixed"]["tiles_per_item"],
                                               "fp64": res["double"]["tiles_per_item"]},
    }


def dense_systems():
    """256 non-periodic water clusters of 17-333 molecules (3N = 153-2,997 unknowns) drawn uniformly."""
    from tholedipoleplugin_amd import systems
    sizes = systems.cluster_sizes_config4()
    return [systems.water_cluster(int(m), seed=20261207 + 1 + k) for k, m in enumerate(sizes)]


def dense_bench(steps):
    """Linear responses (E0 and unit fields along x, y, z) of 256 clusters to a relative residual
    of 1e-10: 16-bit Cholesky on MFMA + FP64 refinement (FP16 and BF16) against batched FP64 CG.
    A solve is the four responses of one cluster; timing covers the whole call (positions upload,
    permanent field, assembly, factorization, refinement, download)."""
    from tholedipoleplugin_amd import _lib
    cl = dense_systems()
    pos = np.concatenate([c.positions for c in cl])
    b = _lib.HipBatchHandle([c.force for c in cl], precision="double", solver="pcg")
    res = {}
    for method in ("fp16", "bf16", "cg"):
        b.responses(pos, method=method, tolerance=1e-10, max_steps=1000)
        t0 = time.perf_counter()
        for _ in range(steps):
            X, st = b.responses(pos, method=method, tolerance=1e-10, max_steps=1000)
        dt = (time.perf_counter() - t0) / steps
        res[method] = dict(seconds=dt, X=X, stats=st)
        print("# dense %s: %.2f ms per set of %d, %s" % (method, dt * 1e3, len(cl), st), file=sys.stderr, flush=True)
    b.close()
    unknowns = np.array([3 * c.n for c in cl], dtype=np.float64)
    flops = float(np.sum(unknowns ** 3) / 3.0)
    fac_ms = res["fp16"]["stats"]["factor_ms"]
    cg = res["cg"]
    out = {
        "workload": "256 water clusters, 17-333 molecules (3N = 153-2,997 unknowns), 4 fields (E0, unit x, y, z), "
                    "relative residual 1e-10",
        "unit": "cluster solves/s",
        "dense_fp16": len(cl) / res["fp16"]["seconds"],
        "dense_bf16": len(cl) / res["bf16"]["seconds"],
        "batched_fp64_cg": len(cl) / cg["seconds"],
        "dense_fp16_vs_cg": cg["seconds"] / res["fp16"]["seconds"],
        "refinement_steps_fp16": res["fp16"]["stats"]["steps"],
        "refinement_steps_bf16": res["bf16"]["stats"]["steps"],
        "cg_iterations_max": cg["stats"]["steps"],
        "fp16_max_err_vs_cg_debye": float(np.max(np.abs(res["fp16"]["X"] - cg["X"]))) / DEBYE,
        "bf16_max_err_vs_cg_debye": float(np.max(np.abs(res["bf16"]["X"] - cg["X"]))) / DEBYE,
        "factor_roofline": {"bound": "mfma", "achieved": flops / (fac_ms * 1e-3) / 1e12,
                            "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                            "frac": flops / (fac_ms * 1e-3) / 1e12 / PEAK_FP16_TFLOPS,
                            "flops": flops, "factor_ms": fac_ms, "assemble_ms": res["fp16"]["stats"]["assemble_ms"]},
    }
    return out


def saline_bench(torch, steps):
    """96,000-site box of water with 0.15 M Na+/Cl- (7.1 nm, Thole a in [0.3, 0.45]), Mutual PCG to
    1e-6 D, with the Ewald option (cutoff 1.0 nm, 96^3 grid, order 5) and with the reference's
    all-pairs semantics: mixed and FP64 evaluations per second and the mixed result's error against
    FP64."""
    from tholedipoleplugin_amd import systems
    s = systems.config2()
    out = {"workload": "96,000 sites: 31,978 waters + 33 Na+/Cl- pairs, 7.1 nm box, Thole a in [0.3, 0.45], "
                       "Mutual PCG to 1e-6 D", "unit": "evals/s"}
    for model, ewald in (("ewald_cutoff1.0_grid96", True), ("allpairs_min_image", False)):
        mixed = time_path(torch, s, "mixed", steps, 2, ewald=ewald)
        print("# saline96000 %s mixed: %.2f ms/step, %s" % (model, mixed["ms_per_step"], mixed["stats"]),
              file=sys.stderr, flush=True)
        fp64 = time_path(torch, s, "double", steps, 2, ewald=ewald)
        print("# saline96000 %s fp64: %.2f ms/step, %s" % (model, fp64["ms_per_step"], fp64["stats"]),
              file=sys.stderr, flush=True)
        a, b = mixed["forces"].reshape(-1), fp64["forces"].reshape(-1)
        out[model] = {
            "mixed": 1000.0 / mixed["ms_per_step"],
            "fp64": 1000.0 / fp64["ms_per_step"],
            "mixed_vs_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
            "pcg_iterations": mixed["stats"]["iterations"],
            "mixed_force_rel_rms_err_vs_fp64": float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))),
            "mixed_energy_rel_err_vs_fp64": abs(mixed["energy"] - fp64["energy"]) / abs(fp64["energy"]),
            "mixed_mu_max_err_vs_fp64_debye": float(np.max(np.abs(mixed["mu"] - fp64["mu"]))) / DEBYE,
        }
    return out


def md_bench(torch, s, steps):
    """Evaluations along a trajectory-like sequence of configurations of the 36,864-site box: before
    each evaluation every coordinate moves by a seeded Gaussian step of 0.0005 nm (about the
    displacement of one ~1 fs MD step), and each Mutual solve starts from the previous evaluation's
    dipoles (tdp_set_warm_start), as an integrator loop would run it.  Ewald option, mixed precision,
    PCG."""
    h = make_handle(s, "mixed", ewald=True)
    h.set_warm_start(True)
    dev = torch.device("cuda:0")
    pos = torch.tensor(s.positions.reshape(-1), dtype=torch.float64, device=dev)
    frc = torch.zeros(pos.numel(), dtype=torch.float64, device=dev)
    en = torch.zeros(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261210)
    warm = 3
    noise = torch.randn((warm + steps, pos.numel()), generator=gen, device=dev, dtype=torch.float64) * 5e-4
    for k in range(warm):
        pos.add_(noise[k])
        h.execute_device(pos.data_ptr(), s.box, frc.data_ptr(), en.data_ptr(), stream)
    torch.cuda.synchronize()
    iters = []
    t0 = time.perf_counter()
    for k in range(steps):
        pos.add_(noise[warm + k])
        h.execute_device(pos.data_ptr(), s.box, frc.data_ptr(), en.data_ptr(), stream)
        iters.append(h.solver_stats()["iterations"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    h.close()
    out = {"workload": "36,864-site box (Ewald option, cutoff 1.0 nm, 64^3 grid), every coordinate moved by "
                       "N(0, 0.0005 nm) before each evaluation, Mutual solve warm-started from the previous "
                       "dipoles, mixed PCG to 1e-6 D",
           "unit": "evals/s", "evals_per_s": steps / dt, "ms_per_step": 1000.0 * dt / steps,
           "pcg_iterations_mean": float(np.mean(iters)), "pcg_iterations_max": int(np.max(iters)), "steps": steps}
    print("# md warm start: %.2f ms/step, PCG iterations %.1f (max %d)" % (out["ms_per_step"], out["pcg_iterations_mean"],
                                                                           out["pcg_iterations_max"]),
          file=sys.stderr, flush=True)
    return out


NO_AUTOTUNE = False


def errors(a, b):
    """Mixed result a against FP64 result b: force relative RMS, energy relative, dipole max (D)."""
    fa, fb = a["forces"].reshape(-1), b["forces"].reshape(-1)
    return (float(np.sqrt(np.mean((fa - fb) ** 2)) / np.sqrt(np.mean(fb ** 2))),
            abs(a["energy"] - b["energy"]) / abs(b["energy"]),
            float(np.max(np.abs(a["mu"] - b["mu"]))) / DEBYE)


def spmv_roofline(prof, n):
    """HBM roofline of the stored-block T*mu kernel from the timed pass: algorithmic bytes per launch
    (each stored block read once: 18 bytes per FP32 entry, 36 per FP64 entry; v once in FP32 and
    FP64; the list partials written once) over its average launch time."""
    st, info = prof["stats"], prof["info"]
    launches = max(st["matvec_full_launches"], 1)
    avg_ms = st["matvec_ms_full"] / launches
    lists = info["tiles"] * 8
    other = n * (16 + 32) + lists * 64 * 32
    algo = info["entries32"] * info["bytes_per_entry32"] + info["entries64"] * info["bytes_per_entry64"] + other
    streamed = info["slots32"] * info["bytes_per_entry32"] + info["slots64"] * info["bytes_per_entry64"] + other
    achieved = algo / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": SPMV_HBM_BYTES,
            "kernel": "k_ew_spmv<true> (stored real-space blocks of T applied to the dipoles, one PCG T*mu)",
            "algorithmic_bytes": algo, "streamed_bytes_with_padding": streamed,
            "entries_fp32": info["entries32"], "entries_fp64": info["entries64"],
            "bytes_per_entry": {"fp32": info["bytes_per_entry32"], "fp64": info["bytes_per_entry64"]},
            "avg_launch_ms": avg_ms, "full_launches": st["matvec_full_launches"],
            "launches": st["matvec_launches"]}


def allpairs_bench(torch, s, steps, warmup, no_fp64, no_cpu, cpu_seconds):
    """The same box with the reference's PME semantics (minimum image over all pairs, no cutoff):
    mixed and FP64 evaluations per second, the T*mu pass against the FP32 compute roofline, and the
    oracle's pair passes as CPU baseline."""
    n = s.n
    mixed = time_path(torch, s, "mixed", steps, warmup)
    print("# allpairs mixed: %.2f ms/step, %s" % (mixed["ms_per_step"], mixed["stats"]), file=sys.stderr, flush=True)
    n_prof = max(1, min(steps, 3))
    prof = time_path(torch, s, "mixed", n_prof, 0, timing=True)
    st = prof["stats"]
    avg_ms = st["matvec_ms_full"] / max(st["matvec_full_launches"], 1)
    flops = FLOPS_PER_PAIR_MATVEC * n * (n - 1) / 2.0
    achieved_tflops = flops / (avg_ms * 1e-3) / 1e12
    out = {"workload": "36,864 sites, all pairs under the minimum image (the reference's PME semantics), Mutual PCG "
                       "to 1e-6 D", "unit": "evals/s", "mixed": 1000.0 / mixed["ms_per_step"],
           "ms_per_step": mixed["ms_per_step"], "pcg_iterations": mixed["stats"]["iterations"],
           "tiles_per_work_item": mixed["stats"]["tiles_per_item"],
           "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": MATVEC_HBM_BYTES,
                        "kernel": "k_sym_matvec_mixed<true> (pair terms of one T*mu pass)",
                        "ceiling": "the published dense FP32 rate; the pair terms run on the vector ALUs",
                        "flops_per_pair": FLOPS_PER_PAIR_MATVEC, "pairs_per_launch": n * (n - 1) // 2,
                        "avg_launch_ms": avg_ms, "full_launches": st["matvec_full_launches"]}}
    if not no_fp64:
        fp64 = time_path(torch, s, "double", steps, warmup)
        print("# allpairs fp64: %.2f ms/step, %s" % (fp64["ms_per_step"], fp64["stats"]), file=sys.stderr,
              flush=True)
        e = errors(mixed, fp64)
        out.update({"fp64": 1000.0 / fp64["ms_per_step"], "mixed_vs_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
                    "mixed_force_rel_rms_err_vs_fp64": e[0], "mixed_energy_rel_err_vs_fp64": e[1],
                    "mixed_mu_max_err_vs_fp64_debye": e[2]})
    if not no_cpu:
        from tholedipoleplugin_amd import _lib
        h = _lib.HipContextHandle(s.force, s.box, precision="double", solver="pcg")
        mu = h.dipoles(s.positions, s.box, _lib.DIPOLES_INDUCED)
        iters = h.solver_stats()["iterations"] + 1
        h.close()
        out["cpu_baseline"] = cpu_baseline(s, mu, iters, cpu_seconds * 3.0)
        print("# allpairs cpu baseline: %s" % out["cpu_baseline"], file=sys.stderr, flush=True)
    return out


def main():
    global NO_AUTOTUNE
    args = parse()
    NO_AUTOTUNE = args.no_autotune
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.clusters_only:
        print(json.dumps({"clusters": cluster_bench(torch, args.steps, args.warmup)}), flush=True)
        return
    if args.dense_only:
        print(json.dumps({"dense_clusters": dense_bench(args.steps)}), flush=True)
        return
    s = build_system()
    n = s.n
    h = make_handle(s, "double", ewald=True)
    alpha, nx, ny, nz = h.pme_parameters()
    h.close()
    print("# workload: %d sites, box %.2f nm, cutoff %.2f nm, alpha %.4f /nm, grid %dx%dx%d; warmup %d, steps %d"
          % (n, s.box[0, 0], s.force.getCutoffDistance(), alpha, nx, ny, nz, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    mixed = time_path(torch, s, "mixed", args.steps, args.warmup, ewald=True)
    print("# ewald mixed: %.2f ms/step, %s" % (mixed["ms_per_step"], mixed["stats"]), file=sys.stderr, flush=True)
    # separate timed pass for the kernel roofline (events around every T*mu launch)
    prof = time_path(torch, s, "mixed", max(1, min(args.steps, 3)), 0, timing=True, ewald=True)
    roof = spmv_roofline(prof, n)
    print("# T*mu (k_ew_spmv): %d full launches, avg %.1f us, %.0f GB/s algorithmic (%.2f of peak)"
          % (roof["full_launches"], roof["avg_launch_ms"] * 1e3, roof["achieved"], roof["frac"]), file=sys.stderr,
          flush=True)
    fp64 = None
    if not args.no_fp64:
        fp64 = time_path(torch, s, "double", args.steps, args.warmup, ewald=True)
        print("# ewald fp64: %.2f ms/step, %s" % (fp64["ms_per_step"], fp64["stats"]), file=sys.stderr, flush=True)
    err = (None, None, None) if fp64 is None else errors(mixed, fp64)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_ewald(s, alpha, (nx, ny, nz))
        ref = {"forces": cpu.pop("forces"), "energy": cpu.pop("energy"), "mu": cpu.pop("induced")}
        e = errors(mixed, ref)
        cpu["mixed_force_rel_rms_err_vs_oracle"], cpu["mixed_energy_rel_err_vs_oracle"] = e[0], e[1]
        cpu["mixed_mu_max_err_vs_oracle_debye"] = e[2]
        if fp64 is not None:
            e = errors(fp64, ref)
            cpu["fp64_force_rel_rms_err_vs_oracle"], cpu["fp64_energy_rel_err_vs_oracle"] = e[0], e[1]
            cpu["fp64_mu_max_err_vs_oracle_debye"] = e[2]
        print("# cpu baseline: %s" % cpu, file=sys.stderr, flush=True)
    out = {
        "metric": "force+induction evals/s (36k sites) & cluster solves/s, mixed vs FP64 @<=1e-5 err",
        "value": 1000.0 / mixed["ms_per_step"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "water12288_36864_sites_box4.95nm_pme_cutoff1.0nm_grid64_order5_mutual_1e-6D",
                   "sites": n, "model": "Ewald option (real space within the cutoff + smooth PME)",
                   "alpha_per_nm": alpha, "grid": [nx, ny, nz], "solver": "pcg", "precision": "mixed",
                   "pcg_iterations": mixed["stats"]["iterations"]},
        "fp64_path": None if fp64 is None else {"evals_per_s": 1000.0 / fp64["ms_per_step"],
                                                "ms_per_step": fp64["ms_per_step"],
                                                "iterations": fp64["stats"]["iterations"]},
        "mixed_vs_fp64_speedup": None if fp64 is None else fp64["ms_per_step"] / mixed["ms_per_step"],
        "mixed_force_rel_rms_err_vs_fp64": err[0],
        "mixed_energy_rel_err_vs_fp64": err[1],
        "mixed_mu_max_err_vs_fp64_debye": err[2],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not args.pmode_only:
        out["allpairs_min_image"] = None if args.no_allpairs else allpairs_bench(
            torch, s, args.steps, args.warmup, args.no_fp64, args.no_cpu_baseline, args.cpu_seconds)
        out["clusters"] = None if args.no_clusters else cluster_bench(torch, max(1, min(args.steps, 10)), 3)
        out["dense_clusters"] = None if args.no_dense else dense_bench(max(1, min(args.steps, 2)))
        out["md_warm_start"] = None if args.no_md else md_bench(torch, s, args.steps)
        out["saline96000"] = None if args.no_saline else saline_bench(torch, max(2, min(args.steps, 6)))
    print(json.dumps(out), flush=True)
    if args.copy_probe_mib > 0:
        x = torch.ones(args.copy_probe_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print("# copy probe: %d MiB read + %d MiB written per launch, 3 launches" % (args.copy_probe_mib,
                                                                                    args.copy_probe_mib),
              file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
