"""Benchmark: 3D OSEM on one MI355X.

Workload (the reference's 3D OSEM single-job configuration): 256x256x81 float32
image, 2.0445 mm in-plane / 2.027 mm axial voxels; cylindrical 3D sinogram of
344 radial bins x 252 angles x 837 sinograms (64 rings, span 11, maximum ring
difference 60, scanner radius 400 mm); phantom = uniform cylinder r=100 mm filling
the axial FOV + 6 point sources (value 50) at seeded positions; measured data =
Poisson(3e8 counts) of the phantom's projection (synthetic, seeded); 21 subsets.

A step is one full OSEM iteration (21 sub-iterations: fused projection + ratio +
likelihood, backprojection, update) with every input resident in HBM.  The
sensitivity images are computed once before timing, as the reference does before
its loop (OsemSinogram3d.cpp:101-117).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--bp tile|gather|scatter]
                    [--no-cpu-baseline] [--calibrate-copy]

Prints one JSON line on stdout; progress goes to stderr."""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C2 = dict(nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344, rings=64,
          rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60)
SUBSETS = 21
COUNTS = 3e8
HBM_PEAK_GBS = 8000.0


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bp", choices=["tile", "gather", "scatter"], default="tile")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 1 GiB device copy first (for PMC byte calibration)")
    ap.add_argument("--workload", choices=["osem", "batch", "serve", "shard", "c1"],
                    default="osem",
                    help="osem: the single-job 256x256x81 OSEM line (default); batch: "
                         "64 seeded jobs of uneven size, batched vs one at a time; serve: "
                         "open-loop arrivals over the batch, p50/p99 latency per rate; "
                         "shard: one rank's share of the OSEM job sharded by sinograms; "
                         "c1: 2D MLEM on 128x128 beside the reference's CPU Mlem2d")
    ap.add_argument("--ranks", type=int, default=8,
                    help="ranks of the sharded job (--workload shard)")
    ap.add_argument("--shard-layout", choices=["balanced", "block", "interleaved"],
                    default="balanced",
                    help="sinograms of a rank: a contiguous run of about equal cost, a "
                         "contiguous run of equal length, or every n-th one")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--rates", default="5,10,20,50,100,200")
    return ap.parse_args()


REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def cpu_baseline(meas_host, threads, launcher):
    """The reference's own CPU path (oracle/_ref/ref_driver, built from its sources)
    timed on one OSEM sub-iteration of the same workload; iterations/s = 1 / (21 x t).
    The driver is started by `launcher` (apirl_amd.spawn), whose helper process was
    created before this process initialised the GPU."""
    drv = REF_DRIVER
    if not os.path.exists(drv) or launcher is None:
        return None
    with tempfile.TemporaryDirectory() as td:
        mp = os.path.join(td, "meas.f32")
        meas_host.tofile(mp)
        args = [drv, "osem3d_sub"] + [f"{k}={v}" for k, v in C2.items()] + \
            [f"subsets={SUBSETS}", "subset=0", f"meas={mp}"]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        try:
            rc, out, err = launcher.run(args, env=env, timeout=600)
        except Exception as e:  # the reference may fail; report it rather than crash
            progress(f"cpu baseline failed: {e}")
            return None
        if rc != 0:
            progress(f"cpu baseline failed (exit {rc}): {err[-500:]}")
            return None
    rec = json.loads(out.splitlines()[0])
    t = rec["seconds"]
    return {"value": 1.0 / (SUBSETS * t), "unit": "OSEM iters/s", "cores": threads,
            "kind": "reference",
            "sample": f"one OSEM sub-iteration (subset 0 of {SUBSETS}: Project + ratio + "
                      f"Backproject + update, {t:.2f} s) of the same workload, "
                      f"OMP_NUM_THREADS={threads}; iterations/s = 1/(21 t)"}


def profile_figure(name, key, entry):
    """A per-launch figure of one entry point committed under profiles/ by its rocprofv3
    passes: lines "<key> <entry> <value>"; None when no measurement is committed."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == key and f[1] == entry:
            return float(f[2])
    return None


def measured_traffic(entry):
    """HBM bytes per launch of one entry point (all of its kernels), as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/pmc_traffic_c2.md."""
    return profile_figure("pmc_traffic_c2.md", "traffic", entry)


# MI355X: 256 CUs; one VALU wave-instruction issued per cycle per CU at 2.4 GHz
VALU_ISSUE_PER_S = 256 * 2.4e9


def valu_issue(entry, ms):
    """The VALU issue bound of the entry point's main kernel: its VALU wave-instructions
    per launch (SQ_INSTS_VALU, profiles/sq_counters_c2.md) at one per cycle per CU,
    against the measured launch time."""
    v = profile_figure("sq_counters_c2.md", "valu", entry)
    if v is None:
        return None
    t = v / VALU_ISSUE_PER_S * 1e3
    return {"wave_instructions": v, "issue_ms": t, "frac": t / ms}


def batch_main(args):
    """Batched jobs: the 64-job mix of four geometries (and its small-job part) run
    one job at a time through the single-job driver, then through one batch."""
    import torch

    from apirl_amd import hip, serving, workloads

    ctx = hip.Context(0)
    out = {}
    mixes = (("mix", tuple(workloads.GEOMETRIES)), ("small", workloads.SMALL),
             ("2d_128", ("2d_128",)))
    for label, kinds in mixes:
        specs = [workloads.draw(k, kinds=kinds, attenuation=True) for k in range(args.jobs)]
        progress(f"{label}: materialising {len(specs)} jobs")
        plans, jobs = serving.materialise_all(ctx, specs, torch)
        # warm-up: one small job each way (kernels loaded, pools primed)
        serving.run_sequential(ctx, plans, jobs[:2], torch)
        t_seq = serving.run_sequential(ctx, plans, jobs, torch)
        progress(f"{label}: one at a time {t_seq:.2f} s")
        ref = [d["out"].clone() for _, d in jobs]
        t_bat, choice = serving.run_batched(ctx, plans, jobs, torch, max_jobs=args.jobs)
        progress(f"{label}: batched {t_bat:.2f} s (autotuned {choice})")
        same = all(torch.equal(torch.nan_to_num(a, nan=-1.0), torch.nan_to_num(d["out"], nan=-1.0))
                   for a, (_, d) in zip(ref, jobs))
        kinds_count = {k: sum(1 for s in specs if s["kind"] == k) for k in kinds}
        out[label] = {"jobs": len(specs), "kinds": kinds_count,
                      "sequential_jobs_per_s": len(specs) / t_seq,
                      "batched_jobs_per_s": len(specs) / t_bat,
                      "speedup": t_seq / t_bat, "autotuned": choice,
                      "results_bit_identical": bool(same)}
        del plans, jobs, ref
        torch.cuda.empty_cache()
    line = {"metric": "batched OSEM/MLEM jobs/s (64 seeded jobs, 4 geometries, 1-21 subsets, "
                      "1-10 iterations)",
            "value": out["mix"]["batched_jobs_per_s"], "unit": "jobs/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded ellipsoids, "
            "Poisson 1e5-1e8 counts, attenuation sinograms on half the jobs)",
            "config": {"workload": "batched jobs",
                       "backprojector": "tile for 3D jobs of 2^21 voxels or more, else gather"},
            "vs_single_job_path": out["mix"]["speedup"], "detail": out}
    print(json.dumps(line), flush=True)


def serve_main(args):
    """Serving: open-loop Poisson arrivals of the job mix, per rate."""
    import torch

    from apirl_amd import hip, serving, workloads

    ctx = hip.Context(0)
    rates = [float(r) for r in args.rates.split(",")]
    res = {}
    mixes = (("mix", tuple(workloads.GEOMETRIES)), ("small", workloads.SMALL),
             ("2d_128", ("2d_128",)))
    for label, kinds in mixes:
        specs = [workloads.draw(k, kinds=kinds, attenuation=True) for k in range(args.jobs)]
        progress(f"{label}: materialising {len(specs)} jobs")
        plans, jobs = serving.materialise_all(ctx, specs, torch)
        serving.run_sequential(ctx, plans, jobs[:2], torch)
        res[label] = {}
        for rate in rates:
            lat, wall = serving.serve(ctx, plans, jobs, rate, torch, max_jobs=args.jobs)
            res[label][str(rate)] = {"p50_s": float(np.percentile(lat, 50)),
                                     "p99_s": float(np.percentile(lat, 99)),
                                     "jobs_per_s": len(jobs) / wall}
            progress(f"{label} rate {rate}: p50 {res[label][str(rate)]['p50_s']:.3f} s, "
                     f"p99 {res[label][str(rate)]['p99_s']:.3f} s, "
                     f"{res[label][str(rate)]['jobs_per_s']:.2f} jobs/s")
        del plans, jobs
        torch.cuda.empty_cache()
    line = {"metric": "served jobs: p50/p99 latency and jobs/s per arrival rate",
            "value": res["mix"][str(rates[0])]["jobs_per_s"], "unit": "jobs/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded job mix)",
            "config": {"workload": "serving, open-loop Poisson arrivals",
                       "jobs_per_rate": args.jobs}, "detail": res}
    print(json.dumps(line), flush=True)


def shard_main(args):
    """One rank's share of the 256x256x81 OSEM job sharded by sinograms over
    args.ranks GPUs (args.shard_layout): the kernels of each rank's
    sub-iteration (fused projection + ratio, gather backprojection, update) timed on
    this GPU at that rank's real shape, beside the whole job's."""
    import torch

    from apirl_amd import hip, phantoms, sharding

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    ctx = hip.Context(0)
    g = C2
    geo = (g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"], g["rfov"],
           g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])
    full = hip.Plan.for_3d(ctx, *geo)
    img_np, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    truth = full.project(torch.from_numpy(img_np).to(dev))
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    meas = torch.poisson(truth * (COUNTS / float(truth.double().sum())), generator=gen)
    del truth
    x = torch.ones(full.image_shape(), dtype=torch.float32, device=dev)

    def sub_iteration_ms(plan, m):
        """Mean kernel time of one sub-iteration over the 21 subsets (HIP events)."""
        sens = torch.empty((SUBSETS,) + plan.image_shape(), dtype=torch.float32, device=dev)
        for s in range(SUBSETS):
            plan.sensitivity(out=sens[s], subset=s, nsubsets=SUBSETS, mode=hip.BP_GATHER)
        bp = torch.empty(plan.image_shape(), dtype=torch.float32, device=dev)
        xs = x.clone()
        times = []
        for rep in range(2):  # the first pass warms up
            for s in range(SUBSETS):
                r = torch.empty(plan.subset_shape(s, SUBSETS), dtype=torch.float32, device=dev)
                bp.zero_()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                plan.project_ratio(xs, m, subset=s, nsubsets=SUBSETS, ratio=r, want_lik=False)
                plan.backproject(r, out=bp, subset=s, nsubsets=SUBSETS, mode=hip.BP_GATHER)
                hip.em_update(ctx, xs, bp, sens[s], 0.0, 0)
                e1.record()
                if rep:
                    times.append((e0, e1))
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in times]))

    progress("whole job")
    full_ms = sub_iteration_ms(full, meas)
    per_rank = []
    for q in range(args.ranks):
        plan, idx = sharding.shard_plan(ctx, *geo, args.ranks, q, args.shard_layout)
        m = meas[torch.from_numpy(idx).to(dev)].contiguous()
        per_rank.append(sub_iteration_ms(plan, m))
        progress(f"rank {q}: {len(idx)} sinograms, {per_rank[-1]:.3f} ms per sub-iteration")
        del plan, m
    worst = max(per_rank)
    image_bytes = 4 * g["nx"] * g["ny"] * g["nz"]
    ring_bytes = 2 * (args.ranks - 1) / args.ranks * image_bytes
    line = {
        "metric": f"OSEM sub-iteration kernel ms, slowest rank's share of {args.ranks} "
                  "(256x256x81, 344x252x837, 21 subsets)",
        "value": worst,
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f32",
        "data": "synthetic (cylinder + 6 point sources, Poisson 3e8 counts, seed 0)",
        "config": {"workload": "one rank's share of the 3D OSEM job sharded by sinograms",
                   "layout": args.shard_layout, "backprojector": "gather",
                   "ranks": args.ranks},
        "per_rank_ms": per_rank,
        "whole_job_ms": full_ms,
        "share_of_whole": worst / full_ms,
        "ideal_share": 1.0 / args.ranks,
        "ring_allreduce_bytes_per_gpu": ring_bytes,
    }
    print(json.dumps(line), flush=True)


C1 = dict(nx=128, dx=2.0, nproj=192, nr=192, rfov=128.0, rscanner=400.0)
C1_ITERS = 60


def c1_main(args):
    """2D MLEM, 60 iterations on 128x128 @ 2 mm from a 192 x 192 sinogram (disc and hot
    spot, 1e6 Poisson counts, numpy seed 0), device-resident with the gather
    backprojector, beside the reference's Mlem2d on the host's cores."""
    launcher = None
    if not args.no_cpu_baseline and os.path.exists(REF_DRIVER):
        from apirl_amd.spawn import Launcher
        launcher = Launcher().start()  # before the GPU is initialised
    import torch

    from apirl_amd import hip, phantoms

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    ctx = hip.Context(0)
    g = C1
    plan = hip.Plan.for_2d(ctx, g["nx"], g["dx"], g["nproj"], g["nr"], g["rfov"], g["rscanner"])
    img = torch.from_numpy(phantoms.disc2d(g["nx"], g["dx"]).reshape(plan.image_shape())).cuda()
    truth = plan.project(img).cpu().numpy()
    meas_np = phantoms.poisson_counts(truth, 1e6, np.random.default_rng(0)).astype(np.float32)
    meas = torch.from_numpy(meas_np.reshape(truth.shape)).cuda()
    init = torch.ones(plan.image_shape(), dtype=torch.float32, device="cuda")
    for _ in range(max(args.warmup, 1)):
        plan.reconstruct(meas, init, C1_ITERS, subsets=1, bp_mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.reconstruct(meas, init, C1_ITERS, subsets=1, bp_mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it_s = args.steps * C1_ITERS / wall
    progress(f"C1: {it_s:.1f} MLEM it/s ({wall / args.steps * 1e3:.2f} ms per 60 iterations)")
    cpu = None
    if launcher is not None:
        threads = min(os.cpu_count() or 1, 16)
        with tempfile.TemporaryDirectory() as td:
            mp, ip = os.path.join(td, "meas.f32"), os.path.join(td, "init.f32")
            meas_np.tofile(mp)
            np.ones(g["nx"] * g["nx"], np.float32).tofile(ip)
            a = [REF_DRIVER, "mlem2d", f"nx={g['nx']}", f"dx={g['dx']}", f"nproj={g['nproj']}",
                 f"nr={g['nr']}", f"rscanner={g['rscanner']}", f"iters={C1_ITERS}",
                 f"meas={mp}", f"init={ip}", f"outdir={td}"]
            rc, out, err = launcher.run(a, env=dict(os.environ, OMP_NUM_THREADS=str(threads)),
                                        timeout=600)
        launcher.close()
        if rc == 0:
            rec = [ln for ln in out.splitlines() if ln.startswith('{"what"')]
            t = json.loads(rec[-1])["seconds"]
            cpu = {"value": C1_ITERS / t, "unit": "MLEM iters/s", "cores": threads,
                   "kind": "reference",
                   "sample": f"the reference's Mlem2d, {C1_ITERS} iterations with its "
                             f"sensitivity ({t:.2f} s), OMP_NUM_THREADS={threads}"}
        else:
            progress(f"cpu baseline failed (exit {rc}): {err[-500:]}")
    line = {
        "metric": "2D MLEM iters/s (128x128, 192x192)",
        "value": it_s,
        "unit": "MLEM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (disc + hot spot, Poisson 1e6 counts, numpy seed 0)",
        "config": {"workload": f"2D MLEM, {C1_ITERS} iterations per step, device-resident "
                               "(sensitivity included)", "backprojector": "gather"},
        "ms_per_step": wall / args.steps * 1e3,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one rank's share of a sharded job: --workload shard)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "serve":
        return serve_main(args)
    if args.workload == "shard":
        return shard_main(args)
    if args.workload == "c1":
        return c1_main(args)
    launcher = None
    if not args.no_cpu_baseline and os.path.exists(REF_DRIVER):
        from apirl_amd.spawn import Launcher
        launcher = Launcher().start()  # before the GPU is initialised
    import torch

    from apirl_amd import hip, phantoms

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    ctx = hip.Context(0)
    g = C2
    progress("building geometry plan")
    plan = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"],
                           g["nr"], g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"],
                           g["mrd"])
    img_np, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    truth = plan.project(torch.from_numpy(img_np).to(dev))
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    lam = truth * (COUNTS / float(truth.double().sum()))
    meas = torch.poisson(lam, generator=gen).contiguous()
    del truth, lam
    nvox = plan.nx * plan.ny * plan.nz
    progress("sensitivity images")
    sens = torch.empty((SUBSETS,) + plan.image_shape(), dtype=torch.float32, device=dev)
    bp_mode = {"tile": hip.BP_TILE, "gather": hip.BP_GATHER, "scatter": hip.BP_SCATTER}[args.bp]
    for s in range(SUBSETS):
        plan.sensitivity(out=sens[s], subset=s, nsubsets=SUBSETS, mode=bp_mode)
    n_int = [plan.intersections(s, SUBSETS) for s in range(SUBSETS)]
    n_ray = [int(np.prod(plan.subset_shape(s, SUBSETS))) for s in range(SUBSETS)]
    ratio = torch.empty(max(n_ray), dtype=torch.float32, device=dev)
    lik = torch.zeros(SUBSETS, dtype=torch.float64, device=dev)  # per sub-iteration
    bp = torch.empty(plan.image_shape(), dtype=torch.float32, device=dev)
    x = torch.ones(plan.image_shape(), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    assert torch.cuda.current_stream().cuda_stream == 0, "events must be on the library stream"

    ev = {"fwd": [], "bwd": [], "upd": []}

    fused = args.bp != "scatter"

    def iteration(record):
        for s in range(SUBSETS):
            r = ratio[: n_ray[s]].view(plan.subset_shape(s, SUBSETS))
            if not fused:
                bp.zero_()
            if record:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e2 = torch.cuda.Event(enable_timing=True)
                e3 = torch.cuda.Event(enable_timing=True)
                e0.record()
            # the reference evaluates the likelihood every sub-iteration
            # (OsemSinogram3d.cpp:190): the fused epilogue reduces it with the ratio
            plan.project_ratio(x, meas, subset=s, nsubsets=SUBSETS, ratio=r, lik=lik[s:s + 1])
            if record:
                e1.record()
            if fused:
                # the backprojection with the update fused in (OSEM's threshold 0, >=)
                plan.backproject_update(r, x, sens[s], 0.0, 0, subset=s, nsubsets=SUBSETS,
                                        mode=bp_mode)
            else:
                plan.backproject(r, out=bp, subset=s, nsubsets=SUBSETS, mode=bp_mode)
            if record:
                e2.record()
            if not fused:
                hip.em_update(ctx, x, bp, sens[s], 0.0, 0)
            if record:
                e3.record()
                ev["fwd"].append((e0, e1, s))
                ev["bwd"].append((e1, e2, s))
                ev["upd"].append((e2, e3, s))

    progress(f"warmup {args.warmup} iteration(s)")
    for _ in range(args.warmup):
        iteration(False)
    torch.cuda.synchronize()
    progress(f"timing {args.steps} iteration(s)")
    t0 = time.perf_counter()
    for k in range(args.steps):
        iteration(True)
        torch.cuda.synchronize()
        progress(f"iteration {k + 1}/{args.steps} done")
    t1 = time.perf_counter()
    torch.cuda.synchronize()
    wall = t1 - t0
    it_s = args.steps / wall

    def avg_ms(key):
        return float(np.mean([a.elapsed_time(b) for a, b, _ in ev[key]]))

    fwd_ms, bwd_ms, upd_ms = avg_ms("fwd"), avg_ms("bwd"), avg_ms("upd")
    lik_last = lik.cpu().numpy()
    if not np.all(np.isfinite(lik_last)):
        raise RuntimeError("likelihood is not finite")
    nint_it = float(sum(n_int))
    nray_it = float(sum(n_ray))
    # algorithmic bytes per launch (averaged over the 21 subsets):
    #   forward fused: 4 N_int (voxel reads) + 4 N_ray (meas read) + 4 N_ray (ratio write)
    #   backprojection: 4 N_int (voxel read-modify-writes counted once) + 4 N_ray + 4 N_vox;
    #   with the update fused in (tile, gather), 12 N_vox: image and sensitivity read,
    #   image write
    fwd_bytes = (4 * nint_it + 8 * nray_it) / SUBSETS
    bwd_bytes = (4 * nint_it + 4 * nray_it) / SUBSETS + (12 if fused else 4) * nvox
    dom = "bwd" if bwd_ms >= fwd_ms else "fwd"
    dom_bytes, dom_ms = (bwd_bytes, bwd_ms) if dom == "bwd" else (fwd_bytes, fwd_ms)
    dom_entry = "forward_fused" if dom == "fwd" else f"backproject_{args.bp}"
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    gint = nint_it * 2 * args.steps / wall / 1e9  # forward + backward traversals
    progress(f"{it_s:.3f} it/s; fwd {fwd_ms:.3f} ms, bwd {bwd_ms:.3f} ms, upd {upd_ms:.3f} ms")

    cpu = None
    if not args.no_cpu_baseline:
        progress("cpu baseline (reference CPU path, one sub-iteration)")
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(meas.cpu().numpy(), threads, launcher)
    if launcher is not None:
        launcher.close()

    line = {
        "metric": "OSEM iters/s (256x256x81, 344x252x837, 21 subsets)",
        "value": it_s,
        "unit": "OSEM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (cylinder + 6 point sources, Poisson 3e8 counts, seed 0)",
        "config": {"workload": "3D OSEM single job: 256x256x81 image, 344r x 252ang x 837 "
                               "sinograms (64 rings, span 11, MRD 60), 21 subsets",
                   "backprojector": args.bp, "parallelism": "single GPU"},
        "likelihood_last_iteration": [float(v) for v in lik_last[[0, -1]]],
        "ray_voxel_gint_per_s": gint,
        "intersections_per_iteration": nint_it,
        "kernel_ms": ({"forward_fused": fwd_ms, "backproject_update": bwd_ms} if fused else
                      {"forward_fused": fwd_ms, "backproject": bwd_ms, "update": upd_ms}),
        "roofline": {"bound": "hbm", "kernel": dom_entry,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_entry),
                     "valu_issue": valu_issue(dom_entry, dom_ms)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
