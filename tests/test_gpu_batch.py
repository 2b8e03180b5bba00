"""Batched jobs (apirl_batch_*): every job's image equals, bit for bit, the image the
single-job driver (apirl_em_reconstruct, with the backprojector the batch runs the
job with: the tile form for 3D jobs of 2^21 voxels or more, else the gather) gives it
alone --
whether the jobs start together, join while others run (continuous batching), or
run with the round graph replayed or launched directly.  NaN voxels (the top
slice, never visited by any LOR) must coincide."""
import numpy as np
import pytest

from apirl_amd import hip, phantoms
from test_gpu_parity import ctx, dev, host, log, torch  # noqa: F401

pytestmark = pytest.mark.gpu

GEOMS = {
    "2d_small": dict(dims=2, nx=64, dx=4.0, nproj=48, nr=64, rfov=128.0, rscanner=400.0),
    "2d_c1": dict(dims=2, nx=128, dx=2.0, nproj=192, nr=192, rfov=128.0, rscanner=400.0),
    "3d_mid": dict(dims=3, nx=64, ny=64, nz=17, dx=2.0445, dz=2.027, nproj=42, nr=86, rings=13,
                   rfov=65.424, zfov=34.459, rscanner=400.0, span=5, mrd=8),
    "3d_tiny": dict(dims=3, nx=32, ny=32, nz=9, dx=4.0, dz=4.0, nproj=24, nr=40, rings=5,
                    rfov=64.0, zfov=36.0, rscanner=300.0, span=3, mrd=2),
}


def make_plan(ctx, g):
    if g["dims"] == 2:
        return hip.Plan.for_2d(ctx, g["nx"], g["dx"], g["nproj"], g["nr"], g["rfov"],
                               g["rscanner"])
    return hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                           g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])


def make_job(torch, plan, g, seed, iterations, subsets):
    rng = np.random.default_rng(seed)
    nz = 1 if g["dims"] == 2 else g["nz"]
    img = phantoms.ellipsoids3d(g["nx"], nz, g["dx"], g.get("dz", 1.0), rng)
    truth = plan.project(dev(torch, img.reshape(plan.image_shape())))
    t = host(truth)
    meas = phantoms.poisson_counts(t, float(10 ** rng.uniform(5, 7)), rng).reshape(t.shape)
    init = torch.ones(plan.image_shape(), dtype=torch.float32, device="cuda")
    return dict(meas=dev(torch, meas), init=init, iterations=iterations, subsets=subsets)


def alone(plan, job):
    img, _ = plan.reconstruct(job["meas"], job["init"], job["iterations"],
                              subsets=job["subsets"], bp_mode=hip.Batch.backprojector(plan))
    return host(img)


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(a)
    d = np.flatnonzero(a[m].view(np.int32) != b[m].view(np.int32))
    assert d.size == 0, f"{d.size} voxels differ"


@pytest.fixture(scope="module")
def mix(torch, ctx):
    plans = {k: make_plan(ctx, g) for k, g in GEOMS.items()}
    spec = [("2d_small", 1, 4, 1), ("2d_c1", 2, 3, 1), ("2d_small", 3, 2, 6), ("3d_mid", 4, 2, 7),
            ("3d_tiny", 5, 3, 4), ("3d_mid", 6, 1, 3), ("2d_c1", 7, 2, 8), ("3d_tiny", 8, 2, 1)]
    jobs = []
    for name, seed, it, sub in spec:
        jobs.append((name, make_job(torch, plans[name], GEOMS[name], seed, it, sub)))
    ref = [alone(plans[n], j) for n, j in jobs]
    torch.cuda.synchronize()
    return plans, jobs, ref


def run_batch(torch, ctx, plans, jobs, late=(), use_graph=True, grid=0, schedule=None):
    b = hip.Batch(ctx, max_jobs=16, max_sinograms=1024)
    b.set_launch(grid, use_graph)
    if schedule is not None:
        b.set_schedule(*schedule)
    outs, ids = {}, {}
    for k, (name, j) in enumerate(jobs):
        if k in late:
            continue
        outs[k] = torch.empty(plans[name].image_shape(), dtype=torch.float32, device="cuda")
        ids[k] = b.submit(plans[name], j["meas"], j["init"], outs[k], j["iterations"], j["subsets"])
    rounds = 0
    while True:
        n = b.step()
        rounds += 1
        if rounds == 3:
            for k in late:  # admitted while the others are running
                name, j = jobs[k]
                outs[k] = torch.empty(plans[name].image_shape(), dtype=torch.float32,
                                      device="cuda")
                ids[k] = b.submit(plans[name], j["meas"], j["init"], outs[k], j["iterations"],
                                  j["subsets"])
            continue
        if n == 0:
            break
    b.synchronize()
    assert all(b.done(ids[k]) for k in ids)
    res = {k: host(outs[k]) for k in outs}
    for k in ids:
        b.release(ids[k])
    b.close()
    return res


def test_batch_equals_each_job_alone(torch, ctx, mix):
    log("case: batched mix of 8 jobs (2D and 3D, 1-8 subsets) vs each alone")
    plans, jobs, ref = mix
    res = run_batch(torch, ctx, plans, jobs)
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


def test_batch_admission_while_running(torch, ctx, mix):
    log("case: jobs admitted after 3 rounds of the others")
    plans, jobs, ref = mix
    res = run_batch(torch, ctx, plans, jobs, late=(2, 5, 6))
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


def test_batch_graph_and_direct_launches_agree(torch, ctx, mix):
    log("case: direct launches, small persistent grid")
    plans, jobs, ref = mix
    res = run_batch(torch, ctx, plans, jobs, use_graph=False, grid=96)
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


@pytest.mark.parametrize("schedule", [(1, hip.Batch.ORDER_LARGEST_FIRST),
                                      (8, hip.Batch.ORDER_SUBMISSION)])
def test_batch_schedules_agree(torch, ctx, mix, schedule):
    """The autotuner's scheduling knobs (virtual blocks per draw, the jobs' order in
    a round) move work between workgroups only: every image equals the job alone."""
    log(f"case: batch schedule grab={schedule[0]} order={schedule[1]}")
    plans, jobs, ref = mix
    res = run_batch(torch, ctx, plans, jobs, schedule=schedule)
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


@pytest.mark.parametrize("budget", ["ample", "one_job"])
def test_serving_loop_results_equal_alone(torch, ctx, budget):
    """Open-loop arrivals through serving.serve (admission by HBM footprint, release
    and reuse of device state): each image equals the job run alone.  With the budget
    set to one job's footprint the jobs are admitted one after another."""
    from apirl_amd import serving, workloads
    log(f"case: serving loop, {budget} HBM budget")
    specs = [workloads.draw(k, kinds=("2d_128", "3d_small"), attenuation=True, max_iterations=2)
             for k in range(6)]
    plans, jobs = serving.materialise_all(ctx, specs, torch)
    ref = []
    for spec, d in jobs:
        p = plans[spec["kind"]]
        img, _ = p.reconstruct(d["meas"], d["init"], spec["iterations"], subsets=spec["subsets"],
                               bp_mode=hip.Batch.backprojector(p), mult=d["mult"])
        ref.append(host(img))
    need = max(hip.Batch.job_bytes(plans[s["kind"]], s["subsets"]) for s in specs)
    lat, wall = serving.serve(ctx, plans, jobs, rate=200.0, torch=torch, max_jobs=8,
                              hbm_budget=1e12 if budget == "ample" else need)
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    for k, (_, d) in enumerate(jobs):
        same_bits(host(d["out"]), ref[k])


def test_batch_with_large_job_equals_alone(torch, ctx):
    """A job of at least 2^21 voxels (256x256x81) runs its sub-iterations with the
    single-job kernels after each grouped round; it and the small jobs beside it
    still equal, bit for bit, each job alone."""
    from apirl_amd import workloads
    log("case: batch with a 256x256x81 job beside small 2D / 3D jobs")
    big = workloads.make_plan(ctx, "3d_c2")
    g = workloads.GEOMETRIES["3d_c2"]
    plans = {"big": big, "3d_tiny": make_plan(ctx, GEOMS["3d_tiny"]),
             "2d_small": make_plan(ctx, GEOMS["2d_small"])}
    jobs = [("big", make_job(torch, big, g, 21, 1, 2)),
            ("3d_tiny", make_job(torch, plans["3d_tiny"], GEOMS["3d_tiny"], 22, 2, 3)),
            ("2d_small", make_job(torch, plans["2d_small"], GEOMS["2d_small"], 23, 3, 2))]
    ref = [alone(plans[n], j) for n, j in jobs]
    torch.cuda.synchronize()
    for use_graph in (True, False):
        res = run_batch(torch, ctx, plans, jobs, use_graph=use_graph)
        for k in range(len(jobs)):
            same_bits(res[k], ref[k])


def test_batch_with_mmr_job_equals_alone(torch, ctx):
    """A job on the mMR geometry (non-uniform radial grid, odd bins at half-step
    angles: the grouped gather reads the job's radial lookup table) beside a
    cylindrical 3D job: each equals, bit for bit, the job alone."""
    log("case: batch with an mMR job beside a cylindrical 3D job")
    g = dict(dims=3, nx=32, ny=32, nz=65, dx=16.0, dz=4.0, nproj=24, nr=86, rings=64,
             rfov=297.0, zfov=260.0, rscanner=328.0, span=3, mrd=2)
    mmr = make_plan(ctx, g)
    mmr.set_scanner("mmr")
    plans = {"mmr": mmr, "3d_tiny": make_plan(ctx, GEOMS["3d_tiny"])}
    jobs = [("mmr", make_job(torch, mmr, g, 31, 2, 3)),
            ("3d_tiny", make_job(torch, plans["3d_tiny"], GEOMS["3d_tiny"], 32, 2, 2))]
    ref = [alone(plans[n], j) for n, j in jobs]
    torch.cuda.synchronize()
    assert np.nanmax(ref[0]) > 0
    res = run_batch(torch, ctx, plans, jobs)
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


def test_batch_with_oversampled_job_equals_alone(torch, ctx):
    """A job with 2 x 2 LORs per bin (apirl_plan_set_oversampling; it runs with the
    single-job kernels, its gather over the plans of one LOR per bin) beside a plain 3D
    job: each equals, bit for bit, the job alone with the gather backprojector."""
    log("case: batch with an oversampled job beside a plain 3D job")
    g = GEOMS["3d_tiny"]
    os_plan = make_plan(ctx, g)
    os_plan.set_oversampling(2, 2)
    plans = {"os": os_plan, "3d_tiny": make_plan(ctx, g)}
    jobs = [("os", make_job(torch, os_plan, g, 41, 3, 2)),
            ("3d_tiny", make_job(torch, plans["3d_tiny"], g, 42, 2, 2))]
    ref = [alone(plans[n], j) for n, j in jobs]
    torch.cuda.synchronize()
    assert np.nanmax(ref[0]) > 0
    res = run_batch(torch, ctx, plans, jobs)
    for k in range(len(jobs)):
        same_bits(res[k], ref[k])


def c3_specs():
    """One job of every geometry of the batched-jobs workload (2D 128x128 / 192x192,
    2D 256x256 / 344x252, 3D 128x128x47 / 192x192x233, 3D 256x256x81 /
    344x252x837), seeded, with uneven subsets and an attenuation sinogram on half."""
    from apirl_amd import workloads
    kinds = list(workloads.GEOMETRIES)
    subsets = {"2d_128": 5, "2d_256": 21, "3d_small": 7, "3d_c2": 21}
    iters = {"2d_128": 3, "2d_256": 2, "3d_small": 2, "3d_c2": 1}
    return [dict(seed=100 + k, kind=kind, subsets=subsets[kind], iterations=iters[kind],
                 counts=float(10 ** (5 + k)), attenuation=bool(k % 2))
            for k, kind in enumerate(kinds)]


def alone_spec(plans, spec, d):
    p = plans[spec["kind"]]
    img, _ = p.reconstruct(d["meas"], d["init"], spec["iterations"], subsets=spec["subsets"],
                           bp_mode=hip.Batch.backprojector(p), mult=d["mult"])
    return host(img)


def test_c3_every_job_type_batched_equals_alone(torch, ctx):
    """The batched-jobs workload's four job types in one batch (the 256x256x81 job on
    the single-job kernels after each grouped round, the others in the grouped
    launches): each image equals, bit for bit, the job run alone."""
    from apirl_amd import serving
    log("case: one job of each C3 geometry in one batch")
    specs = c3_specs()
    plans, jobs = serving.materialise_all(ctx, specs, torch)
    ref = [alone_spec(plans, s, d) for s, d in jobs]
    torch.cuda.synchronize()
    dt, choice = serving.run_batched(ctx, plans, jobs, torch, max_jobs=8, tune=True)
    log(f"  batched in {dt:.3f} s, autotuned {choice}")
    for k, (spec, d) in enumerate(jobs):
        log(f"  {spec['kind']}: S={spec['subsets']} T={spec['iterations']}")
        same_bits(host(d["out"]), ref[k])


@pytest.mark.parametrize("rate", [5.0, 200.0])
def test_c4_serving_job_types_at_two_rates(torch, ctx, rate):
    """Serving (open-loop arrivals, admission by HBM footprint) over the four job
    types at a low and a high arrival rate: every image equals the job alone."""
    from apirl_amd import serving
    log(f"case: serving the C3 job types at {rate} jobs/s")
    specs = c3_specs()
    plans, jobs = serving.materialise_all(ctx, specs, torch)
    ref = [alone_spec(plans, s, d) for s, d in jobs]
    torch.cuda.synchronize()
    lat, wall = serving.serve(ctx, plans, jobs, rate=rate, torch=torch, max_jobs=8)
    log(f"  p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s, "
        f"{len(jobs) / wall:.2f} jobs/s")
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    for k, (_, d) in enumerate(jobs):
        same_bits(host(d["out"]), ref[k])
