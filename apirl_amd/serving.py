"""Batched throughput and a serving loop over apirl_batch (continuous batching).

* run_sequential: the plain single-job path, one apirl_em_reconstruct per job, each
  with the backprojector the batch would run it with (Batch.backprojector).
* run_batched: all jobs submitted to one batch (up to its capacity, later ones as
  slots free up), rounds stepped until every job is done.
* autotune: picks the batch's persistent grid (workgroups per grouped launch), graph
  replay on or off, the virtual blocks (256 rays / voxels, or one gather tile) a
  workgroup draws at a time and the jobs' order in a round, by timing real rounds of
  the batch it tunes -- results do not depend on any of them, so the rounds spent
  tuning are not wasted.
* serve: open-loop Poisson arrivals; a job is admitted when the batch has a slot and
  its device state (apirl_batch_job_bytes) fits the HBM budget, joins the next
  round, and leaves when done; latency = completion - arrival."""
import time

import numpy as np

from . import hip, workloads


def _sync(torch):
    torch.cuda.synchronize()


def run_sequential(ctx, plans, jobs, torch):
    t0 = time.perf_counter()
    for spec, d in jobs:
        p = plans[spec["kind"]]
        p.reconstruct(d["meas"], d["init"], spec["iterations"], subsets=spec["subsets"],
                      bp_mode=hip.Batch.backprojector(p), mult=d["mult"], out=d["out"])
    _sync(torch)
    return time.perf_counter() - t0


def autotune(batch, grids=(512, 1024, 2048, 4096), grabs=(1, 2, 4, 8),
             orders=(hip.Batch.ORDER_SUBMISSION, hip.Batch.ORDER_LARGEST_FIRST), rounds=2):
    """Times `rounds` rounds per candidate on the live batch, one knob after the other
    (grid and graph replay, then the grab size, then the job order), each keeping the
    best so far; returns the choice."""
    state = {"grid": grids[0], "graph": True, "grab": 2, "order": orders[0]}
    best = [None]
    done = [False]

    def trial(**kw):
        cand = dict(state, **kw)
        batch.set_launch(cand["grid"], cand["graph"])
        batch.set_schedule(cand["grab"], cand["order"])
        batch.synchronize()
        t0 = time.perf_counter()
        active = 1
        for _ in range(rounds):
            active = batch.step()
        batch.synchronize()
        dt = (time.perf_counter() - t0) / rounds
        if best[0] is None or dt < best[0]:
            best[0] = dt
            state.update(cand)
        done[0] = active == 0

    for use_graph in (True, False):
        for grid in grids:
            if not done[0]:
                trial(grid=grid, graph=use_graph)
    for grab in grabs:
        if not done[0]:
            trial(grab=grab)
    for order in orders:
        if not done[0]:
            trial(order=order)
    batch.set_launch(state["grid"], state["graph"])
    batch.set_schedule(state["grab"], state["order"])
    return {"grid": state["grid"], "graph": state["graph"], "grab": state["grab"],
            "order": state["order"], "round_s": best[0]}


def run_batched(ctx, plans, jobs, torch, max_jobs=64, tune=True):
    batch = hip.Batch(ctx, max_jobs=max_jobs, max_sinograms=1024)
    pending = list(range(len(jobs)))
    ids = {}
    t0 = time.perf_counter()
    choice = None

    def admit():
        while pending and len(ids) - sum(1 for k in ids if done[k]) < max_jobs:
            k = pending.pop(0)
            spec, d = jobs[k]
            ids[k] = batch.submit(plans[spec["kind"]], d["meas"], d["init"], d["out"],
                                  spec["iterations"], spec["subsets"], mult=d["mult"])
            done[k] = False

    done = {}
    admit()
    if tune:
        choice = autotune(batch)
    while True:
        active = batch.step()
        if pending:
            for k in list(ids):
                if not done[k] and batch.done(ids[k]):
                    done[k] = True
                    batch.release(ids[k])
            admit()
        elif active == 0:
            break
    batch.synchronize()
    dt = time.perf_counter() - t0
    batch.close()
    return dt, choice


def serve(ctx, plans, jobs, rate, torch, max_jobs=64, hbm_budget=200e9, seed=0, grid=0,
          use_graph=True):
    """Open-loop arrivals at `rate` jobs/s of `jobs` (specs with materialised inputs).
    Returns latencies (s) in arrival order and the wall time of the run."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    batch = hip.Batch(ctx, max_jobs=max_jobs, max_sinograms=1024)
    batch.set_launch(grid, use_graph)
    need = [hip.Batch.job_bytes(plans[s["kind"]], s["subsets"]) for s, _ in jobs]
    if max(need) > hbm_budget:
        raise ValueError(f"a job needs {max(need)} bytes, above the HBM budget {hbm_budget}")
    queue, running = [], {}
    finished = np.full(len(jobs), np.nan)
    used = 0
    nxt = 0
    _sync(torch)
    t0 = time.perf_counter()
    while nxt < len(jobs) or queue or running:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        # admission: capacity and HBM footprint, in arrival order
        while queue and len(running) < max_jobs and used + need[queue[0]] <= hbm_budget:
            k = queue.pop(0)
            spec, d = jobs[k]
            running[k] = batch.submit(plans[spec["kind"]], d["meas"], d["init"], d["out"],
                                      spec["iterations"], spec["subsets"], mult=d["mult"])
            used += need[k]
        if running:
            batch.step()
            for k in list(running):
                if batch.done(running[k]):
                    finished[k] = time.perf_counter() - t0
                    batch.release(running.pop(k))
                    used -= need[k]
        elif nxt < len(jobs):
            time.sleep(max(0.0, min(arrivals[nxt] - now, 0.01)))
    batch.synchronize()
    wall = time.perf_counter() - t0
    batch.close()
    return finished - arrivals, wall


def materialise_all(ctx, specs, torch):
    plans = {k: workloads.make_plan(ctx, k) for k in sorted({s["kind"] for s in specs})}
    jobs = [(s, workloads.materialise(s, plans[s["kind"]], torch)) for s in specs]
    _sync(torch)
    return plans, jobs
