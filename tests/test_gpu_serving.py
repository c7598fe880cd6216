"""GPU: serving admission by device memory and overlap of admitted jobs.

* With a memory budget of 1.5 job footprints, four equal jobs submitted
  together are admitted one at a time as memory frees up: none is rejected,
  every result equals the job run alone, the serving workers never hold
  more than the budget (peak of tpcf_ctx_memory), and each job's time on
  the device (admission to completion) starts after the previous one ended.
* With room for all, three equal jobs submitted together are admitted at
  once by three workers and run at the same time: their admission-to-done
  intervals overlap, and the results equal the jobs run alone.
"""
import importlib

import numpy as np
import pytest

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")
pytestmark = pytest.mark.gpu


def jobs_of(ctx, k, n, seed):
    keep, jobs = [], []
    for j in range(k):
        d = synth.neyman_scott(n, nclust=20, sigma=2.0, size=150.0, seed=seed + 2 * j)
        r = synth.uniform_cube(n, 150.0, seed=seed + 2 * j + 1)
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        keep += [cd, cr]
        jobs.append(tp.four_term_job(cd, cr))
    return jobs, keep


def intervals(stats):
    # submit times are within microseconds of each other; admission and
    # completion relative to each job's own submit
    return sorted((s.queue_ms, s.wall_ms) for s in stats)


def test_queued_admission_under_memory_budget():
    print("case serving: budget of 1.5 footprints, four jobs", flush=True)
    with tp.Context(0) as ctx:
        jobs, keep = jobs_of(ctx, 4, 6000, seed=700)
        spec = tp.r123_spec(20, 1.0, 150.0, "log")
        fp = max(ctx.footprint(j, spec) for j in jobs)
        budget = int(1.5 * fp)
        ctx.set_memory_budget(budget)
        tickets = [ctx.submit(j, spec) for j in jobs]
        stats = [ctx.wait(t) for t, _ in tickets]
        mem = ctx.memory()
        print(f"  footprint {fp}, budget {budget}, peak {mem['peak']}, queue ms "
              f"{[round(s.queue_ms) for s in stats]}", flush=True)
        assert mem["budget"] == budget
        assert 0 < mem["peak"] <= budget
        iv = intervals(stats)
        for (a0, d0), (a1, d1) in zip(iv, iv[1:]):
            assert a1 >= d0 - 1.0, iv  # admitted after the previous job ended
        assert iv[-1][0] > 0
        for (_, out), job in zip(tickets, jobs):
            alone, _ = ctx.count(job, spec)
            assert np.array_equal(out, alone)


def test_admitted_jobs_run_at_the_same_time():
    print("case serving: three jobs at once", flush=True)
    with tp.Context(0) as ctx:
        jobs, keep = jobs_of(ctx, 3, 12000, seed=800)
        spec = tp.r123_spec(20, 1.0, 150.0, "log")
        tickets = [ctx.submit(j, spec) for j in jobs]
        stats = [ctx.wait(t) for t, _ in tickets]
        iv = intervals(stats)
        print(f"  (admit, done) ms {[(round(a), round(b)) for a, b in iv]}", flush=True)
        # every job was admitted before any of the others finished
        assert max(a for a, _ in iv) < min(d for _, d in iv)
        mem = ctx.memory()
        assert mem["peak"] <= mem["budget"]
        for (_, out), job in zip(tickets, jobs):
            alone, _ = ctx.count(job, spec)
            assert np.array_equal(out, alone)
