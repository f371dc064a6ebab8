"""Served step-job stream: each result is the same alone and in the mix.

af_serve admits jobs as the pool frees up, steps every resident job with the
batched step graph (the resident set changes between steps while earlier
steps still run) and records each job's final MaxDivergence and
KineticEnergy.  Every job is re-run alone (New, the same device initial
condition, StepN(N)) and both numbers must match bit for bit; the pool is
sized so that jobs queue, and the arrivals overlap running work.
"""
import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import workload

pytestmark = pytest.mark.gpu


def small_stream(n, seed, rate):
    jobs = workload.served_jobs(n, seed, rate=rate, sizes=(8, 12, 16, 24, 32), steps=(1, 6))
    return jobs


def alone(af, prec, j):
    e = af.New(af.Config(NX=j["NX"], NY=j["NY"], NZ=j["NZ"], H=j["H"], DT=j["DT"],
                         Viscosity=j["Viscosity"], Precision=prec))
    e.init_band_limited(j["Seed"])
    e.StepN(j["Steps"])
    out = (e.MaxDivergence(), e.KineticEnergy())
    e.Close()
    return out


@pytest.mark.parametrize("policy", [afm.SERVE_FCFS, afm.SERVE_LEAST_WORK_PREEMPT,
                                    afm.SERVE_LEAST_WORK_STEPS, afm.SERVE_LEAST_WORK])
@pytest.mark.parametrize("prec", [afm.PRECISION_FP64, afm.PRECISION_MIXED])
@pytest.mark.parametrize("realtime", [False, True])
@pytest.mark.parametrize("group_bytes", [0, 40000])
def test_served_results_equal_alone(af, prec, realtime, policy, group_bytes):
    jobs = small_stream(14, 5 + prec, rate=200.0)
    biggest = max(af.SlotBytes(af.Config(NX=j["NX"], NY=j["NY"], NZ=j["NZ"], Precision=prec),
                               prec) for j in jobs)
    # policy 2 with a budget of about two jobs' cells: the resident set is
    # stepped a few jobs at a time, and the subset changes between steps.
    # 40 kB solve groups: most resident jobs solve in a group of their own,
    # and the group count changes from step to step.
    b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=3 * biggest, MaxSlots=16,
                                ServePolicy=policy, StepCellBudget=2 * 24 ** 3,
                                SolveGroupBytes=group_bytes))
    res, st = af.Serve(b, jobs, realtime=realtime)
    assert st["Steps"] > 0 and 1 <= st["MaxResident"] <= len(jobs)
    assert st["PeakBytes"] <= 3 * biggest
    assert st["CellUpdates"] == sum(j["NX"] * j["NY"] * j["NZ"] * j["Steps"] for j in jobs)
    assert b.info()["Slots"] == 0 and b.info()["BytesInUse"] == 0
    for j, r in zip(jobs, res):
        assert r["Status"] == afm.AF_OK
        assert r["Done"] >= r["Admitted"] >= 0
        if realtime:
            assert r["Admitted"] >= j["Arrival"]
        d, ke = alone(af, prec, j)
        assert r["MaxDivergence"] == d, j
        assert r["KineticEnergy"] == ke, j
    b.Close()


@pytest.mark.parametrize("prec", [afm.PRECISION_FP64, afm.PRECISION_MIXED])
def test_served_staged_sweep_results_equal_alone(af, prec):
    """Jobs 64 and 96 cells wide run the LDS-staged sweep beside narrow ones
    on the gather; as jobs arrive the staged work-list outgrows its grid and
    the step graph is rebuilt between steps.  Every result equals the job
    run alone."""
    jobs = workload.served_jobs(10, 11 + prec, rate=400.0, sizes=(16, 64, 96), steps=(1, 4))
    assert any(j["NX"] >= 64 for j in jobs) and any(j["NX"] < 64 for j in jobs)
    b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=1 << 30, MaxSlots=16,
                                SweepStaged=2))
    res, st = af.Serve(b, jobs, realtime=True)
    assert st["CellUpdates"] == sum(j["NX"] * j["NY"] * j["NZ"] * j["Steps"] for j in jobs)
    assert b.info()["SweepStaged"] == 1
    for j, r in zip(jobs, res):
        assert r["Status"] == afm.AF_OK
        d, ke = alone(af, prec, j)
        assert r["MaxDivergence"] == d, j
        assert r["KineticEnergy"] == ke, j
    b.Close()


@pytest.mark.parametrize("prec", [afm.PRECISION_FP64, afm.PRECISION_MIXED])
def test_preempted_jobs_resume_with_the_same_results(af, prec):
    """Least-remaining-work-first with preemption: two long 32^3 jobs fill
    the pool; short 8^3 jobs arriving while they run cannot be placed, so a
    long job is snapshotted to host memory and requeued, the short jobs run,
    and the long job resumes from its snapshot.  Every result equals the job
    run alone, and a short job finishes before either long job does (under
    first-come-first-served none could start until a long job ended).  Later
    short jobs may wait for a slot: each job is preempted at most once, so
    once both long jobs have been, an arrival waits for one to finish."""
    big = dict(NX=32, NY=32, NZ=32, H=1 / 32, Viscosity=1e-4, DT=0.002, Steps=600)
    jobs = [dict(big, Seed=1, Arrival=0.0), dict(big, Seed=2, Arrival=0.0)]
    for i in range(6):
        jobs.append(dict(NX=8, NY=8, NZ=8, H=1 / 8, Viscosity=1e-4, DT=0.004, Steps=3,
                         Seed=10 + i, Arrival=0.03 + 0.002 * i))
    pool = 2 * af.SlotBytes(af.Config(NX=32, NY=32, NZ=32, Precision=prec), prec)
    b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=pool, MaxSlots=16,
                                ServePolicy=afm.SERVE_LEAST_WORK_PREEMPT))
    res, st = af.Serve(b, jobs, realtime=True)
    assert st["Preemptions"] >= 1
    assert st["CellUpdates"] == sum(j["NX"] * j["NY"] * j["NZ"] * j["Steps"] for j in jobs)
    assert b.info()["Slots"] == 0 and b.info()["BytesInUse"] == 0
    for j, r in zip(jobs, res):
        assert r["Status"] == afm.AF_OK
        d, ke = alone(af, prec, j)
        assert r["MaxDivergence"] == d, j
        assert r["KineticEnergy"] == ke, j
    assert min(r["Done"] for r in res[2:]) < min(r["Done"] for r in res[:2])
    b.Close()


@pytest.mark.parametrize("group_bytes", [0, 40000])
def test_checked_served_jobs_meet_their_targets(af, group_bytes):
    """A checked mixed batch serves jobs that each carry a max|div| target
    (here the job's own final max|div| on the FP64 path): every job ends at
    or below it, and equals the same checked engine run alone, with one
    solve group or several."""
    jobs = small_stream(10, 13, rate=300.0)
    fix = dict(MixedSweeps=24, MixedFixSweeps=1, MixedMaxCycles=2)
    for j in jobs:
        j["MixedTarget"] = alone(af, afm.PRECISION_FP64, j)[0]
    prec = afm.PRECISION_MIXED
    b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=256 << 20, MaxSlots=16,
                                MixedChecked=1, SolveGroupBytes=group_bytes, **fix))
    res, st = af.Serve(b, jobs, realtime=True)
    assert st["CellUpdates"] == sum(j["NX"] * j["NY"] * j["NZ"] * j["Steps"] for j in jobs)
    b.Close()
    for j, r in zip(jobs, res):
        assert r["Status"] == afm.AF_OK
        assert r["MaxDivergence"] <= j["MixedTarget"] * (1 + 1e-9), j
        e = af.New(af.Config(NX=j["NX"], NY=j["NY"], NZ=j["NZ"], H=j["H"], DT=j["DT"],
                             Viscosity=j["Viscosity"], Precision=prec,
                             MixedTarget=j["MixedTarget"], **fix))
        e.init_band_limited(j["Seed"])
        e.StepN(j["Steps"])
        assert (r["MaxDivergence"], r["KineticEnergy"]) == (e.MaxDivergence(), e.KineticEnergy())
        e.Close()


def test_least_work_stepping_finishes_short_jobs_first(af):
    """Policy 2 with a budget of one job's cells: each batched step advances
    only the resident job with the least remaining work, so four jobs of 8,
    2, 5 and 1 steps queued together take 16 batched steps and finish in
    the order 1, 2, 5, 8 -- and each result equals the job run alone."""
    jobs = [dict(NX=16, NY=16, NZ=16, H=1 / 16, Viscosity=1e-4, DT=0.004, Steps=s, Seed=30 + i,
                 Arrival=0.0) for i, s in enumerate((8, 2, 5, 1))]
    b = af.Batch(af.BatchConfig(MemoryBudget=64 << 20, MaxSlots=8,
                                ServePolicy=afm.SERVE_LEAST_WORK_STEPS, StepCellBudget=16 ** 3))
    res, st = af.Serve(b, jobs, realtime=False)
    assert st["Steps"] == 16 and st["MaxResident"] == 4
    assert res[3]["Done"] <= res[1]["Done"] <= res[2]["Done"] <= res[0]["Done"]
    # batched step indices: 1 step at 0, 2 steps at 1-2, 5 at 3-7, 8 at 8-15
    assert [(r["FirstStep"], r["DoneStep"]) for r in res] == [(8, 15), (1, 2), (3, 7), (0, 0)]
    # after the run every slot advances again
    s = b.Add(af.Config(NX=8, NY=8, NZ=8))
    b.StepN(1)
    b.Remove(s)
    b.Close()
    for j, r in zip(jobs, res):
        assert r["Status"] == afm.AF_OK
        assert (r["MaxDivergence"], r["KineticEnergy"]) == alone(af, afm.PRECISION_FP64, j)


def test_step_max_wait_bounds_how_long_a_job_is_passed_over(af):
    """The same four jobs (8, 2, 5 and 1 steps, a budget of one job per
    batched step) with StepMaxWait = 2: a job passed over by two consecutive
    steps goes first in the next (longest wait first, ties by arrival), so
    the 8-step job first advances at step 2 instead of 8 and the 5-step job
    at step 3; the run still takes 16 batched steps, with every result equal
    to the job run alone."""
    jobs = [dict(NX=16, NY=16, NZ=16, H=1 / 16, Viscosity=1e-4, DT=0.004, Steps=s, Seed=30 + i,
                 Arrival=0.0) for i, s in enumerate((8, 2, 5, 1))]
    for policy in (afm.SERVE_LEAST_WORK_STEPS, afm.SERVE_LEAST_WORK):
        b = af.Batch(af.BatchConfig(MemoryBudget=64 << 20, MaxSlots=8, ServePolicy=policy,
                                    StepCellBudget=16 ** 3, StepMaxWait=2))
        res, st = af.Serve(b, jobs, realtime=False)
        b.Close()
        assert st["Steps"] == 16 and st["MaxResident"] == 4
        assert res[3]["FirstStep"] == res[3]["DoneStep"] == 0
        assert res[0]["FirstStep"] == 2
        assert all(0 <= r["FirstStep"] <= 3 for r in res)
        assert max(r["DoneStep"] for r in res) == 15
        for j, r in zip(jobs, res):
            assert r["Status"] == afm.AF_OK
            assert (r["MaxDivergence"], r["KineticEnergy"]) == alone(af, afm.PRECISION_FP64, j)


def test_serve_progress_reports(af):
    """The progress callback runs between batched steps with the running
    statistics: steps and completed jobs never decrease and stay within the
    run's totals."""
    jobs = small_stream(8, 3, rate=500.0)
    b = af.Batch(af.BatchConfig(MemoryBudget=64 << 20, MaxSlots=16))
    seen = []
    res, st = af.Serve(b, jobs, realtime=False, progress=lambda s, d: seen.append((s, d)),
                       progress_interval=0.0)
    assert seen and len(seen) == st["Steps"]
    steps = [s["Steps"] for s, _ in seen]
    done = [d for _, d in seen]
    assert steps == sorted(steps) and done == sorted(done)
    assert steps[-1] == st["Steps"] and done[-1] <= len(jobs)
    assert all(s["CellUpdates"] <= st["CellUpdates"] for s, _ in seen)
    assert all(r["Status"] == afm.AF_OK for r in res)
    b.Close()


def test_served_job_larger_than_pool_is_refused(af):
    jobs = small_stream(4, 9, rate=1000.0)
    jobs[2] = dict(jobs[2], NX=64, NY=64, NZ=64, H=1 / 64)
    need = max(af.SlotBytes(af.Config(NX=j["NX"], NY=j["NY"], NZ=j["NZ"]), 0) for j in jobs
               if j["NX"] < 64)
    b = af.Batch(af.BatchConfig(MemoryBudget=2 * need))
    res, _ = af.Serve(b, jobs, realtime=False)
    assert [r["Status"] for r in res] == [0, 0, afm.AF_ERR_OOM, 0]
    for i in (0, 1, 3):
        d, ke = alone(af, afm.PRECISION_FP64, jobs[i])
        assert (res[i]["MaxDivergence"], res[i]["KineticEnergy"]) == (d, ke)
    b.Close()


def test_serve_rejects_bad_input(af):
    b = af.Batch(af.BatchConfig(MemoryBudget=16 << 20))
    jobs = small_stream(3, 1, rate=100.0)
    with pytest.raises(af.AeroFluxError):
        af.Serve(b, list(reversed(jobs)))  # unsorted arrivals
    with pytest.raises(af.AeroFluxError):
        af.Serve(b, [dict(jobs[0], Steps=0)])
    s = b.Add(af.Config(NX=8, NY=8, NZ=8))
    with pytest.raises(af.AeroFluxError):
        af.Serve(b, jobs)  # batch not empty
    b.Remove(s)
    res, st = af.Serve(b, [])
    assert res == [] and st["Steps"] == 0
    b.Close()
    # a checked batch needs every job's target
    b = af.Batch(af.BatchConfig(Precision=afm.PRECISION_MIXED, MemoryBudget=16 << 20,
                                MixedChecked=1))
    with pytest.raises(af.AeroFluxError, match="MixedTarget"):
        af.Serve(b, jobs)
    b.Close()
