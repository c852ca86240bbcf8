"""Batched, graph-captured sweeps: every job's results are bit-identical to
the same job run alone (the batch runs each job's own partition and
reductions), mixed SU(2)/SU(3), FP64 and mixed precision, uneven sizes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [
    # group, dims, precision, omega, tol, eps, seed
    (2, (8, 8, 8, 8), 1, 1.7, 1e-12, 0.3, 11),
    (2, (4, 8, 4, 8), 1, 1.7, 1e-12, 0.5, 12),
    (3, (4, 4, 4, 8), 1, 1.6, 1e-11, 0.4, 13),
    (2, (8, 8, 8, 8), 0, 1.7, 1e-12, 0.3, 14),
    (3, (4, 4, 4, 4), 0, 1.6, 1e-11, 0.4, 15),
]


def _job(mag, spec):
    group, dims, prec, omega, tol, eps, seed = spec
    job = mag.MagJob(group=group, dims=dims, precision=prec, omega=omega, tol=tol,
                     max_sweeps=3000, reunit_every=10, tol_mixed=1e-8)
    job.generate(eps, seed)
    return job


def test_batch_equals_alone(mag):
    alone = []
    for spec in SPECS:
        with _job(mag, spec) as job:
            job.run()
            F, R, A = job.history()
            alone.append((F, A, job.download(), job.status(), job.fp64_start()))
    batch = mag.MagBatch()
    jobs = [_job(mag, s) for s in SPECS]
    for j in jobs:
        batch.add(j)
    batch.run()
    for j, (F, A, links, status, start) in zip(jobs, alone):
        Fb, Rb, Ab = j.history()
        assert np.array_equal(Fb, F)
        assert np.array_equal(Ab, A)
        assert np.array_equal(j.download(), links)
        assert j.status() == status
        assert j.fp64_start() == start
        assert status[1]  # converged
    batch.close()
    for j in jobs:
        j.close()


def test_batch_fp64_matches_oracle(mag, oracle):
    dims = (8, 8, 4, 8)
    links = oracle.gen_su2_config(dims, 0.4, 3)
    ref = links.copy()
    F_ref = [oracle.su2_link_sweep(ref, dims, 1.7)[0] for _ in range(40)]
    with mag.MagBatch() as batch:
        job = mag.MagJob(group=mag.SU2, dims=dims, omega=1.7, tol=0.0, max_sweeps=40)
        job.upload(links)
        batch.add(job)
        batch.replays(2)  # 2 x 20 sweeps, then the cap
        batch.poll()
        F, _, _ = job.history()
        got = job.download()
        batch.remove(job)
        job.close()
    assert np.max(np.abs(np.array(F) - F_ref) / np.abs(F_ref)) < 1e-12
    assert np.max(np.abs(got - ref)) < 1e-12


def test_batched_job_rejects_direct_sweeps(mag):
    with mag.MagBatch() as batch:
        job = mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4))
        job.generate(0.3, 1)
        batch.add(job)
        with pytest.raises(mag.MagError) as e:
            job.sweeps(1)
        assert e.value.code == -7
        batch.remove(job)
        job.sweeps(1)
        job.close()
