"""Link-form output in the reference's convention: with track_gauge the SU(2)
link form accumulates the gauge G(x) of the fixing, and
mag_job_download_links_canonical returns G_c(x) U_in G_c(x+mu)^+ with
G_c = GetGaugeMatrix(s(x)) (matrix.h:960-963), s(x) the U(1)-invariant
spin of G(x): what gauge_tranformation_spins (mag.cpp:1045-1068) writes for
those spins.  The link form and the spin form run the same omega step, so
the canonical links after n sweeps equal the oracle's spin-form sweeps of
spins starting at e3, gauge-transformed by the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 4, 8)


def test_canonical_links_equal_spin_form_of_the_oracle(mag, oracle):
    links = oracle.gen_su2_config(DIMS, 0.4, 61)
    V = int(np.prod(DIMS))
    n, omega = 12, 1.7
    spins = np.tile([0.0, 0.0, 1.0], (V, 1))
    for _ in range(n):
        oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 2, omega)
        oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 2, omega)
    ref = oracle.su2_gauge_transform_spins(links, spins, DIMS)
    with mag.MagJob(group=mag.SU2, dims=DIMS, omega=omega, tol=0.0, max_sweeps=n,
                    track_gauge=1) as job:
        job.upload(links)
        job.sweeps(n)
        s = job.download_spins()
        can = job.download_canonical()
        plain = job.download()
    assert np.max(np.abs(s - spins)) < 1e-11
    assert np.max(np.abs(can - ref)) < 1e-11
    # the plain download differs from the canonical one by a diagonal U(1)
    # gauge only: same functional, and not the same links
    assert abs(oracle.su2_functional(can, DIMS) - oracle.su2_functional(plain, DIMS)) < 1e-13
    assert np.max(np.abs(can - plain)) > 1e-3


@pytest.mark.parametrize("precision,warmup", [(0, 0), (1, 0), (1, 40)])
def test_canonical_links_after_a_run(mag, oracle, precision, warmup):
    """To convergence, FP64 or mixed (the FP32 phase's red gauges and the
    recovered black gauges join the tracked total at its end): the canonical
    links are the spin-form transform of the input by the downloaded spins,
    with the run's functional and the input's plaquette."""
    links = oracle.gen_su2_config(DIMS, 0.5, 62)
    with mag.MagJob(group=mag.SU2, dims=DIMS, omega=1.7, tol=1e-12, max_sweeps=4000,
                    precision=precision, reunit_every=10, tol_mixed=1e-11, fp64_warmup=warmup,
                    track_gauge=1) as job:
        job.upload(links)
        assert job.run()
        _, _, F = job.status()
        s = job.download_spins()
        can = job.download_canonical()
    assert np.max(np.abs(np.linalg.norm(s, axis=1) - 1)) < 1e-12
    ref = mag.gauge_transform_spins(DIMS, links, s)
    assert np.max(np.abs(can - ref)) < 1e-11
    assert abs(oracle.su2_functional(can, DIMS) - F) < 1e-12
    assert abs(oracle.su2_plaquette(can, DIMS) - oracle.su2_plaquette(links, DIMS)) < 1e-12


def test_track_gauge_arguments(mag):
    with pytest.raises(mag.MagError):
        mag.MagJob(group=mag.SU3, dims=(4, 4, 4, 4), track_gauge=1)
    with mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4)) as job:
        job.generate(0.3, 1)
        with pytest.raises(mag.MagError):
            job.download_canonical()  # no tracking
    with mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4), track_gauge=1) as job, \
            mag.MagBatch() as batch:
        job.generate(0.3, 1)
        with pytest.raises(mag.MagError):
            batch.add(job)


def test_tracked_gauge_every_shape(mag, oracle):
    """Every launch shape accumulates the same gauge (the FP64 half-sweeps
    with the tracked gauge, one or two tiles per thread): spins and
    canonical links bit-identical across shapes and equal to the oracle's
    spin-form sweeps."""
    links = oracle.gen_su2_config(DIMS, 0.4, 63)
    V = int(np.prod(DIMS))
    n, omega = 8, 1.7
    spins = np.tile([0.0, 0.0, 1.0], (V, 1))
    for _ in range(n):
        oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 2, omega)
        oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 2, omega)
    ref = None
    try:
        for nt, spt in mag.SHAPES:
            print(f"[canonical] shape {nt}x{spt}", flush=True)
            mag.tune_set(mag.TUNE_FIXED, nt, spt)
            with mag.MagJob(group=mag.SU2, dims=DIMS, omega=omega, tol=0.0, max_sweeps=n,
                            track_gauge=1) as job:
                job.upload(links)
                job.sweeps(n)
                got = (job.download_spins(), job.download_canonical(), job.download())
                assert job.shape() == (nt, spt)
            assert np.max(np.abs(got[0] - spins)) < 1e-11
            if ref is None:
                ref = got
            for x, y in zip(got, ref):
                assert np.array_equal(x, y), (nt, spt)
    finally:
        mag.tune_set(mag.TUNE_AUTO)
