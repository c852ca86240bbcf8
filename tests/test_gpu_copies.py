"""Gauge copies in the spin form: one SU(2) configuration, several spin
fields swept together with the links read once per half-sweep.

Parity: each copy against the oracle's spin-form checkerboard sweep
(or_su2_spin_halfsweep, omega rule) and against the link-form job of the
same gauge copy, per-sweep functional to 1e-12 relative; final links
compared through the U(1)-invariant a0^2 + a3^2 of every link (the two
forms' outputs differ by a diagonal gauge)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 4, 8)
EPS, SEED, OMEGA = 0.5, 4242, 1.7


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _spin_job(mag, copies, n_max=3000, tol=0.0, dims=DIMS):
    return mag.MagJob(group=mag.SU2, dims=dims, form=mag.FORM_SPIN, copies=copies,
                      omega=OMEGA, tol=tol, max_sweeps=n_max)


def _u1_invariant(links):
    return links[:, 0] ** 2 + links[:, 3] ** 2


def test_copies_match_oracle_spin_sweeps(mag, oracle):
    n = 30
    with _spin_job(mag, 3, n_max=n) as job:
        job.generate(EPS, SEED)
        job.sweeps(n)
        hist = [job.copy_history(c)[0] for c in range(3)]
        spins = [job.copy_spins(c) for c in range(3)]
    U0 = oracle.gen_su2_config(DIMS, EPS, SEED)
    for c in range(3):
        s = oracle.gen_copy_spins(DIMS, SEED, c)
        ref = []
        for _ in range(n):
            oracle.su2_spin_halfsweep(U0, s, DIMS, 0, 2, OMEGA)
            oracle.su2_spin_halfsweep(U0, s, DIMS, 1, 2, OMEGA)
            ref.append(oracle.su2_functional_spin(U0, s, DIMS))
        assert len(hist[c]) == n
        assert _rel(hist[c], ref) < 1e-12, c
        assert np.max(np.abs(spins[c] - s)) < 1e-11, c


def test_streaming_spin_sweeps_match_oracle(mag, oracle):
    """SU(2) spin form at 48^4, FP64: the links (680 MB) exceed the 256 MB
    MALL and the half-sweeps load them with the nontemporal hint.  The
    sweeps equal the oracle's: functional per sweep to 1e-12, spins to
    1e-10 (the largest difference over 5.3M sites; the 8^4 test's is below
    1e-11)."""
    dims, n = (48, 48, 48, 48), 3
    print(f"[streaming spin] SU(2) {dims}, {n} sweeps", flush=True)
    with _spin_job(mag, 1, n_max=n, dims=dims) as job:
        job.generate(EPS, SEED)
        job.sweeps(n)
        hist = job.copy_history(0)[0]
        spins = job.copy_spins(0)
    U0 = oracle.gen_su2_config(dims, EPS, SEED)
    s = oracle.gen_copy_spins(dims, SEED, 0)
    ref = []
    for _ in range(n):
        oracle.su2_spin_halfsweep(U0, s, dims, 0, 2, OMEGA)
        oracle.su2_spin_halfsweep(U0, s, dims, 1, 2, OMEGA)
        ref.append(oracle.su2_functional_spin(U0, s, dims))
    assert _rel(hist, ref) < 1e-12
    assert np.max(np.abs(spins - s)) < 1e-10


def test_streaming_su3_site_fp32_gauge_equivalent(mag):
    """SU(3) site form, FP32 phase, at 40^3x24: FP32 links of 295 MB, past
    the MALL, loaded with the nontemporal hint.  After 20 sweeps the copy's
    links G U G^+ (gauges put on SU(3) in FP64) keep the input's plaquette
    to 1e-12, and the functional has gone up."""
    dims = (40, 40, 40, 24)
    print(f"[streaming site] SU(3) {dims}", flush=True)
    with mag.MagJob(group=mag.SU3, dims=dims, form=mag.FORM_SPIN, copies=1, omega=1.6, tol=0.0,
                    max_sweeps=40, precision=mag.MIXED, reunit_every=10, tol_mixed=0.0,
                    fp64_warmup=0) as j:
        j.generate(0.4, 45)
        P0 = j.plaquette()
        j.sweeps(20)
        assert j.fp64_start() < 0  # still in the FP32 phase
        F = j.copy_history(0)[0]
        links = j.download_copy(0)
    assert len(F) == 20 and F[-1] > F[0]
    assert abs(mag.plaquette(mag.SU3, dims, links) - P0) < 1e-12


def test_copies_match_link_form_copies(mag):
    n = 40
    with _spin_job(mag, 3, n_max=n) as job:
        job.generate(EPS, SEED)
        F0 = job.initial_functional()
        job.sweeps(n)
        hist = [job.copy_history(c)[0] for c in range(3)]
        links = [job.download_copy(c) for c in range(3)]
        P0 = job.plaquette()
    for c in range(3):
        with mag.MagJob(group=mag.SU2, dims=DIMS, omega=OMEGA, tol=0.0, max_sweeps=n) as lj:
            lj.generate(EPS, SEED, copy=c)
            if c == 0:
                assert abs(lj.initial_functional() - F0) < 1e-14
            lj.sweeps(n)
            F, _, _ = lj.history()
            ref = lj.download()
            assert abs(lj.plaquette() - P0) < 1e-13
        assert _rel(hist[c], F) < 1e-12, c
        assert np.max(np.abs(_u1_invariant(links[c]) - _u1_invariant(ref))) < 1e-11, c
        assert np.max(np.abs(np.linalg.norm(links[c], axis=1) - 1)) < 1e-13


def test_copy_trajectory_independent_of_set(mag, oracle):
    """Copy 2 of a 3-copy set and the same spins alone in a 1-copy job run
    the same arithmetic and partition: bit-identical histories."""
    n = 25
    with mag.MagJob(group=mag.SU2, dims=DIMS) as lj:
        lj.generate(EPS, SEED)
        U0 = lj.download()  # the same generator kernel as the spin job's links
    with _spin_job(mag, 3, n_max=n) as job:
        job.generate(EPS, SEED)
        start = job.copy_spins(2)
        assert np.max(np.abs(start - oracle.gen_copy_spins(DIMS, SEED, 2))) < 1e-13
        job.sweeps(n)
        F3, R3, A3 = job.copy_history(2)
        s3 = job.copy_spins(2)
    with _spin_job(mag, 1, n_max=n) as one:
        one.upload(U0)
        one.set_copy_spins(0, start)
        one.sweeps(n)
        F1, R1, A1 = one.copy_history(0)
        s1 = one.copy_spins(0)
    assert np.array_equal(F3, F1) and np.array_equal(R3, R1) and np.array_equal(A3, A1)
    assert np.array_equal(s3, s1)


def test_copies_converge_and_best_copy(mag):
    dims = (8, 8, 8, 8)
    with _spin_job(mag, 4, n_max=4000, tol=1e-12, dims=dims) as job:
        job.generate(EPS, SEED)
        assert job.run()
        st = [job.copy_status(c) for c in range(4)]
        best, Fb = job.best_copy()
    assert all(conv for _, conv, _ in st)
    assert Fb == max(F for _, _, F in st)
    assert st[best][2] == Fb
    # copy 1 alone in the link form reaches the same maximum
    with mag.MagJob(group=mag.SU2, dims=dims, omega=OMEGA, tol=1e-12, max_sweeps=4000) as lj:
        lj.generate(EPS, SEED, copy=1)
        assert lj.run()
        s, conv, F = lj.status()
    assert abs(F - st[1][2]) / F < 1e-10


def test_copy_sets_in_batch_equal_alone(mag):
    specs = [dict(form=mag.FORM_SPIN, copies=2, dims=(8, 8, 8, 8), seed=5),
             dict(form=mag.FORM_LINK, copies=1, dims=(8, 4, 8, 8), seed=6),
             dict(form=mag.FORM_SPIN, copies=3, dims=(4, 8, 4, 8), seed=7)]

    def make(sp):
        j = mag.MagJob(group=mag.SU2, dims=sp["dims"], form=sp["form"], copies=sp["copies"],
                       omega=OMEGA, tol=1e-12, max_sweeps=3000)
        j.generate(EPS, sp["seed"])
        return j

    alone = []
    for sp in specs:
        with make(sp) as j:
            j.run()
            alone.append([j.copy_history(c) for c in range(sp["copies"])])
    jobs = [make(sp) for sp in specs]
    with mag.MagBatch() as b:
        for j in jobs:
            b.add(j)
        b.run()
        for j, sp, ref in zip(jobs, specs, alone):
            for c in range(sp["copies"]):
                got = j.copy_history(c)
                for x, y in zip(got, ref[c]):
                    assert np.array_equal(x, y)
            b.remove(j)
    for j in jobs:
        j.close()


def test_mixed_fp64_warmup(mag):
    """The first fp64_warmup sweeps of a mixed job are the FP64 job's sweeps
    (same kernels); the FP32 phase follows; alone and in a batch agree."""
    dims = (8, 8, 8, 8)

    def make(prec, warm):
        j = mag.MagJob(group=mag.SU2, dims=dims, precision=prec, omega=OMEGA, tol=1e-12,
                       max_sweeps=3000, reunit_every=10, tol_mixed=1e-8, fp64_warmup=warm)
        j.generate(EPS, 31)
        return j

    with make(mag.F64, 0) as ref:
        ref.sweeps(40)
        Fr, _, _ = ref.history()
    with make(mag.MIXED, 40) as j:
        j.run()
        F, R, A = j.history()
        start = j.fp64_start()
        st = j.status()
    assert np.array_equal(F[:40], Fr[:40])
    assert start > 40 and st[1]
    with make(mag.MIXED, 40) as jb, mag.MagBatch() as b:
        b.add(jb)
        b.run()
        Fb, Rb, Ab = jb.history()
        assert np.array_equal(Fb, F) and np.array_equal(Ab, A)
        assert jb.fp64_start() == start
        b.remove(jb)


def test_spin_form_argument_errors(mag):
    with mag.MagJob(group=mag.SU3, dims=(4, 4, 4, 4), form=mag.FORM_SPIN, copies=2) as j3:
        j3.generate(0.3, 1)
        with pytest.raises(mag.MagError):  # SU(3) site variables are gauges, not spins
            j3.copy_spins(0)
    with pytest.raises(mag.MagError):
        mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4), form=mag.FORM_SPIN, copies=9)
    with pytest.raises(mag.MagError):
        mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4), form=mag.FORM_SPIN, precision=mag.MIXED,
                   fp64_warmup=50)
    with pytest.raises(mag.MagError):
        mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4), copies=2)
    with pytest.raises(mag.MagError):
        mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4), precision=mag.MIXED, fp64_warmup=30)
    with _spin_job(mag, 2, dims=(4, 4, 4, 4)) as j:
        j.generate(0.3, 1)
        with pytest.raises(mag.MagError):
            j.copy_status(2)
        with pytest.raises(mag.MagError):
            j.generate(0.3, 1, copy=1)


def test_spin_form_fp32_phase_reads(mag, oracle):
    """The FP32 phase of the spin form sweeps FP32 spins.  Spins and copy
    links read during that phase come from them: the spins (unit length in
    FP64) give the oracle's spin functional on the FP32-rounded links equal
    to the job's last recorded functional, and the copy links equal the
    oracle's gauge_tranformation_spins of the FP64 links with those spins
    (up to the diagonal U(1) gauge: the a0^2 + a3^2 of every link)."""
    dims = (8, 8, 8, 8)
    n = 30
    with mag.MagJob(group=mag.SU2, dims=dims, form=mag.FORM_SPIN, copies=2, precision=mag.MIXED,
                    omega=OMEGA, tol=0.0, max_sweeps=4 * n, tol_mixed=0.0,
                    fp64_warmup=0) as j:
        j.generate(EPS, 91)
        j.sweeps(n)
        assert j.fp64_start() < 0  # still in the FP32 phase
        for c in range(2):
            F = j.copy_history(c)[0]
            assert len(F) == n
            s = j.copy_spins(c)
            got = j.download_copy(c)
            links = oracle.gen_su2_config(dims, EPS, 91)
            assert np.max(np.abs(np.linalg.norm(s, axis=1) - 1)) < 1e-14
            links32 = links.astype(np.float32).astype(np.float64)
            Fo = oracle.su2_functional_spin(links32, s, dims)
            assert abs(Fo - F[-1]) / abs(F[-1]) < 1e-12, c
            ref = oracle.su2_gauge_transform_spins(links, s, dims)
            assert np.max(np.abs(_u1_invariant(got) - _u1_invariant(ref))) < 1e-12, c


def test_spin_form_mixed_reaches_fp64_maximum(mag):
    """Mixed spin form: FP64 warm-up, then sweeps that read FP32-rounded
    links (spins and arithmetic FP64), then FP64 polish.  Every copy ends at
    the FP64 run's maximum; alone and in a batch agree bit for bit."""
    dims = (16, 16, 8, 8)

    def make(prec):
        j = mag.MagJob(group=mag.SU2, dims=dims, form=mag.FORM_SPIN, copies=3, precision=prec,
                       omega=OMEGA, tol=1e-12, max_sweeps=4000, tol_mixed=1e-12,
                       fp64_warmup=100)
        j.generate(EPS, 77)
        return j

    with make(mag.F64) as ref:
        assert ref.run()
        Fr = [ref.copy_status(c)[2] for c in range(3)]
    with make(mag.MIXED) as j:
        assert j.run()
        st = [j.copy_status(c) for c in range(3)]
        hist = [j.copy_history(c) for c in range(3)]
        start = j.fp64_start()
    assert start > 100
    for c in range(3):
        assert st[c][1]
        assert abs(st[c][2] - Fr[c]) / Fr[c] < 1e-10, c
    with make(mag.MIXED) as jb, mag.MagBatch() as b:
        b.add(jb)
        b.run()
        for c in range(3):
            for x, y in zip(jb.copy_history(c), hist[c]):
                assert np.array_equal(x, y)
        assert jb.fp64_start() == start
        b.remove(jb)


# ------------------------------------------------------------ SU(3) site form
SU3_DIMS = (4, 4, 4, 8)


def test_su3_site_copies_match_link_form(mag, oracle):
    """SU(3) site form: copy c (site gauges G_c on the links of copy 0)
    follows the link-form job of gauge copy c: F per sweep to 1e-12, links
    G_c U G_c^+ to 1e-11; and the link-form job itself follows the oracle."""
    n = 30
    with mag.MagJob(group=mag.SU3, dims=SU3_DIMS, form=mag.FORM_SPIN, copies=3, omega=1.6,
                    tol=0.0, max_sweeps=n) as job:
        job.generate(0.4, 606)
        F0 = [job.copy_history(c) for c in range(3)]
        assert all(len(h[0]) == 0 for h in F0)
        job.sweeps(n)
        hist = [job.copy_history(c)[0] for c in range(3)]
        links = [job.download_copy(c) for c in range(3)]
    for c in range(3):
        with mag.MagJob(group=mag.SU3, dims=SU3_DIMS, omega=1.6, tol=0.0, max_sweeps=n) as lj:
            lj.generate(0.4, 606, copy=c)
            lj.sweeps(n)
            F, _, _ = lj.history()
            ref = lj.download()
        assert _rel(hist[c], F) < 1e-12, c
        assert np.max(np.abs(links[c] - ref)) < 1e-11, c
    # oracle on copy 2's links
    W = oracle.gen_su3_config(SU3_DIMS, 0.4, 606, copy=2)
    Fo = [oracle.su3_link_sweep(W, SU3_DIMS, 1.6)[0] for _ in range(n)]
    assert _rel(hist[2], Fo) < 1e-12
    assert np.max(np.abs(links[2] - W)) < 1e-11


def test_su3_site_fp32_phase_cap_mid_period(mag):
    """The FP32 phase of the SU(3) site form keeps its sweeps' partials in
    slots reduced once per period of reunit_every sweeps.  With the sweep
    cap inside a period (no warm-up, periods ending at 10, 20, ..., 50, cap
    45) every copy stops exactly at the cap, and run(), sweeps in chunks
    that end mid-period, and a batch record bit-identical histories."""
    dims = (8, 8, 8, 8)

    def make():
        j = mag.MagJob(group=mag.SU3, dims=dims, form=mag.FORM_SPIN, copies=2, omega=1.6,
                       tol=0.0, max_sweeps=45, precision=mag.MIXED, reunit_every=10,
                       tol_mixed=0.0, fp64_warmup=0)
        j.generate(0.4, 808)
        return j

    with make() as a:
        a.run()
        ref = [a.copy_history(c) for c in range(2)]
        for c in range(2):
            sw, conv, _ = a.copy_status(c)
            assert sw == 45 and not conv
            assert len(ref[c][0]) == 45
            assert np.all(np.isfinite(ref[c][0]))
    with make() as b:
        for _ in range(7):
            b.sweeps(7)  # 49 sweeps asked for; the cap holds at 45
        for c in range(2):
            assert b.copy_status(c)[0] == 45
            for x, y in zip(b.copy_history(c), ref[c]):
                assert np.array_equal(x, y)
    with make() as d, mag.MagBatch() as bt:
        bt.add(d)
        bt.run()
        for c in range(2):
            assert d.copy_status(c)[0] == 45
            for x, y in zip(d.copy_history(c), ref[c]):
                assert np.array_equal(x, y)
        bt.remove(d)


@pytest.mark.parametrize("prec", [0, 1])
def test_su3_site_copies_batch_equals_alone(mag, prec):
    def make():
        j = mag.MagJob(group=mag.SU3, dims=SU3_DIMS, form=mag.FORM_SPIN, copies=2, omega=1.6,
                       tol=1e-11, max_sweeps=3000, precision=prec, tol_mixed=1e-11,
                       fp64_warmup=40)
        j.generate(0.4, 707)
        return j

    with make() as a:
        a.run()
        ref = [a.copy_history(c) for c in range(2)]
        st = [a.copy_status(c) for c in range(2)]
        start = a.fp64_start()
    other = mag.MagJob(group=mag.SU2, dims=(8, 8, 8, 8), omega=1.7, tol=1e-12, max_sweeps=3000)
    other.generate(0.5, 3)
    with make() as b, mag.MagBatch() as bt:
        bt.add(b)
        bt.add(other)
        bt.run()
        for c in range(2):
            for x, y in zip(b.copy_history(c), ref[c]):
                assert np.array_equal(x, y)
        assert b.fp64_start() == start
        bt.remove(b)
        bt.remove(other)
    other.close()
    if prec == 1:
        with mag.MagJob(group=mag.SU3, dims=SU3_DIMS, form=mag.FORM_SPIN, copies=2, omega=1.6,
                        tol=1e-11, max_sweeps=3000) as f:
            f.generate(0.4, 707)
            f.run()
            for c in range(2):
                Ff = f.copy_status(c)[2]
                assert abs(st[c][2] - Ff) / Ff < 1e-9
