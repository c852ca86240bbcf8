"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 link-form sweeps: the device kernels (built without floating-point
contraction, with the omega rule's angle functions as a fixed operation
sequence) and the oracle perform the same IEEE operations site by site, so
the swept links are bit-identical (asserted: max |dU| == 0).  The
functional per sweep differs only by the order of the sum over sites:
1e-12 relative (measured about 1e-13).  Gauge invariants 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SU2_DIMS = (8, 8, 8, 8)
SU3_DIMS = (4, 4, 4, 8)


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_su2_generator_matches_oracle(mag, oracle):
    dims = (4, 6, 4, 8)
    ref = oracle.gen_su2_config(dims, 0.3, 20240601)
    with mag.MagJob(group=mag.SU2, dims=dims) as job:
        job.generate(0.3, 20240601)
        got = job.download()
    assert np.max(np.abs(got - ref)) < 1e-13


def test_su3_generator_matches_oracle(mag, oracle):
    dims = (4, 4, 4, 6)
    ref = oracle.gen_su3_config(dims, 0.4, 77)
    with mag.MagJob(group=mag.SU3, dims=dims) as job:
        job.generate(0.4, 77)
        got = job.download()
    assert np.max(np.abs(got - ref)) < 1e-13


@pytest.mark.parametrize("group,copy", [(2, 1), (2, 3), (3, 2)])
def test_gauge_copy_generator_matches_oracle(mag, oracle, group, copy):
    dims = (4, 4, 6, 4)
    gen = oracle.gen_su2_config if group == 2 else oracle.gen_su3_config
    ref = gen(dims, 0.4, 5150, copy=copy)
    with mag.MagJob(group=group, dims=dims) as job:
        job.generate(0.4, 5150, copy=copy)
        got = job.download()
    assert np.max(np.abs(got - ref)) < 1e-13


# omega = 2 (pure reflection) conserves F and is not contracting: rounding
# differences grow about 3x per sweep, so the trajectory is compared over a
# few sweeps only (F conservation is checked over many below).
@pytest.mark.parametrize("omega,n", [(1.0, 40), (1.7, 40), (2.0, 3)])
def test_su2_fp64_sweeps_match_oracle(mag, oracle, omega, n):
    dims = SU2_DIMS
    links = oracle.gen_su2_config(dims, 0.3, 1234)
    ref = links.copy()
    F_ref, R_ref, A_ref = [], [], []
    for _ in range(n):
        F, mx, av = oracle.su2_link_sweep(ref, dims, omega)
        F_ref.append(F)
        R_ref.append(mx)
        A_ref.append(av)
    with mag.MagJob(group=mag.SU2, dims=dims, omega=omega, tol=0.0, max_sweeps=n) as job:
        job.upload(links)
        assert abs(job.initial_functional() - oracle.su2_functional(links, dims)) < 1e-14
        job.sweeps(n)
        F, R, A = job.history()
        got = job.download()
        P = job.plaquette()
    assert len(F) == n
    assert _rel(F, F_ref) < 1e-12
    assert np.max(np.abs(np.array(A) - np.array(A_ref))) < 1e-12
    assert np.array_equal(got, ref)  # bit-identical links
    assert abs(P - oracle.su2_plaquette(links, dims)) < 1e-12


def test_su2_reflection_conserves_functional(mag, oracle):
    dims = SU2_DIMS
    links = oracle.gen_su2_config(dims, 0.3, 5)
    F0 = oracle.su2_functional(links, dims)
    with mag.MagJob(group=mag.SU2, dims=dims, omega=2.0, tol=0.0, max_sweeps=50) as job:
        job.upload(links)
        job.sweeps(50)
        F, _, _ = job.history()
    assert np.max(np.abs(np.array(F) - F0)) < 1e-12


def test_su2_fp64_converges_like_oracle(mag, oracle):
    """8^4 SU(2), eps 0.3, omega 1.7, stop at |dF/F| < 1e-12 or 1000 sweeps."""
    dims = SU2_DIMS
    links = oracle.gen_su2_config(dims, 0.3, 42)
    ref = links.copy()
    F_prev = oracle.su2_functional(ref, dims)
    F_ref = []
    for _ in range(1000):
        F, _, _ = oracle.su2_link_sweep(ref, dims, 1.7)
        F_ref.append(F)
        if abs(F - F_prev) < 1e-12 * abs(F):
            break
        F_prev = F
    with mag.MagJob(group=mag.SU2, dims=dims, omega=1.7, tol=1e-12, max_sweeps=1000) as job:
        job.upload(links)
        assert job.run()
        sweeps, conv, Ffin = job.status()
        F, _, _ = job.history()
        got = job.download()
    assert conv
    m = min(len(F), len(F_ref))
    assert abs(len(F) - len(F_ref)) <= 1
    assert _rel(F[:m], F_ref[:m]) < 1e-12
    if len(F) == len(F_ref):
        assert np.array_equal(got, ref)
    for R, T in [(1, 1), (2, 2)]:
        w0 = oracle.su2_wilson_loop(links, dims, R, T)
        w1 = oracle.su2_wilson_loop(got, dims, R, T)
        assert abs(w1 - w0) < 1e-12


def test_su3_fp64_converges_like_oracle(mag, oracle):
    """SU(3) 8^3x16, eps 0.4, omega 1.6, tol 1e-11 run to convergence: the
    same number of sweeps as the oracle's run of the same sweeps (bit-
    identical trajectories), the same final links."""
    dims = (8, 8, 8, 16)
    links = oracle.gen_su3_config(dims, 0.4, 2024)
    ref = links.copy()
    F_prev = oracle.su3_functional(ref, dims)
    n_ref = 0
    for _ in range(20000):
        F, _, _ = oracle.su3_link_sweep(ref, dims, 1.6)
        n_ref += 1
        if abs(F - F_prev) < 1e-11 * abs(F):
            break
        F_prev = F
    with mag.MagJob(group=mag.SU3, dims=dims, omega=1.6, tol=1e-11, max_sweeps=20000) as job:
        job.upload(links)
        assert job.run()
        sweeps, conv, Ffin = job.status()
        got = job.download()
    print(f"[su3 8^3x16 to tol] oracle {n_ref} sweeps, GPU {sweeps}", flush=True)
    assert conv and sweeps == n_ref
    assert np.array_equal(got, ref)
    assert abs(Ffin - F) < 1e-12 * abs(F)


@pytest.mark.parametrize("omega", [1.0, 1.6])
def test_su3_fp64_sweeps_match_oracle(mag, oracle, omega):
    dims = SU3_DIMS
    links = oracle.gen_su3_config(dims, 0.4, 4321)
    n = 25
    ref = links.copy()
    F_ref, A_ref = [], []
    for _ in range(n):
        F, mx, av = oracle.su3_link_sweep(ref, dims, omega)
        F_ref.append(F)
        A_ref.append(av)
    with mag.MagJob(group=mag.SU3, dims=dims, omega=omega, tol=0.0, max_sweeps=n) as job:
        job.upload(links)
        assert abs(job.initial_functional() - oracle.su3_functional(links, dims)) < 1e-14
        job.sweeps(n)
        F, R, A = job.history()
        got = job.download()
        P = job.plaquette()
    assert _rel(F, F_ref) < 1e-12
    assert np.max(np.abs(np.array(A) - np.array(A_ref))) < 1e-11
    assert np.array_equal(got, ref)  # bit-identical links
    assert abs(P - oracle.su3_plaquette(links, dims)) < 1e-12


# Full-size lattices of the measured workloads, plain FP64, sweep by sweep
# against the oracle: SU(2) 32^4 (eps 0.5, omega 1.7) and SU(3) 24^3x48
# (eps 0.4, omega 1.6); the gauge-invariant 1x1 and 4x4 Wilson loops of the
# swept links equal those of the input.
# 40^4: links past the 256 MB MALL (327 MB), the streaming kernel variant
@pytest.mark.parametrize("group,dims,eps,omega,n", [
    (2, (32, 32, 32, 32), 0.5, 1.7, 30),
    (3, (24, 24, 24, 48), 0.4, 1.6, 20),
    (2, (40, 40, 40, 40), 0.5, 1.7, 6),
])
def test_full_size_fp64_sweeps_match_oracle(mag, oracle, group, dims, eps, omega, n):
    print(f"[full-size parity] SU({group}) {dims} omega {omega}, {n} sweeps", flush=True)
    gen = oracle.gen_su2_config if group == 2 else oracle.gen_su3_config
    sweep = oracle.su2_link_sweep if group == 2 else oracle.su3_link_sweep
    wloop = oracle.su2_wilson_loop if group == 2 else oracle.su3_wilson_loop
    links = gen(dims, eps, 2024)
    ref = links.copy()
    F_ref = [sweep(ref, dims, omega)[0] for _ in range(n)]
    with mag.MagJob(group=group, dims=dims, omega=omega, tol=0.0, max_sweeps=n) as job:
        job.generate(eps, 2024)
        assert np.max(np.abs(job.download() - links)) < 1e-13  # the same input
        job.upload(links)
        job.sweeps(n)
        F, _, _ = job.history()
        got = job.download()
    dU = np.max(np.abs(got - ref))
    print(f"[full-size parity] max rel dF {_rel(F, F_ref):.2e}, max |dU| {dU:.2e}", flush=True)
    assert len(F) == n
    assert _rel(F, F_ref) < 1e-12
    assert dU == 0.0  # bit-identical
    for R, T in [(1, 1), (4, 4)]:
        w0 = wloop(links, dims, R, T)
        assert abs(wloop(got, dims, R, T) - w0) < 1e-12 * max(1.0, abs(w0))


def test_su2_mixed_functional_close_to_fp64(mag, oracle):
    """Mixed: FP32 sweeps + FP64 master rebuild + FP64 polish; the final
    functional within 1e-10 relative of the FP64 oracle, gauge invariants
    to 1e-12 (links are rebuilt from the FP64 master)."""
    dims = SU2_DIMS
    links = oracle.gen_su2_config(dims, 0.3, 42)
    ref = links.copy()
    F_prev = oracle.su2_functional(ref, dims)
    for _ in range(1000):
        F, _, _ = oracle.su2_link_sweep(ref, dims, 1.7)
        if abs(F - F_prev) < 1e-12 * abs(F):
            break
        F_prev = F
    with mag.MagJob(group=mag.SU2, dims=dims, precision=mag.MIXED, omega=1.7, tol=1e-12,
                    max_sweeps=1000, reunit_every=10, tol_mixed=1e-8) as job:
        job.upload(links)
        assert job.run()
        _, conv, Fm = job.status()
        got = job.download()
        start = job.fp64_start()
    assert conv and start > 0
    assert abs(Fm - F) / abs(F) < 1e-10
    assert abs(oracle.su2_functional(got, dims) - F) / abs(F) < 1e-10
    for R, T in [(1, 1), (2, 2)]:
        assert abs(oracle.su2_wilson_loop(got, dims, R, T) -
                   oracle.su2_wilson_loop(links, dims, R, T)) < 1e-12


def test_su3_mixed_functional_close_to_fp64(mag, oracle):
    dims = SU3_DIMS
    links = oracle.gen_su3_config(dims, 0.4, 99)
    ref = links.copy()
    F_prev = oracle.su3_functional(ref, dims)
    for _ in range(2000):
        F, _, _ = oracle.su3_link_sweep(ref, dims, 1.6)
        if abs(F - F_prev) < 1e-11 * abs(F):
            break
        F_prev = F
    with mag.MagJob(group=mag.SU3, dims=dims, precision=mag.MIXED, omega=1.6, tol=1e-11,
                    max_sweeps=2000, reunit_every=10, tol_mixed=1e-8) as job:
        job.upload(links)
        assert job.run()
        _, conv, Fm = job.status()
        got = job.download()
    assert conv
    assert abs(Fm - F) / abs(F) < 1e-10
    assert abs(oracle.su3_plaquette(got, dims) - oracle.su3_plaquette(links, dims)) < 1e-12


@pytest.mark.parametrize("group,dims,eps,omega", [
    (2, (48, 48, 48, 48), 0.5, 1.7),
    (3, (40, 40, 40, 24), 0.4, 1.6),
])
def test_streaming_mixed_sweeps_stay_gauge_equivalent(mag, group, dims, eps, omega):
    """Mixed link form on lattices whose FP32 links exceed the 256 MB MALL
    (SU(2) 48^4 340 MB, SU(3) 40^3x24 295 MB), where the half-sweeps load
    and store the links with the nontemporal hint: after two periods of FP32
    sweeps with FP64 rebuilds the final FP64 links are gauge-equivalent to
    the input (plaquette to 1e-12) and the functional has gone up."""
    print(f"[streaming mixed] SU({group}) {dims}", flush=True)
    with mag.MagJob(group=group, dims=dims, precision=mag.MIXED, omega=omega, tol=0.0,
                    max_sweeps=40, reunit_every=10, tol_mixed=0.0) as job:
        job.generate(eps, 31)
        P0 = job.plaquette()
        F0 = job.functional()
        job.sweeps(20)
        job.finalize()
        P1 = job.plaquette()
        F1 = job.functional()
        s, _, _ = job.status()
    assert s == 20
    assert abs(P1 - P0) < 1e-12
    assert F1 > F0


def test_one_shot_functionals(mag, oracle):
    dims = (4, 4, 4, 4)
    l2 = oracle.gen_su2_config(dims, 0.5, 5)
    l3 = oracle.gen_su3_config(dims, 0.4, 6)
    assert abs(mag.su2_functional(dims, l2) - oracle.su2_functional(l2, dims)) < 1e-14
    assert abs(mag.su3_functional(dims, l3) - oracle.su3_functional(l3, dims)) < 1e-14
    assert abs(mag.plaquette(2, dims, l2) - oracle.su2_plaquette(l2, dims)) < 1e-13
    assert abs(mag.plaquette(3, dims, l3) - oracle.su3_plaquette(l3, dims)) < 1e-13


def test_errors_are_reported(mag):
    with pytest.raises(mag.MagError) as e:
        mag.MagJob(group=mag.SU2, dims=(7, 8, 8, 8))
    assert e.value.code == -2
    with mag.MagJob(group=mag.SU2, dims=(4, 4, 4, 4)) as job:
        bad = np.zeros((1024, 4))
        bad[3, 1] = np.nan
        with pytest.raises(mag.MagError) as e:
            job.upload(bad)
        assert e.value.code == -4
        with pytest.raises(mag.MagError) as e:
            job.sweeps(1)
        assert e.value.code == -7
