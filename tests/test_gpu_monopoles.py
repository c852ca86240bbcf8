"""Abelian projection and monopole currents on the device against the
oracle's restatement of MatrixConverter (matrix.h:1059-1100) and of
calculate_current (monopoles.cpp:44-122, 368-457, 646-687, 844-884).

From host angles the plaquette and current kernels do the reference's
additions, subtractions and divisions in its order: bit-exact.  From a
job's links the angles go through the device atan2 (1-2 ulp from glibc's):
angles to 1e-14, currents to 1e-12 and equal after rounding to integers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_job_su2(mag, oracle, job, copy, links):
    dims = job.params.dims
    ang = job.abelian_angles(copy)
    ref = oracle.su2_abelian_angles(links, dims)
    assert np.abs(ang - ref).max() < 1e-14
    J = job.monopole_currents(copy)
    Jr = oracle.monopole_currents(oracle.u1_monopole_plaquettes(ref, dims), dims)
    assert np.abs(J - Jr).max() < 1e-12
    assert np.array_equal(np.round(J), np.round(Jr))
    return np.round(J)


@pytest.mark.parametrize("eps", [0.5, 2.0])
def test_su2_job_currents(mag, oracle, eps):
    dims = (8, 8, 8, 8)
    with mag.MagJob(group=mag.SU2, dims=dims, omega=1.7, tol=1e-12, max_sweeps=400) as job:
        job.generate(eps, 2024)
        try:
            job.run()
        except mag.NotConverged:
            pass
        links = job.download()
        J = _check_job_su2(mag, oracle, job, 0, links)
    if eps == 2.0:
        assert np.count_nonzero(J) > 0  # rough links carry monopoles


def test_su2_spin_copy_currents(mag, oracle):
    dims = (8, 8, 4, 8)
    with mag.MagJob(group=mag.SU2, dims=dims, form=mag.FORM_SPIN, copies=2, omega=1.7,
                    tol=0.0, max_sweeps=40) as job:
        job.generate(1.5, 77)
        job.sweeps(40)
        for c in range(2):
            _check_job_su2(mag, oracle, job, c, job.download_copy(c))


def test_su3_job_currents(mag, oracle):
    dims = (4, 4, 4, 8)
    with mag.MagJob(group=mag.SU3, dims=dims, omega=1.6, tol=0.0, max_sweeps=30) as job:
        job.generate(1.2, 99)
        job.sweeps(30)
        links = job.download()
        ang = job.abelian_angles()
        J = job.monopole_currents()
    ref = oracle.su3_abelian_angles(links, dims)
    assert np.abs(ang - ref).max() < 1e-14
    P = oracle.su3_monopole_plaquettes(ref, dims)
    Jr = np.stack([oracle.monopole_currents(np.ascontiguousarray(P[c]), dims) for c in range(3)])
    assert np.abs(J - Jr).max() < 1e-12
    Jn = np.round(J)
    assert np.array_equal(Jn, np.round(Jr))
    assert np.all(Jn.sum(axis=0) == 0)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 5, 3), (16, 8, 8, 12)])
def test_u1_host_angles_bit_exact(mag, oracle, dims):
    rng = np.random.default_rng(11)
    th = rng.uniform(-np.pi, np.pi, 4 * int(np.prod(dims)))
    J = mag.u1_monopole_currents(dims, th)
    Jr = oracle.monopole_currents(oracle.u1_monopole_plaquettes(th, dims), dims)
    assert np.array_equal(J, Jr)
    assert np.count_nonzero(np.round(J)) > 0


# L3 = 3 and 6: runs of t shorter than the kernel's run length, and a
# partial last run
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 5, 3), (6, 4, 5, 6)])
def test_su3_host_angles_bit_exact(mag, oracle, dims):
    rng = np.random.default_rng(12)
    n = 4 * int(np.prod(dims))
    a = rng.uniform(-np.pi, np.pi, (n, 2))
    ang = np.ascontiguousarray(np.stack([a[:, 0], a[:, 1], -(a[:, 0] + a[:, 1])], axis=1))
    J = mag.su3_monopole_currents(dims, ang)
    P = oracle.su3_monopole_plaquettes(ang, dims)
    Jr = np.stack([oracle.monopole_currents(np.ascontiguousarray(P[c]), dims) for c in range(3)])
    assert np.array_equal(J, Jr)


def test_monopole_argument_errors(mag):
    dims = (4, 4, 4, 4)
    th = np.zeros(4 * 256)
    th[17] = np.inf
    with pytest.raises(mag.MagError) as e:
        mag.u1_monopole_currents(dims, th)
    assert e.value.code == -4  # MAG_ERR_NONFINITE
    th[17] = np.nan
    with pytest.raises(mag.MagError) as e:
        mag.su3_monopole_currents(dims, np.zeros((4 * 256, 3)) + th[:, None])
    assert e.value.code == -4  # MAG_ERR_NONFINITE
    th[17] = 2e3
    with pytest.raises(mag.MagError) as e:
        mag.u1_monopole_currents(dims, th)
    assert e.value.code == -1  # MAG_ERR_ARG
    with pytest.raises(mag.MagError):
        mag.u1_monopole_currents((4, 4, 1, 4), np.zeros(4 * 64))
    with mag.MagJob(group=mag.SU2, dims=dims) as job:
        job.generate(0.5, 1)
        with pytest.raises(mag.MagError):
            job.monopole_currents(1)
