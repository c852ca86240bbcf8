"""Launch shapes of the link-form half-sweep: partials are per wave, so every
(threads per workgroup, tiles per thread) gives bit-identical sweeps; the
autotuner picks one of them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(2, (8, 8, 8, 12), 0, 0.5, 1.7), (2, (8, 8, 8, 12), 1, 0.5, 1.7),
         (3, (4, 6, 4, 8), 0, 0.4, 1.6), (3, (4, 6, 4, 8), 1, 0.4, 1.6)]


def _run(mag, group, dims, prec, eps, omega, n=25):
    with mag.MagJob(group=group, dims=dims, precision=prec, omega=omega, tol=0.0,
                    max_sweeps=n, reunit_every=5, tol_mixed=0.0) as j:
        j.generate(eps, 99)
        j.sweeps(n)
        F, R, A = j.history()
        return F, R, A, j.download(), j.shape(fp32=bool(prec))


@pytest.mark.parametrize("group,dims,prec,eps,omega", CASES)
def test_every_shape_gives_the_same_sweeps(mag, group, dims, prec, eps, omega):
    try:
        mag.tune_set(mag.TUNE_DEFAULT)
        ref = _run(mag, group, dims, prec, eps, omega)
        assert ref[4] == (256, 1)
        for nt, spt in mag.SHAPES:
            mag.tune_set(mag.TUNE_FIXED, nt, spt)
            got = _run(mag, group, dims, prec, eps, omega)
            assert got[4] == (nt, spt)
            for x, y in zip(got[:4], ref[:4]):
                assert np.array_equal(x, y), (nt, spt)
        mag.tune_set(mag.TUNE_AUTO)
        mag.tune_clear()
        got = _run(mag, group, dims, prec, eps, omega)
        assert got[4] in mag.SHAPES
        for x, y in zip(got[:4], ref[:4]):
            assert np.array_equal(x, y)
    finally:
        mag.tune_set(mag.TUNE_AUTO)


def test_bad_shape_is_refused(mag):
    with pytest.raises(mag.MagError):
        mag.tune_set(mag.TUNE_FIXED, 96, 1)
    with pytest.raises(mag.MagError):
        mag.tune_set(7)
