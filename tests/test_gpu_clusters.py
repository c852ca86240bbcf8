"""Monopole clusters on the device (mag_current_clusters,
mag_job_monopole_clusters) against the oracle's restatement of the
reference's walk (calculate_clusters, monopoles.cpp:1120-1303, with the
observables of 1343-1449 and 1510-1531): the records are integers, so
bit-exact; the cluster order is the reference's (by first site)."""
import ctypes

import numpy as np
import pytest

from test_oracle_clusters import add_line, add_loop, random_currents, union_find_clusters

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,sigma", [((8, 8, 8, 8), 0.6), ((8, 8, 8, 8), 0.8),
                                        ((6, 4, 8, 4), 0.9), ((2, 4, 2, 6), 1.0),
                                        ((5, 4, 3, 6), 0.8)])
def test_host_currents_match_oracle(mag, oracle, dims, sigma):
    print(f"[clusters] host J dims={dims} sigma={sigma}", flush=True)
    J = random_currents(dims, sigma, 21)
    want = oracle.monopole_clusters(J, dims)
    assert len(want) > 0
    np.testing.assert_array_equal(mag.current_clusters(dims, J), want)


def test_host_currents_hand_built(mag, oracle):
    dims = (4, 4, 6, 8)
    J = np.zeros(4 * int(np.prod(dims)))
    add_loop(J, dims, [1, 1, 1, 1], 0, 1)
    add_loop(J, dims, [2, 2, 4, 5], 2, 3, q=2)
    add_line(J, dims, [3, 0, 0, 0], 3)
    add_line(J, dims, [0, 3, 3, 3], 0, q=-1)
    add_loop(J, dims, [0, 3, 3, 3], 1, 2)  # joins the x line
    got = mag.current_clusters(dims, J)
    np.testing.assert_array_equal(got, oracle.monopole_clusters(J, dims))
    np.testing.assert_array_equal(got, union_find_clusters(J, dims))
    assert len(got) == 4  # the last loop joins the x line: 8 links on 7 sites
    assert [r[1] for r in got if r[2] == -4] == [8] and [r[8] for r in got if r[2] == -4] == [7]
    np.testing.assert_array_equal(mag.current_clusters(dims, np.zeros_like(J)), np.zeros((0, 9)))


def test_su3_host_currents_three_colours(mag, oracle):
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(5)
    a3 = rng.normal(0.0, 0.8, (4 * int(np.prod(dims)), 3))
    a3 -= a3.mean(axis=1, keepdims=True)
    J3 = mag.su3_monopole_currents(dims, a3)
    got = mag.current_clusters(dims, J3)
    for k in range(3):
        np.testing.assert_array_equal(got[k], oracle.monopole_clusters(np.ascontiguousarray(J3[k]), dims))


def test_job_clusters_su2(mag, oracle):
    dims = (8, 8, 8, 8)
    with mag.MagJob(group=mag.SU2, dims=dims, omega=1.7, tol=0.0, max_sweeps=200) as job:
        job.generate(1.0, 2024)
        job.sweeps(200)
        J = job.monopole_currents()
        rec = job.monopole_clusters()
    assert len(rec) > 0
    np.testing.assert_array_equal(rec, oracle.monopole_clusters(J, dims))


def test_job_clusters_spin_copies(mag, oracle):
    dims = (8, 8, 4, 8)
    with mag.MagJob(group=mag.SU2, dims=dims, form=mag.FORM_SPIN, copies=2, omega=1.7,
                    tol=0.0, max_sweeps=40) as job:
        job.generate(1.5, 77)
        job.sweeps(40)
        for c in range(2):
            np.testing.assert_array_equal(job.monopole_clusters(c),
                                          oracle.monopole_clusters(job.monopole_currents(c), dims))


def test_job_clusters_su3(mag, oracle):
    dims = (4, 4, 4, 8)
    with mag.MagJob(group=mag.SU3, dims=dims, omega=1.6, tol=0.0, max_sweeps=30) as job:
        job.generate(1.2, 99)
        job.sweeps(30)
        J3 = job.monopole_currents()
        recs = job.monopole_clusters()
        ms = job.clusters_timed(3)
    assert ms > 0
    for k in range(3):
        np.testing.assert_array_equal(recs[k], oracle.monopole_clusters(np.ascontiguousarray(J3[k]), dims))


def test_capacity_and_argument_errors(mag, oracle):
    dims = (6, 6, 6, 6)
    V = int(np.prod(dims))
    J = random_currents(dims, 0.6, 4)
    want = oracle.monopole_clusters(J, dims)
    assert len(want) > 2
    L = mag.lib()
    d = (ctypes.c_int * 4)(*dims)
    rec = np.zeros((2, 9), dtype=np.int32)
    n = np.zeros(1, dtype=np.int64)
    rc = L.mag_current_clusters(d, 1, J.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                rec.ctypes.data, 2, n.ctypes.data)
    assert rc == -1 and n[0] == len(want)
    np.testing.assert_array_equal(rec, want[:2])
    bad = np.zeros(4 * V)
    bad[4 * 7 + 2] = 1.0  # a lone link: charges do not balance
    with pytest.raises(mag.MagError) as e:
        mag.current_clusters(dims, bad)
    assert e.value.code == -1
    half = np.zeros(4 * V)
    add_loop(half, dims, [1, 1, 1, 1], 0, 1, q=0.4)  # currents of charge 0
    with pytest.raises(mag.MagError) as e:
        mag.current_clusters(dims, half)
    assert e.value.code == -1
    nan = J.copy()
    nan[5] = np.nan
    with pytest.raises(mag.MagError) as e:
        mag.current_clusters(dims, nan)
    assert e.value.code == -4
