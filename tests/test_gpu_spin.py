"""GPU parity of the spin-form path (the reference's mag.h call surface)
against the oracle, on the same checkerboard order.

Reflection sweeps (overrelaxation_update) conserve the spin functional and
do not contract, so rounding differences between the GPU and the oracle
grow sweep by sweep; trajectories with reflection sweeps are compared over
a few sweeps, and longer runs through invariants and the drivers'
criteria.  Relaxation sweeps contract and are compared over many.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 4, 8)


@pytest.fixture(scope="module")
def cfg(oracle):
    links = oracle.gen_su2_config(DIMS, 0.4, 314)
    spins = oracle.gen_spins_uniform(DIMS, 271)
    return links, spins


def test_generate_spins_matches_oracle(mag, oracle):
    got = mag.generate_spins_uniform(DIMS, 99)
    ref = oracle.gen_spins_uniform(DIMS, 99)
    assert np.max(np.abs(got - ref)) < 1e-15


def test_functionals_and_transform(mag, oracle, cfg):
    links, spins = cfg
    Fs = mag.functional_spin(DIMS, links, spins)
    assert abs(Fs - oracle.su2_functional_spin(links, spins, DIMS)) < 1e-14
    Ut = mag.gauge_transform_spins(DIMS, links, spins)
    ref = oracle.su2_gauge_transform_spins(links, spins, DIMS)
    assert np.max(np.abs(Ut - ref)) < 1e-13
    assert abs(mag.su2_functional(DIMS, Ut) - Fs) < 1e-13


def test_spin_at_pole_has_no_gauge_matrix(mag, oracle, cfg):
    links, spins = cfg
    s = spins.copy()
    s[5] = [0.0, 0.0, 1.0]
    with pytest.raises(mag.MagError) as e:
        mag.gauge_transform_spins(DIMS, links, s)
    assert e.value.code == -4


# relaxation and the geodesic omega rule (which puts each spin back on the
# sphere) over 20 / 40 sweeps at 1e-12; reflection conserves F and does not
# contract, so rounding differences grow: 3 sweeps
@pytest.mark.parametrize("mode,param,n,tol", [(1, 0.0, 20, 1e-12), (2, 1.6, 40, 1e-12),
                                              (0, 0.0, 3, 1e-12)])
def test_checkerboard_sweeps_match_oracle(mag, oracle, cfg, mode, param, n, tol):
    links, spins = cfg
    ref = spins.copy()
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        for _ in range(n):
            mx, av = s.sweep(mode, param)
            m0, s0 = oracle.su2_spin_halfsweep(links, ref, DIMS, 0, mode, param)
            m1, s1 = oracle.su2_spin_halfsweep(links, ref, DIMS, 1, mode, param)
            assert abs(mx - max(m0, m1)) < tol
            assert abs(av - (s0 + s1) / np.prod(DIMS)) < tol
        got = s.get_spins()
    assert np.max(np.abs(got - ref)) < tol


def test_heat_bath_sweep_matches_oracle(mag, oracle, cfg):
    links, spins = cfg
    ref = spins.copy()
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        for k in range(3):
            s.sweep(3, 1.3, seed=555)
            oracle.su2_heatbath_halfsweep(links, ref, DIMS, 0, 1.3, 555, k)
            oracle.su2_heatbath_halfsweep(links, ref, DIMS, 1, 1.3, 555, k)
        got = s.get_spins()
    assert np.max(np.abs(got - ref)) < 1e-11
    assert np.max(np.abs(np.linalg.norm(got, axis=1) - 1)) < 1e-13


def _oracle_final(oracle, links, spins, or_steps, tol_max, tol_avg, max_iter):
    """make_maximization_final (mag.cpp:1800-1816) composed from the
    oracle's checkerboard half-sweeps: the GPU driver's order."""
    V = np.prod(DIMS)
    it = 0
    while True:
        m0, s0 = oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 1)
        m1, s1 = oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 1)
        conv = max(m0, m1) < tol_max and (s0 + s1) / V < tol_avg
        for _ in range(or_steps):
            oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 0)
            oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 0)
        it += 1
        if conv or it >= max_iter:
            return it, conv


def test_maximize_final_relaxation_only_matches_oracle(mag, oracle, cfg):
    links, spins = cfg
    ref = spins.copy()
    it_ref, conv = _oracle_final(oracle, links, ref, 0, 1e-8, 1e-12, 5000)
    assert conv
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        it = s.maximize_final(0, 1e-8, 1e-12, 5000)
        got = s.get_spins()
    assert it == it_ref
    assert np.max(np.abs(got - ref)) < 1e-10


def test_maximize_final_with_reflections(mag, oracle, cfg):
    links, spins = cfg
    ref = spins.copy()
    _oracle_final(oracle, links, ref, 2, 0.0, 0.0, 2)  # two iterations
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        with pytest.raises(mag.NotConverged):
            s.maximize_final(2, 0.0, 0.0, 2)
        got = s.get_spins()
        assert np.max(np.abs(got - ref)) < 1e-10
        # to convergence: the criterion holds afterwards, F_s went up
        F0 = oracle.su2_functional_spin(links, spins, DIMS)
        it = s.maximize_final(6, 1e-10, 1e-14, 20000)
        assert it > 0
        F1 = s.functional_spin()
        assert F1 > F0
        mx, av = s.sweep(1)
        assert mx < 1e-8


def _oracle_approximate(oracle, links, spins, or_steps, digits, max_iter):
    """make_maximization_approximate (mag.cpp:1733-1750) composed from the
    oracle's checkerboard half-sweeps and its criterion (or_approx_same):
    the GPU driver's order."""
    F_old = oracle.su2_functional_spin(links, spins, DIMS)
    it = 0
    while True:
        oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 1)
        oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 1)
        for _ in range(or_steps):
            oracle.su2_spin_halfsweep(links, spins, DIMS, 0, 0)
            oracle.su2_spin_halfsweep(links, spins, DIMS, 1, 0)
        F = oracle.su2_functional_spin(links, spins, DIMS)
        it += 1
        same = oracle.approx_same(F, F_old, digits)
        F_old = F
        if same or it >= max_iter:
            return it, same, F


@pytest.mark.parametrize("or_steps,digits", [(4, 4), (1, 6), (0, 9), (0, 11)])
def test_maximize_approximate_matches_oracle(mag, oracle, cfg, or_steps, digits):
    """The GPU approximate driver stops at the oracle's iteration with the
    oracle's spins; 11 digits exercise the inexact single-precision
    powf(10, 11) of the reference's criterion.  Reflection sweeps do not
    contract rounding differences, so runs with them stop at few digits;
    relaxation-only runs go to 9 and 11."""
    links, spins = cfg
    ref = spins.copy()
    it_ref, same, F_ref = _oracle_approximate(oracle, links, ref, or_steps, digits, 20000)
    assert same and it_ref > 1
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        it = s.maximize_approximate(or_steps, digits, 20000)
        got = s.get_spins()
        F = s.functional_spin()
    assert it == it_ref
    assert np.max(np.abs(got - ref)) < 1e-10
    assert abs(F - F_ref) < 1e-12 * abs(F_ref)


def test_annealing_then_fixing_end_to_end(mag, oracle, cfg):
    """mag_fixation's new-trial path (apps/mag_su2/mag_fixation.cpp:165-204):
    uniform spins -> simulated annealing -> final maximisation -> gauge
    transformation; the fixed links' functional equals the spin functional
    and the plaquette is unchanged."""
    links, _ = cfg
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.generate_spins(7)
        s.simulated_annealing(2.5, 0.5, 0.5, 2, 3, 11)
        s.maximize_final(4, 1e-9, 1e-13, 20000)
        Fs = s.functional_spin()
        s.gauge_transform()
        fixed = s.download_links()
        F = s.functional()
    assert abs(F - Fs) < 1e-12
    assert abs(oracle.su2_functional(fixed, DIMS) - F) < 1e-13
    assert abs(oracle.su2_plaquette(fixed, DIMS) - oracle.su2_plaquette(links, DIMS)) < 1e-12
    assert F > 0.5


def test_annealing_schedule_matches_oracle_short(mag, oracle, cfg):
    links, spins = cfg
    ref = spins.copy()
    # therm 1 at T=2.0, then T = 2.0, 1.875, 1.75 (T_final/2 = 1.7), OR_steps 1
    k = 0
    oracle.su2_heatbath_halfsweep(links, ref, DIMS, 0, 2.0, 9, k)
    oracle.su2_heatbath_halfsweep(links, ref, DIMS, 1, 2.0, 9, k)
    k += 1
    T = 2.0
    while T > 3.4 / 2:
        oracle.su2_heatbath_halfsweep(links, ref, DIMS, 0, T, 9, k)
        oracle.su2_heatbath_halfsweep(links, ref, DIMS, 1, T, 9, k)
        k += 1
        oracle.su2_spin_halfsweep(links, ref, DIMS, 0, 0)
        oracle.su2_spin_halfsweep(links, ref, DIMS, 1, 0)
        T -= 0.5 / 4 if (T <= 1.6 + 0.5 and T >= 1.2) else 0.5
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(spins)
        s.simulated_annealing(2.0, 3.4, 0.5, 1, 1, 9)
        got = s.get_spins()
    assert np.max(np.abs(got - ref)) < 1e-10
