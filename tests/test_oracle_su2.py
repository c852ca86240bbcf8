"""Pin the CPU oracle's SU(2) part against closed forms, invariants and an
independent brute-force formulation (the reference ships no golden vectors
for this path: ilyakudrov/general_code tests print numbers without
assertions, tests/cpu/code/test_mag.cpp:159-222).

Each check is chosen so that a plausible restatement mistake (a wrong sign
in a product, a transposed adjoint matrix, a dropped neighbour term, the
wrong parity order) fails it.
"""
import numpy as np
import pytest

S1 = np.array([[0, 1], [1, 0]], complex)
S2 = np.array([[0, -1j], [1j, 0]])
S3 = np.array([[1, 0], [0, -1]], complex)
SIG = [S1, S2, S3]


def mat(q):
    """a0 + i a.sigma (lib/cpu/include/matrix.h:8-13) as a 2x2 matrix."""
    return q[0] * np.eye(2) + 1j * (q[1] * S1 + q[2] * S2 + q[3] * S3)


def quat(M):
    a0 = 0.5 * np.trace(M).real
    a = [0.5 * np.trace(-1j * s @ M).real for s in SIG]
    return np.array([a0, *a])


def rand_su2(rng, n=None):
    q = rng.normal(size=(n, 4) if n else 4)
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def adjoint(q):
    A = mat(q)
    return np.array([[0.5 * np.trace(SIG[i] @ A @ SIG[j] @ A.conj().T).real
                      for j in range(3)] for i in range(3)])


def test_su2_products_match_2x2_matrices(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b = rand_su2(rng), rand_su2(rng)
        A, B = mat(a), mat(b)
        assert np.allclose(mat(oracle.su2_mul(a, b)), A @ B, atol=1e-15)
        assert np.allclose(mat(oracle.su2_mul_dag(a, b)), A @ B.conj().T, atol=1e-15)
        assert np.allclose(mat(oracle.su2_dag_mul(a, b)), A.conj().T @ B, atol=1e-15)
        assert abs(np.linalg.norm(oracle.su2_mul(a, b)) - 1) < 1e-15


def test_link_functional_closed_form(oracle):
    rng = np.random.default_rng(2)
    for q in rand_su2(rng, 100):
        F = oracle.su2_link_functional(q)
        M = mat(q)
        assert abs(F - (1 - 2 * (q[1] ** 2 + q[2] ** 2))) < 1e-15
        assert abs(F - (abs(M[0, 0]) ** 2 - abs(M[0, 1]) ** 2)) < 1e-15
    assert oracle.su2_link_functional(np.array([1.0, 0, 0, 0])) == 1.0


def test_contribution_is_adjoint_matrix(oracle):
    rng = np.random.default_rng(3)
    for q in rand_su2(rng, 20):
        R = adjoint(q)
        b = rng.normal(size=3)
        h0 = rng.normal(size=3)
        assert np.allclose(oracle.contribution1(h0, q, b), h0 + R @ b, atol=1e-14)
        assert np.allclose(oracle.contribution1_conj(h0, q, b), h0 + R.T @ b, atol=1e-14)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14


def test_reflect_and_parallel(oracle):
    rng = np.random.default_rng(4)
    for _ in range(50):
        s = rng.normal(size=3)
        s /= np.linalg.norm(s)
        V = rng.normal(size=3) * 3
        v = V / np.linalg.norm(V)
        r = oracle.reflect_fast(s, V)
        assert np.allclose(r, 2 * np.dot(s, v) * v - s, atol=1e-14)
        assert np.allclose(oracle.reflect_fast(r, V), s, atol=1e-14)
        assert abs(np.dot(r, V) - np.dot(s, V)) < 1e-13
        p, d = oracle.parallel(s, V)
        assert np.allclose(p, v, atol=1e-15)
        assert abs(d - (1 - np.dot(s, v))) < 1e-15


def test_gauge_matrix_spin_roundtrip(oracle):
    """row3 of the rotation of G = GetGaugeMatrix(s) is s (matrix.h:948-963),
    and G^+ s3 G = s.sigma for the reference's convention."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        s = rng.normal(size=3)
        s /= np.linalg.norm(s)
        g = oracle.gauge_matrix(s)
        assert abs(np.linalg.norm(g) - 1) < 1e-14
        a0, a1, a2, a3 = g
        row3 = np.array([2 * (a1 * a3 + a0 * a2), 2 * (a2 * a3 - a0 * a1),
                         a0 ** 2 + a3 ** 2 - a1 ** 2 - a2 ** 2])
        assert np.allclose(row3, s, atol=1e-13)
        G = mat(g)
        X = G.conj().T @ S3 @ G
        assert np.allclose(X, s[0] * S1 + s[1] * S2 + s[2] * S3, atol=1e-13)
    assert np.isnan(oracle.gauge_matrix(np.array([0.0, 0.0, 1.0]))).any()


def _lattice(oracle, dims, eps, seed):
    return oracle.gen_su2_config(dims, eps, seed)


def _contribution_site_bruteforce(links, spins, dims, site):
    """h(x) = sum_mu M(U_x,mu) s(x+mu) + M(U_x-mu,mu)^T s(x-mu) from 2x2
    matrices (independent of the oracle's closed-form expressions)."""
    L = dims
    x = np.unravel_index(site, L[::-1])[::-1]  # (x, y, z, t)
    h = np.zeros(3)
    for mu in range(4):
        xf = list(x)
        xf[mu] = (xf[mu] + 1) % L[mu]
        xb = list(x)
        xb[mu] = (xb[mu] - 1) % L[mu]
        f = np.ravel_multi_index(xf[::-1], L[::-1])
        b = np.ravel_multi_index(xb[::-1], L[::-1])
        h += adjoint(links[4 * site + mu]) @ spins[f]
        h += adjoint(links[4 * b + mu]).T @ spins[b]
    return h


def test_contribution_site_bruteforce(oracle):
    dims = (2, 4, 2, 4)
    links = _lattice(oracle, dims, 0.7, 11)
    spins = oracle.gen_spins_uniform(dims, 12)
    for site in range(int(np.prod(dims))):
        h = oracle.contribution_site(links, spins, dims, site)
        assert np.allclose(h, _contribution_site_bruteforce(links, spins, dims, site), atol=1e-13)


def test_functional_spin_equals_link_functional_after_transform(oracle):
    """The reference author's own check (tests/cpu/code/test_mag.cpp:206-222):
    F_spin(U, s) = F(gauge_tranformation_spins(U, s))."""
    dims = (4, 4, 4, 4)
    links = _lattice(oracle, dims, 0.5, 13)
    spins = oracle.gen_spins_uniform(dims, 14)
    Fs = oracle.su2_functional_spin(links, spins, dims)
    Ut = oracle.su2_gauge_transform_spins(links, spins, dims)
    assert abs(Fs - oracle.su2_functional(Ut, dims)) < 1e-13
    assert abs(oracle.su2_plaquette(Ut, dims) - oracle.su2_plaquette(links, dims)) < 1e-13
    # identity configuration and aligned spins: F = 1
    ident = np.tile([1.0, 0, 0, 0], (4 * 256, 1))
    assert oracle.su2_functional(ident, dims) == 1.0


def test_lexicographic_reflection_sweep_conserves_functional(oracle):
    dims = (4, 4, 4, 4)
    links = _lattice(oracle, dims, 0.5, 15)
    spins = oracle.gen_spins_uniform(dims, 16)
    F0 = oracle.su2_functional_spin(links, spins, dims)
    for _ in range(5):
        oracle.su2_or_sweep_lexi(links, spins, dims)
        assert abs(oracle.su2_functional_spin(links, spins, dims) - F0) < 1e-14


def test_relaxation_is_monotone_and_maximisation_converges(oracle):
    dims = (4, 4, 4, 4)
    links = _lattice(oracle, dims, 0.3, 17)
    spins = oracle.gen_spins_uniform(dims, 18)
    F = oracle.su2_functional_spin(links, spins, dims)
    for _ in range(10):
        oracle.su2_relax_sweep_lexi(links, spins, dims)
        F1 = oracle.su2_functional_spin(links, spins, dims)
        assert F1 >= F - 1e-15
        F = F1
    it = oracle.su2_maximize_final(links, spins, dims, 4, 1e-10, 1e-14, 10000)
    assert it > 0
    dmax, davg = oracle.su2_relax_sweep_lexi(links, spins, dims)
    assert dmax < 1e-9
    Fa = oracle.su2_functional_spin(links, spins, dims)
    assert Fa >= F - 1e-15
    # approximate criterion terminates too
    spins2 = oracle.gen_spins_uniform(dims, 19)
    assert oracle.su2_maximize_approximate(links, spins2, dims, 4, 7, 10000) > 0


def test_approximate_criterion_hand_cases(oracle):
    """The stopping test of make_maximization_approximate (mag.cpp:1744-1746):
    trunc(powf(10, d) * F) == trunc(powf(10, d) * F_old) and equal sign
    bits.  powf is single precision: 10^d is exact in float up to d = 10,
    and powf(10, 11) = 99999997952 (the nearest float to 1e11).  The cases
    below were chosen by hand so that pow in place of powf, floor in place
    of trunc, or a dropped sign test each flips one of them."""
    # d = 11, F_old = 0.3, F = 0.3 + 4.5e-12:
    #   with 1e11: 30000000000 vs 30000000000.45 -> same
    #   with powf: 29999999385.6 vs 29999999386.05 -> 29999999385 != 29999999386
    assert oracle.approx_same(0.3000000000045, 0.3, 11) is False
    # F = 0.3 - 3e-12: 1e11 gives 29999999999 != 30000000000, powf gives
    # 29999999385.3 and 29999999385.6 -> same
    assert oracle.approx_same(0.3 - 3e-12, 0.3, 11) is True
    # d = 10 (powf exact): the decimal digits decide
    assert oracle.approx_same(0.3000000000045, 0.3, 10) is True
    assert oracle.approx_same(0.3 - 3e-12, 0.3, 10) is False
    # trunc, not floor: -1.0 vs -0.5 -> -1 != -0 (floor: -1 == -1)
    assert oracle.approx_same(-0.001, -0.0005, 3) is False
    # both truncate to zero but the signs differ
    assert oracle.approx_same(-1e-13, 1e-13, 3) is False
    assert oracle.approx_same(1e-13, 2e-13, 3) is True
    # ordinary cases at 7 digits
    assert oracle.approx_same(0.61234567, 0.61234568, 7) is True
    assert oracle.approx_same(0.61234567, 0.61234557, 7) is False


def test_maximize_approximate_stops_at_the_first_equal_digits(oracle):
    """or_su2_maximize_approximate is the loop {relax, OR x n, F_s} until
    the criterion holds: replayed step by step with the oracle's own
    sweeps, it stops at the first iteration whose F_s matches the previous
    one and returns that iteration count."""
    dims = (4, 4, 4, 4)
    links = oracle.gen_su2_config(dims, 0.5, 23)
    spins = oracle.gen_spins_uniform(dims, 23)
    ref = spins.copy()
    it = oracle.su2_maximize_approximate(links, spins, dims, 3, 6, 10000)
    F_old = oracle.su2_functional_spin(links, ref, dims)
    k = 0
    while True:
        oracle.su2_relax_sweep_lexi(links, ref, dims)
        for _ in range(3):
            oracle.su2_or_sweep_lexi(links, ref, dims)
        F = oracle.su2_functional_spin(links, ref, dims)
        k += 1
        # criterion by hand in decimal (powf(10, 6) is exact)
        same = int(F * 1e6) == int(F_old * 1e6) and (F < 0) == (F_old < 0)
        F_old = F
        if same:
            break
    assert it == k > 1
    assert np.array_equal(spins, ref)


def test_omega_rule_angle_functions_against_libm(oracle):
    """The omega rule's atan2 (y > 0) and sin / cos on [0, 2 pi] are a fixed
    sequence of IEEE operations shared with the device; they agree with the
    C library's atan2 / sin / cos to 2 ulp, at random points and at the ends
    and switch points of their argument reductions."""
    rng = np.random.default_rng(77)
    ys = np.concatenate([rng.random(3000), 10.0 ** rng.uniform(-12, 3, 2000)])
    xs = np.concatenate([rng.normal(size=3000), rng.choice([-1, 1], 2000) * 10.0 ** rng.uniform(-12, 3, 2000)])
    # t = y/|x| on the quarter switch points and the octant boundary
    edge_t = np.array([0.125, 0.375, 0.625, 0.875, 1.0, 1e-300, 0.5])
    ys = np.concatenate([ys, edge_t, edge_t, np.ones(7)])
    xs = np.concatenate([xs, np.ones(7), -np.ones(7), 1 / edge_t[:7]])
    for y, x in zip(ys, xs):
        ref = np.arctan2(y, x)
        got = oracle.atan2_pos(y, x)
        assert abs(got - ref) <= 2 * np.spacing(ref), (y, x, got, ref)
    pio4 = np.pi / 4
    angles = np.concatenate([rng.random(5000) * 2 * np.pi, [0.0, 2 * np.pi, np.pi],
                             pio4 * np.arange(1, 8), np.nextafter(pio4 * np.arange(1, 8), 0)])
    for a in angles:
        s, c = oracle.sincos(a)
        assert abs(s - np.sin(a)) <= 2 * np.spacing(abs(np.sin(a))) + 1e-300, (a, s)
        assert abs(c - np.cos(a)) <= 2 * np.spacing(abs(np.cos(a))) + 1e-300, (a, c)


def test_omega_gauge_rotates_by_omega_theta(oracle):
    rng = np.random.default_rng(6)
    for _ in range(30):
        h = rng.normal(size=3)
        hh = h / np.linalg.norm(h)
        theta = np.arccos(hh[2])
        for w in (0.5, 1.0, 1.7, 2.0):
            g = oracle.omega_gauge(h, w)
            G = mat(g)
            X = G.conj().T @ S3 @ G
            n = np.array([0.5 * np.trace(s @ X).real for s in SIG])
            assert abs(g[3]) == 0.0
            assert abs(np.arccos(np.clip(n[2], -1, 1)) - min(w * theta, 2 * np.pi - w * theta)) < 1e-12
            # n stays in the plane spanned by e3 and h
            assert abs(np.dot(n, np.cross([0, 0, 1.0], hh))) < 1e-12
        g1 = oracle.omega_gauge(h, 1.0)
        G1 = mat(g1)
        X = G1.conj().T @ S3 @ G1
        n = np.array([0.5 * np.trace(s @ X).real for s in SIG])
        assert np.allclose(n, hh, atol=1e-13)


def test_block_fields_match_trace_formula(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        Bc = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        B = np.array([Bc[0, 0].real, Bc[0, 0].imag, Bc[0, 1].real, Bc[0, 1].imag,
                      Bc[1, 0].real, Bc[1, 0].imag, Bc[1, 1].real, Bc[1, 1].imag])
        xf = [0.5 * np.trace(s @ Bc @ S3 @ Bc.conj().T).real for s in SIG]
        xb = [0.5 * np.trace(s @ Bc.conj().T @ S3 @ Bc).real for s in SIG]
        assert np.allclose(oracle.block_field_fwd(B), xf, atol=1e-13)
        assert np.allclose(oracle.block_field_bwd(B), xb, atol=1e-13)


@pytest.mark.parametrize("omega", [1.0, 1.3, 1.7, 2.0])
def test_link_sweep_equals_spin_sweep(oracle, omega):
    """Two independent formulations of the same checkerboard sweep: spins
    with read-only links (reference primitives, geodesic rule) and the
    link form the GPU runs.  F per sweep must agree."""
    dims = (8, 8, 4, 8)
    links = _lattice(oracle, dims, 0.5, 4242)
    spins = np.tile([0.0, 0.0, 1.0], (2048, 1))
    W = links.copy()
    # long enough for a spin-length instability of the omega step (rounding
    # grown by ~2 omega - 1 per update) to show: it reaches 1e-9 by sweep 20
    n = 4 if omega == 2.0 else 60
    for _ in range(n):
        oracle.su2_spin_halfsweep(links, spins, dims, 0, 2, omega)
        oracle.su2_spin_halfsweep(links, spins, dims, 1, 2, omega)
        F, _, _ = oracle.su2_link_sweep(W, dims, omega)
        assert abs(oracle.su2_functional_spin(links, spins, dims) - F) < 1e-13
        assert np.max(np.abs(np.linalg.norm(spins, axis=1) - 1)) < 1e-14


def test_spin_modes_reduce_to_reference_updates(oracle):
    """omega=1 of the geodesic rule is spin::parallel, omega=2 is
    spin::reflect_fast (to rounding: the rule goes through atan2/sincos)."""
    dims = (4, 4, 4, 4)
    links = _lattice(oracle, dims, 0.5, 23)
    s0 = oracle.gen_spins_uniform(dims, 24)
    for omega, mode in [(1.0, 1), (2.0, 0)]:
        a, b = s0.copy(), s0.copy()
        oracle.su2_spin_halfsweep(links, a, dims, 0, 2, omega)
        oracle.su2_spin_halfsweep(links, b, dims, 0, mode)
        assert np.max(np.abs(a - b)) < 1e-13


def test_link_sweep_monotone_and_gauge_invariant(oracle):
    dims = (4, 4, 4, 8)
    links = _lattice(oracle, dims, 0.5, 25)
    W = links.copy()
    F = oracle.su2_functional(W, dims)
    for _ in range(30):
        F1, rmax, ravg = oracle.su2_link_sweep(W, dims, 1.7)
        assert F1 >= F - 1e-14
        assert 0 <= ravg <= rmax <= 2
        F = F1
    assert np.max(np.abs(np.linalg.norm(W, axis=1) - 1)) < 1e-13
    assert abs(oracle.su2_plaquette(W, dims) - oracle.su2_plaquette(links, dims)) < 1e-13
    for R, T in [(1, 1), (2, 3), (4, 4)]:
        assert abs(oracle.su2_wilson_loop(W, dims, R, T) -
                   oracle.su2_wilson_loop(links, dims, R, T)) < 1e-13


def test_link_sweep_thread_independent(oracle):
    dims = (4, 4, 8, 8)
    links = _lattice(oracle, dims, 0.5, 27)
    a, b = links.copy(), links.copy()
    n0 = oracle.get_threads()
    oracle.set_threads(1)
    ra = [oracle.su2_link_sweep(a, dims, 1.7) for _ in range(3)]
    oracle.set_threads(max(n0, 2))
    rb = [oracle.su2_link_sweep(b, dims, 1.7) for _ in range(3)]
    oracle.set_threads(n0)
    assert ra == rb
    assert np.array_equal(a, b)


@pytest.mark.parametrize("copy", [0, 1, 3])
def test_copy_spins_follow_link_form_copy(oracle, copy):
    """Spin form on the links of copy 0, started at gen_copy_spins(copy),
    is gauge copy `copy` of the link form: the same F at the start (the
    invariant F_s(U, s) = F(G(s) U G(s')^+), mag.cpp:925-953 vs 993-1001)
    and after every omega sweep."""
    dims = (4, 4, 4, 6)
    U0 = oracle.gen_su2_config(dims, 0.5, 77)
    Uc = oracle.gen_su2_config(dims, 0.5, 77, copy=copy)
    spins = oracle.gen_copy_spins(dims, 77, copy)
    assert np.max(np.abs(np.linalg.norm(spins, axis=1) - 1)) < 1e-14
    if copy == 0:
        assert np.array_equal(spins, np.tile([0.0, 0.0, 1.0], (len(spins), 1)))
    assert abs(oracle.su2_functional_spin(U0, spins, dims) - oracle.su2_functional(Uc, dims)) < 1e-14
    for _ in range(10):
        oracle.su2_spin_halfsweep(U0, spins, dims, 0, 2, 1.7)
        oracle.su2_spin_halfsweep(U0, spins, dims, 1, 2, 1.7)
        F, _, _ = oracle.su2_link_sweep(Uc, dims, 1.7)
        assert abs(oracle.su2_functional_spin(U0, spins, dims) - F) < 1e-13
    # a wrong pairing (copy spins on another copy's links) does not match
    if copy:
        wrong = oracle.gen_copy_spins(dims, 77, copy + 1)
        assert abs(oracle.su2_functional_spin(U0, wrong, dims) -
                   oracle.su2_functional(Uc, dims)) > 1e-4
