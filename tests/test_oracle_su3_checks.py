"""Pin the oracle's SU(3) part, the gauge-invariant checks, the Philox
generator and the synthetic configurations (closed forms, library routines,
known-answer vectors, invariants)."""
import numpy as np
import pytest
import scipy.linalg


def m3(v18):
    v = np.asarray(v18).reshape(3, 3, 2)
    return v[..., 0] + 1j * v[..., 1]


def v18(M):
    return np.stack([M.real, M.imag], axis=-1).reshape(18)


def gell_mann():
    """lambda_1..8 of get_generators_su3 (matrix.h:987-1046)."""
    L = np.zeros((8, 3, 3), complex)
    L[0][0, 1] = L[0][1, 0] = 1
    L[1][0, 1], L[1][1, 0] = -1j, 1j
    L[2][0, 0], L[2][1, 1] = 1, -1
    L[3][0, 2] = L[3][2, 0] = 1
    L[4][0, 2], L[4][2, 0] = -1j, 1j
    L[5][1, 2] = L[5][2, 1] = 1
    L[6][1, 2], L[6][2, 1] = -1j, 1j
    L[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return L


# ---------------------------------------------------------------- Philox
def test_philox_known_answers(oracle):
    """Random123 philox4x32-10 known-answer vectors (kat_vectors)."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---------------------------------------------------------------- SU(3) algebra
def test_su3_exp_matches_expm(oracle):
    rng = np.random.default_rng(1)
    L = gell_mann()
    for _ in range(20):
        c = rng.normal(size=8)
        got = m3(oracle.gen_su3_exp(c, 0.4))
        H = 0.4 * np.einsum("a,aij->ij", c, L) / 2
        ref = scipy.linalg.expm(1j * H)
        assert np.max(np.abs(got - ref)) < 1e-14


def test_su3_reunitarize_and_proj(oracle):
    rng = np.random.default_rng(2)
    for _ in range(20):
        U = scipy.linalg.expm(1j * np.einsum("a,aij->ij", rng.normal(size=8), gell_mann()))
        P = U + 1e-3 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        for f in (oracle.su3_reunitarize, oracle.su3_proj):
            R = m3(f(v18(P)))
            assert np.max(np.abs(R @ R.conj().T - np.eye(3))) < 1e-13
            assert abs(np.linalg.det(R) - 1) < 1e-12
            assert np.max(np.abs(R - U)) < 1e-2
        # Gram-Schmidt keeps the direction of row 0
        R = m3(oracle.su3_reunitarize(v18(P)))
        r0 = P[0] / np.linalg.norm(P[0])
        assert np.max(np.abs(R[0] - r0)) < 1e-14


def test_su3_proj_is_the_special_unitary_polar_factor(oracle):
    """su3::proj (matrix.h:279-292): Newton steps A <- 3/2 A - 1/2 A A^+ A
    converge to the unitary polar factor W of M = W P, and the factor
    1 - (det A - 1)/3 removes the determinant's phase (to third order per
    step), so six steps give W det(W)^(-1/3) (principal root) to rounding.
    W comes from scipy.linalg.polar.  Gram-Schmidt of the same M (another
    unitary matrix, off by O(|M - U|)) fails the 1e-12 bound."""
    rng = np.random.default_rng(12)
    for amp in (1e-3, 1e-2, 3e-2):
        for _ in range(10):
            U = scipy.linalg.expm(1j * np.einsum("a,aij->ij", rng.normal(size=8), gell_mann()))
            M = U + amp * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
            W, _ = scipy.linalg.polar(M)
            d = np.linalg.det(W)
            ref = W * np.exp(-1j * np.angle(d) / 3)
            got = m3(oracle.su3_proj(v18(M)))
            assert np.max(np.abs(got - ref)) < 1e-12, amp
            gs = m3(oracle.su3_reunitarize(v18(M)))
            assert np.max(np.abs(gs - ref)) > 1e3 * np.max(np.abs(got - ref))
    # an SU(3) matrix is its own projection; a positive multiple projects to it
    U = scipy.linalg.expm(1j * np.einsum("a,aij->ij", rng.normal(size=8), gell_mann()))
    assert np.max(np.abs(m3(oracle.su3_proj(v18(U))) - U)) < 1e-14
    assert np.max(np.abs(m3(oracle.su3_proj(v18(0.98 * U))) - U)) < 1e-13


def test_su3_products(oracle):
    rng = np.random.default_rng(3)
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    B = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    assert np.allclose(m3(oracle.su3_mul(v18(A), v18(B))), A @ B, atol=1e-13)
    assert np.allclose(m3(oracle.su3_mul_dag(v18(A), v18(B))), A @ B.conj().T, atol=1e-13)


def test_su3_functional_closed_form(oracle):
    dims = (2, 2, 2, 2)
    ident = np.tile(v18(np.eye(3)), (4 * 16, 1))
    assert oracle.su3_functional(ident, dims) == 1.0
    links = oracle.gen_su3_config(dims, 0.4, 5)
    M = np.array([m3(v) for v in links])
    ref = np.sum(np.abs(np.einsum("lii->li", M)) ** 2) / (12 * 16)
    assert abs(oracle.su3_functional(links, dims) - ref) < 1e-14


# ---------------------------------------------------------------- SU(3) CM sweep
def embed_su2(q):
    a0, a1, a2, a3 = q
    M = np.eye(3, dtype=complex)
    M[:2, :2] = [[a0 + 1j * a3, a2 + 1j * a1], [-a2 + 1j * a1, a0 - 1j * a3]]
    return M


@pytest.mark.parametrize("omega", [1.0, 1.7])
def test_su3_sweep_on_embedded_su2_is_su2_sweep(oracle, omega):
    """SU(2) links embedded in the (0,1) block: subgroup (0,1) performs the
    SU(2) update and the subgroups (1,2), (0,2) see h_perp = 0, h3 > 0 (no
    update), so the SU(3) sweep reproduces the SU(2) sweep link by link with
    F3 = (F2 + 2)/3."""
    dims = (4, 4, 4, 4)
    l2 = oracle.gen_su2_config(dims, 0.5, 31)
    l3 = np.array([v18(embed_su2(q)) for q in l2])
    for _ in range(6):
        F2, _, a2 = oracle.su2_link_sweep(l2, dims, omega)
        F3, _, a3 = oracle.su3_link_sweep(l3, dims, omega)
        assert abs(F3 - (F2 + 2) / 3) < 1e-14
        assert abs(a3 - a2) < 1e-14
    emb = np.array([v18(embed_su2(q)) for q in l2])
    assert np.max(np.abs(emb - l3)) < 1e-13


def test_su3_sweep_monotone_unitary_invariant(oracle):
    dims = (4, 4, 4, 4)
    links = oracle.gen_su3_config(dims, 0.4, 33)
    W = links.copy()
    F = oracle.su3_functional(W, dims)
    for _ in range(15):
        F1, rmax, ravg = oracle.su3_link_sweep(W, dims, 1.6)
        assert F1 >= F - 1e-14
        F = F1
    M = np.array([m3(v) for v in W[::7]])
    assert np.max(np.abs(np.einsum("lij,lkj->lik", M, M.conj()) - np.eye(3))) < 1e-13
    assert abs(oracle.su3_plaquette(W, dims) - oracle.su3_plaquette(links, dims)) < 1e-13
    for R, T in [(1, 1), (2, 2)]:
        assert abs(oracle.su3_wilson_loop(W, dims, R, T) -
                   oracle.su3_wilson_loop(links, dims, R, T)) < 1e-13


def test_su3_local_update_does_not_lower_subgroup_functional(oracle):
    """For each subgroup the omega in [1,2] step cannot lower the local
    functional: checked on single-site lattices through the whole sweep
    (every half-sweep is monotone)."""
    dims = (2, 2, 2, 2)
    links = oracle.gen_su3_config(dims, 1.0, 35)
    F = oracle.su3_functional(links, dims)
    for _ in range(20):
        F1, _, _ = oracle.su3_link_sweep(links, dims, 1.9)
        assert F1 >= F - 1e-14
        F = F1


# ---------------------------------------------------------------- checks
def _abelian_t_field(dims, b):
    """U_x(x,y,z,t) = exp(i b t s3), all other links 1."""
    V = int(np.prod(dims))
    links = np.tile([1.0, 0, 0, 0], (4 * V, 1))
    for s in range(V):
        t = s // (dims[0] * dims[1] * dims[2])
        links[4 * s] = [np.cos(b * t), 0, 0, np.sin(b * t)]
    return links


def test_plaquette_and_wilson_loop_closed_forms(oracle):
    dims = (4, 4, 4, 8)
    b = 2 * np.pi / dims[3]
    links = _abelian_t_field(dims, b)
    # only the (x,t) plaquette is non-trivial: cos(b)
    assert abs(oracle.su2_plaquette(links, dims) - (5 + np.cos(b)) / 6) < 1e-14
    for R, T in [(1, 1), (2, 3), (3, 2)]:
        ref = (2 + np.cos(b * R * T)) / 3
        assert abs(oracle.su2_wilson_loop(links, dims, R, T) - ref) < 1e-14
    # SU(3) embedding of the same field
    l3 = np.array([v18(embed_su2(q)) for q in links])
    assert abs(oracle.su3_plaquette(l3, dims) - (5 + (2 * np.cos(b) + 1) / 3) / 6) < 1e-14
    ref3 = (2 + (2 * np.cos(b * 2 * 2) + 1) / 3) / 3
    assert abs(oracle.su3_wilson_loop(l3, dims, 2, 2) - ref3) < 1e-14


def test_pure_gauge_configs(oracle):
    """eps = 0: links are a pure gauge g(x) g(x+mu)^+, plaquette 1."""
    dims = (4, 4, 4, 4)
    l2 = oracle.gen_su2_config(dims, 0.0, 41)
    assert abs(oracle.su2_plaquette(l2, dims) - 1) < 1e-14
    assert abs(oracle.su2_wilson_loop(l2, dims, 2, 2) - 1) < 1e-14
    l3 = oracle.gen_su3_config(dims, 0.0, 42)
    assert abs(oracle.su3_plaquette(l3, dims) - 1) < 1e-14


def test_generated_configs_unitary_and_seeded(oracle):
    dims = (4, 4, 4, 4)
    a = oracle.gen_su2_config(dims, 0.5, 7)
    b = oracle.gen_su2_config(dims, 0.5, 7)
    c = oracle.gen_su2_config(dims, 0.5, 8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.max(np.abs(np.linalg.norm(a, axis=1) - 1)) < 1e-14
    l3 = oracle.gen_su3_config(dims, 0.4, 7)
    M = np.array([m3(v) for v in l3[::5]])
    assert np.max(np.abs(np.einsum("lij,lkj->lik", M, M.conj()) - np.eye(3))) < 1e-14
    assert np.max(np.abs(np.linalg.det(M) - 1)) < 1e-13
    # the gauge transformation randomises the functional (MAG not fixed)
    assert oracle.su2_functional(a, dims) < 0.3


def test_spins_uniform_on_sphere(oracle):
    dims = (8, 8, 8, 8)
    s = oracle.gen_spins_uniform(dims, 3)
    assert np.max(np.abs(np.linalg.norm(s, axis=1) - 1)) < 1e-14
    # uniform on S^2: each component has mean 0 and variance 1/3
    assert np.all(np.abs(s.mean(0)) < 0.02)
    assert np.all(np.abs(s.var(0) - 1 / 3) < 0.02)


def test_heat_bath_distribution(oracle):
    """s.h_hat follows p(x) ~ exp(k x) on [-1,1], k = |h|/T
    (mag.cpp:402-433): its mean is the Langevin function coth(k) - 1/k."""
    dims = (8, 8, 8, 8)
    links = oracle.gen_su2_config(dims, 0.4, 5)
    spins = oracle.gen_spins_uniform(dims, 6)
    V = int(np.prod(dims))
    T = 1.5
    # fields of the parity-0 sites are fixed during that half-sweep
    par = np.array([sum(np.unravel_index(s, dims[::-1])) % 2 for s in range(V)])
    sites = np.nonzero(par == 0)[0]
    H = np.array([oracle.contribution_site(links, spins, dims, s) for s in sites])
    oracle.su2_heatbath_halfsweep(links, spins, dims, 0, T, 99, 0)
    hn = np.linalg.norm(H, axis=1)
    x = np.einsum("ij,ij->i", spins[sites], H) / hn
    k = hn / T
    lang = 1 / np.tanh(k) - 1 / k
    var = 1 - 2 * lang / k - lang ** 2  # variance of x under exp(k x)
    z = np.sum(x - lang) / np.sqrt(np.sum(var))
    assert abs(z) < 5
    assert np.max(np.abs(np.linalg.norm(spins, axis=1) - 1)) < 1e-13


# ---------------------------------------------------------------- gauge copies
@pytest.mark.parametrize("group", [2, 3])
def test_gauge_copies_are_gauge_equivalent(oracle, group):
    """Copy 0 is the plain configuration; copy k > 0 differs link by link but
    holds the same gauge-invariant content (plaquette, Wilson loops)."""
    dims = (4, 4, 4, 6)
    gen = oracle.gen_su2_config if group == 2 else oracle.gen_su3_config
    plaq = oracle.su2_plaquette if group == 2 else oracle.su3_plaquette
    wl = oracle.su2_wilson_loop if group == 2 else oracle.su3_wilson_loop
    base = gen(dims, 0.5, 91)
    assert np.array_equal(gen(dims, 0.5, 91, copy=0), base)
    c1 = gen(dims, 0.5, 91, copy=1)
    c2 = gen(dims, 0.5, 91, copy=2)
    assert np.max(np.abs(c1 - base)) > 0.1 and np.max(np.abs(c2 - c1)) > 0.1
    for c in (c1, c2):
        assert abs(plaq(c, dims) - plaq(base, dims)) < 1e-13
        assert abs(wl(c, dims, 2, 3) - wl(base, dims, 2, 3)) < 1e-13
    # a different seed is a different field, not a gauge copy
    assert abs(plaq(gen(dims, 0.5, 92), dims) - plaq(base, dims)) > 1e-6
