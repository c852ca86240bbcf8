"""Oracle checks for the abelian projection and the monopole currents
(MatrixConverter, matrix.h:1059-1100; monopoles.cpp:44-122, 368-457,
844-884).  The reference holds no golden vectors for them; the restatement
is pinned by closed forms and by what the mathematics fixes: the currents
are integers, conserved on the dual lattice, invariant under abelian gauge
transformations of the angles, equal to the Levi-Civita form of the
DeGrand-Toussaint current, and (SU(3)) sum to zero over the colours."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

DIMS = [(4, 4, 4, 4), (4, 6, 5, 3)]


def _site_grid(dims):
    V = int(np.prod(dims))
    s = np.arange(V)
    c = np.stack([s % dims[0], (s // dims[0]) % dims[1], (s // (dims[0] * dims[1])) % dims[2],
                  s // (dims[0] * dims[1] * dims[2])], axis=1)
    return c


def _shift(dims, mu, n=1):
    """site index of x + n mu for every site x"""
    c = _site_grid(dims)
    c = c.copy()
    c[:, mu] = (c[:, mu] + n) % dims[mu]
    return c[:, 0] + dims[0] * (c[:, 1] + dims[1] * (c[:, 2] + dims[2] * c[:, 3]))


def _random_angles(dims, rng, scale=np.pi):
    return rng.uniform(-scale, scale, size=4 * int(np.prod(dims)))


def test_su2_angles_closed_form():
    rng = np.random.default_rng(1)
    dims = (2, 2, 2, 2)
    n = 4 * 16
    phi = rng.uniform(-np.pi, np.pi, n)
    links = np.zeros((n, 4))
    links[:, 0], links[:, 3] = np.cos(phi), np.sin(phi)
    assert np.allclose(O.su2_abelian_angles(links, dims), phi, atol=1e-15, rtol=0)
    a = rng.normal(size=(n, 4))
    a /= np.linalg.norm(a, axis=1)[:, None]
    assert np.abs(O.su2_abelian_angles(a, dims) - np.arctan2(a[:, 3], a[:, 0])).max() < 1e-15


def test_su3_angles_closed_form():
    rng = np.random.default_rng(2)
    dims = (2, 2, 2, 2)
    n = 4 * 16
    al, be = rng.uniform(-np.pi, np.pi, (2, n))
    ph = np.stack([al, be, -(al + be)], axis=1)
    m = np.zeros((n, 3, 3), complex)
    for c in range(3):
        m[:, c, c] = np.exp(1j * ph[:, c])
    links18 = np.stack([m.real, m.imag], axis=-1).reshape(n, 18)
    ang = O.su3_abelian_angles(np.ascontiguousarray(links18), dims)
    # each angle equals its diagonal phase modulo 2 pi, and the three sum to
    # a multiple of 2 pi with the wrapped third removed
    d = np.angle(np.exp(1j * (ang - ph)))
    assert np.abs(d).max() < 1e-12
    s = ang.sum(axis=1)
    assert np.abs(s / (2 * np.pi) - np.round(s / (2 * np.pi))).max() < 1e-12
    # general unitary links: phases of the diagonal less a third of their
    # sum wrapped into [-pi, pi)
    z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, _ = np.linalg.qr(z)
    l18 = np.ascontiguousarray(np.stack([q.real, q.imag], axis=-1).reshape(n, 18))
    ang = O.su3_abelian_angles(l18, dims)
    raw = np.angle(np.stack([q[:, c, c] for c in range(3)], axis=1))
    tot = raw.sum(axis=1)
    tot = (tot + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(ang - (raw - tot[:, None] / 3)).max() < 1e-14


@pytest.mark.parametrize("dims", DIMS)
def test_u1_plaquettes_brute_force(dims):
    rng = np.random.default_rng(3)
    th = _random_angles(dims, rng)
    P = O.u1_monopole_plaquettes(th, dims)
    t = th.reshape(-1, 4)
    for (mu, nu), pl in zip(itertools.combinations(range(4), 2), range(6)):
        raw = t[:, mu] + t[_shift(dims, mu), nu] - t[_shift(dims, nu), mu] - t[:, nu]
        ref = raw - 2 * np.pi * np.round(raw / (2 * np.pi))
        assert np.abs(P[pl] - ref).max() < 1e-12
        assert np.all(np.abs(P[pl]) <= np.pi)


def _levi_civita_currents(P, dims):
    """-(1/4 pi) eps_{mu nu rho sigma} [P_rs(x + mu + nu) - P_rs(x + mu)],
    P_rs = -P_sr, eps_{0123} = +1"""
    V = int(np.prod(dims))
    plane = {pair: i for i, pair in enumerate(itertools.combinations(range(4), 2))}

    def Pf(r, s):
        return P[plane[(r, s)]] if r < s else -P[plane[(s, r)]]

    J = np.zeros((V, 4))
    for perm in itertools.permutations(range(4)):
        mu, nu, rho, sig = perm
        inv = sum(1 for i in range(4) for j in range(i + 1, 4) if perm[i] > perm[j])
        eps = -1.0 if inv % 2 else 1.0
        f = Pf(rho, sig)
        xm = _shift(dims, mu)
        # x + mu + nu
        c = _site_grid(dims)
        c[:, mu] = (c[:, mu] + 1) % dims[mu]
        c[:, nu] = (c[:, nu] + 1) % dims[nu]
        xmn = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * (c[:, 2] + dims[2] * c[:, 3]))
        J[:, mu] += -eps * (f[xmn] - f[xm]) / (4 * np.pi)
    return J.reshape(-1)


@pytest.mark.parametrize("dims", DIMS)
def test_u1_currents(dims):
    rng = np.random.default_rng(4)
    th = _random_angles(dims, rng)
    P = O.u1_monopole_plaquettes(th, dims)
    J = O.monopole_currents(P, dims)
    Jr = np.round(J)
    # integers, some of them non-zero
    assert np.abs(J - Jr).max() < 1e-12
    assert np.count_nonzero(Jr) > 10
    # conserved: sum_mu J_mu(x) - J_mu(x - mu) = 0 on every dual site
    j = Jr.reshape(-1, 4)
    div = sum(j[:, mu] - j[_shift(dims, mu, -1), mu] for mu in range(4))
    assert np.all(div == 0)
    # the Levi-Civita form
    assert np.abs(J - _levi_civita_currents(P, dims)).max() < 1e-12
    # abelian gauge invariance: theta_mu(x) + a(x) - a(x + mu)
    a = rng.uniform(-np.pi, np.pi, int(np.prod(dims)))
    t2 = th.reshape(-1, 4).copy()
    for mu in range(4):
        t2[:, mu] += a - a[_shift(dims, mu)]
    J2 = O.monopole_currents(O.u1_monopole_plaquettes(np.ascontiguousarray(t2.reshape(-1)), dims),
                             dims)
    assert np.array_equal(np.round(J2), Jr)


def test_single_dirac_sheet():
    """Angles pi - d on the x links of one (y, z) column over all t, 0
    elsewhere, put a monopole loop where the sheet ends: every current is an
    integer and their total |J| is positive and even (closed loops)."""
    dims = (4, 4, 4, 4)
    th = np.zeros(4 * 256)
    s = np.arange(256)
    c = _site_grid(dims)
    sel = (c[:, 1] == 1) & (c[:, 2] == 2) & (c[:, 0] == 0)
    th[4 * s[sel] + 0] = np.pi - 0.1
    th[4 * s[sel & (c[:, 3] == 0)] + 0] += 0.3  # a kink so the sheet has an edge
    J = O.monopole_currents(O.u1_monopole_plaquettes(th, dims), dims)
    assert np.abs(J - np.round(J)).max() < 1e-12
    tot = int(np.abs(np.round(J)).sum())
    assert tot > 0 and tot % 2 == 0


@pytest.mark.parametrize("dims", DIMS)
def test_su3_currents(dims):
    rng = np.random.default_rng(5)
    V = int(np.prod(dims))
    a = rng.uniform(-np.pi, np.pi, (4 * V, 2))
    ang = np.ascontiguousarray(np.stack([a[:, 0], a[:, 1], -(a[:, 0] + a[:, 1])], axis=1))
    P = O.su3_monopole_plaquettes(ang, dims)
    # the colour sum of every plaquette is brought to 0 (mod rounding)
    assert np.abs(P.sum(axis=0)).max() < 1e-12
    Js = []
    for c in range(3):
        # each colour: the U(1) plaquettes of that colour, up to the 2 pi moved
        # by the colour-sum rule
        Pc = O.u1_monopole_plaquettes(np.ascontiguousarray(ang[:, c]), dims)
        d = (P[c] - Pc) / (2 * np.pi)
        assert np.abs(d - np.round(d)).max() < 1e-12
        J = O.monopole_currents(np.ascontiguousarray(P[c]), dims)
        assert np.abs(J - np.round(J)).max() < 1e-12
        j = np.round(J).reshape(-1, 4)
        div = sum(j[:, mu] - j[_shift(dims, mu, -1), mu] for mu in range(4))
        assert np.all(div == 0)
        Js.append(np.round(J))
    assert np.all(Js[0] + Js[1] + Js[2] == 0)
    assert np.count_nonzero(Js[0]) > 10
