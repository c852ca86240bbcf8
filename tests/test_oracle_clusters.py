"""Oracle checks for the monopole clusters (calculate_clusters,
monopoles.cpp:1120-1303; cluster_length / length_mu / currents_directions,
1343-1449; site_number, 1510-1531) and for the app's cluster tables
(apps/monopoles_su2/monopoles_su2.cpp:110-197, group_percolating
monopoles.cpp:1533-1590).  The reference holds no golden vectors for them;
the walk is pinned by an independent union-find (scipy connected
components), by totals that must equal sums over all links, and by
hand-built currents with known clusters, windings and charges; the
percolating-group search by hand cases, including one its enumeration
skips.  The C++ drop-in's tables are checked against the oracle's."""
import os
import subprocess

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lround(x):
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def _fwd(dims, mu):
    """site index of x + mu for every site x"""
    V = int(np.prod(dims))
    s = np.arange(V)
    c = [s % dims[0], (s // dims[0]) % dims[1], (s // (dims[0] * dims[1])) % dims[2],
         s // (dims[0] * dims[1] * dims[2])]
    c[mu] = (c[mu] + 1) % dims[mu]
    return c[0] + dims[0] * (c[1] + dims[1] * (c[2] + dims[2] * c[3]))


def union_find_clusters(J, dims):
    """The same records from scipy's connected components of the graph of
    links with |J| > 0.3 (an algorithm independent of the reference's walk)."""
    V = int(np.prod(dims))
    Jm = J.reshape(V, 4)
    src, dst, mus = [], [], []
    for mu in range(4):
        on = np.nonzero(np.abs(Jm[:, mu]) > 0.3)[0]
        src.append(on)
        dst.append(_fwd(dims, mu)[on])
        mus.append(np.full(on.size, mu))
    src, dst, mus = np.concatenate(src), np.concatenate(dst), np.concatenate(mus)
    if src.size == 0:
        return np.zeros((0, 9), dtype=np.int32)
    g = coo_matrix((np.ones(src.size), (src, dst)), shape=(V, V))
    _, lab = connected_components(g, directed=False)
    q = _lround(Jm[src, mus])
    out = []
    for c in np.unique(lab[src]):
        e = lab[src] == c
        sites = np.unique(np.concatenate([src[e], dst[e]]))
        lmu = [int(q[e & (mus == mu)].sum()) for mu in range(4)]
        time = int(np.abs(q[e & (mus == 3)]).sum())
        length = int(np.abs(q[e]).sum())
        out.append([int(sites.min()), length, *lmu, length - time, time, sites.size])
    out.sort(key=lambda r: r[0])
    return np.array(out, dtype=np.int32)


def _site(dims, c):
    return c[0] + dims[0] * (c[1] + dims[1] * (c[2] + dims[2] * c[3]))


def _step(dims, c, mu, n=1):
    c = list(c)
    c[mu] = (c[mu] + n) % dims[mu]
    return c


def add_loop(J, dims, c, mu, nu, q=1):
    """the boundary of the plaquette (c, mu, nu): charge q around it"""
    J[4 * _site(dims, c) + mu] += q
    J[4 * _site(dims, _step(dims, c, mu)) + nu] += q
    J[4 * _site(dims, _step(dims, c, nu)) + mu] -= q
    J[4 * _site(dims, c) + nu] -= q


def add_line(J, dims, c, mu, q=1):
    """a straight line through c along mu, closed around the lattice"""
    for k in range(dims[mu]):
        J[4 * _site(dims, _step(dims, c, mu, k)) + mu] += q


def random_currents(dims, sigma, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(0.0, sigma, 4 * int(np.prod(dims)))
    return O.monopole_currents(O.u1_monopole_plaquettes(a, dims), dims)


@pytest.mark.parametrize("dims,sigma", [((6, 6, 6, 6), 0.6), ((6, 6, 6, 6), 0.8),
                                        ((4, 6, 4, 8), 0.7), ((2, 4, 2, 4), 1.0)])
def test_walk_equals_connected_components(dims, sigma):
    J = random_currents(dims, sigma, 11)
    rec = O.monopole_clusters(J, dims)
    assert len(rec) > 0
    np.testing.assert_array_equal(rec, union_find_clusters(J, dims))


def test_su3_colours_equal_connected_components():
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(5)
    a3 = rng.normal(0.0, 0.8, (4 * int(np.prod(dims)), 3))
    a3 -= a3.mean(axis=1, keepdims=True)  # su3_angles sum to zero
    P = O.su3_monopole_plaquettes(a3, dims)
    for k in range(3):
        J = O.monopole_currents(np.ascontiguousarray(P[k]), dims)
        np.testing.assert_array_equal(O.monopole_clusters(J, dims), union_find_clusters(J, dims))


def test_totals_are_sums_over_links():
    dims = (6, 6, 6, 6)
    J = random_currents(dims, 0.8, 3)
    rec = O.monopole_clusters(J, dims)
    q = _lround(J).reshape(-1, 4)
    q[np.abs(J.reshape(-1, 4)) <= 0.3] = 0
    assert rec[:, 1].sum() == np.abs(q).sum()
    for mu in range(4):
        assert rec[:, 2 + mu].sum() == q[:, mu].sum()
    assert rec[:, 7].sum() == np.abs(q[:, 3]).sum()
    assert (rec[:, 6] + rec[:, 7] == rec[:, 1]).all()
    assert (np.diff(rec[:, 0]) > 0).all()  # the reference's order: first sites ascending


def test_hand_built_clusters():
    dims = (4, 4, 6, 8)
    V = int(np.prod(dims))
    J = np.zeros(4 * V)
    add_loop(J, dims, [1, 1, 1, 1], 0, 1)          # space-space plaquette
    add_loop(J, dims, [2, 2, 4, 5], 2, 3, q=2)      # charge 2, z-t plane
    add_line(J, dims, [3, 0, 0, 0], 3)              # wraps t once
    add_line(J, dims, [0, 3, 3, 3], 0, q=-1)        # wraps x once, backwards
    rec = O.monopole_clusters(J, dims)
    s = lambda c: _site(dims, c)
    want = sorted([
        [s([1, 1, 1, 1]), 4, 0, 0, 0, 0, 4, 0, 4],
        [s([2, 2, 4, 5]), 8, 0, 0, 0, 0, 4, 4, 4],
        [s([3, 0, 0, 0]), 8, 0, 0, 0, 8, 0, 8, 8],
        [s([0, 3, 3, 3]), 4, -4, 0, 0, 0, 4, 0, 4],
    ])
    np.testing.assert_array_equal(rec, np.array(want, dtype=np.int32))
    # two loops through one site are one cluster
    J2 = np.zeros(4 * V)
    add_loop(J2, dims, [1, 1, 1, 1], 0, 1)
    add_loop(J2, dims, [1, 1, 1, 1], 2, 3)
    rec2 = O.monopole_clusters(J2, dims)
    np.testing.assert_array_equal(rec2, [[s([1, 1, 1, 1]), 8, 0, 0, 0, 0, 6, 2, 7]])
    np.testing.assert_array_equal(rec2, union_find_clusters(J2, dims))


def test_summary_windings_and_percolating_group():
    dims = (4, 4, 4, 6)
    V = int(np.prod(dims))
    J = np.zeros(4 * V)
    add_line(J, dims, [0, 0, 0, 0], 3, 1)
    add_line(J, dims, [2, 2, 0, 0], 3, -1)
    add_line(J, dims, [0, 2, 2, 0], 3, 1)
    add_loop(J, dims, [1, 1, 1, 1], 0, 1)
    add_loop(J, dims, [3, 1, 3, 1], 0, 2)
    summ = O.cluster_summary(O.monopole_clusters(J, dims), dims)
    assert summ["unwrapped"] == {4: 2}
    assert summ["wrapped"] == [(6, [0, 0, 0, 1], True), (6, [0, 0, 0, -1], True), (6, [0, 0, 0, 1], False)]
    assert summ["asymmetry"] == pytest.approx((8 / 3 - 18) / (8 / 3 + 18), rel=1e-15)


def test_group_percolating_enumeration():
    assert O.group_percolating([[0, 0, 0, 1], [0, 0, 0, -1]]) == [0, 1]
    assert O.group_percolating([[1, 0, 0, 0], [1, 0, 0, 0], [-1, 0, 0, 0]]) == [0, 2]
    assert O.group_percolating([[1, 0, 0, 0], [0, 1, 0, 0]]) == []
    # order 3 over 5: {0, 1, 4} is reached by running the last position
    w = [[2, 0, 0, 0], [-1, 0, 0, 0], [5, 0, 0, 0], [5, 0, 0, 0], [-1, 0, 0, 0]]
    assert O.group_percolating(w) == [0, 1, 4]
    # {0, 2, 3} of four: the inner position moves only while its successor
    # is more than 2 ahead, so the reference never tries it and finds none
    w = [[2, 0, 0, 0], [5, 0, 0, 0], [-1, 0, 0, 0], [-1, 0, 0, 0]]
    assert O.group_percolating(w) == []
    # with one more cluster in between the same group is reached
    w = [[2, 0, 0, 0], [5, 0, 0, 0], [-1, 0, 0, 0], [7, 0, 0, 0], [-1, 0, 0, 0]]
    assert O.group_percolating(w) == [0, 2, 4]


@pytest.fixture(scope="module")
def tables_exe(tmp_path_factory, mag):
    exe = str(tmp_path_factory.mktemp("bin") / "clusters")
    pkg = os.path.join(ROOT, "general_code_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(pkg, "cpp"), os.path.join(ROOT, "tests", "cpp", "clusters_main.cpp"),
                    "-L", pkg, "-lmag_gpu", f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    return exe


def cpp_tables(exe, rec, dims):
    inp = " ".join(map(str, [*dims, len(rec), *np.asarray(rec).reshape(-1).tolist()]))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    unwrapped, wrapped, asym = {}, [], None
    for line in out.splitlines():
        f = line.split()
        if f[0] == "U":
            unwrapped[int(f[1])] = int(f[2])
        elif f[0] == "W":
            wrapped.append((int(f[1]), [int(v) for v in f[2:6]], f[6] == "1"))
        else:
            asym = float(f[1])
    return {"unwrapped": unwrapped, "wrapped": wrapped, "asymmetry": asym}


def _records(rows, dims):
    """records with given (length, windings) rows, lengths_mu = windings * L"""
    return np.array([[i, L, *[w[mu] * dims[mu] for mu in range(4)], L, 0, 4] for i, (L, w) in enumerate(rows)],
                    dtype=np.int32)


def test_cpp_tables_match_oracle(tables_exe):
    dims = (6, 6, 6, 6)
    for sigma, seed in [(0.6, 1), (0.8, 2), (1.0, 3)]:
        rec = O.monopole_clusters(random_currents(dims, sigma, seed), dims)
        assert cpp_tables(tables_exe, rec, dims) == O.cluster_summary(rec, dims)
    d2 = (4, 4, 4, 6)
    for rows in ([(40, [2, 0, 0, 0]), (30, [5, 0, 0, 0]), (20, [-1, 0, 0, 0]), (10, [-1, 0, 0, 0])],
                 [(40, [2, 0, 0, 0]), (30, [5, 0, 0, 0]), (20, [-1, 0, 0, 0]), (15, [7, 0, 0, 0]),
                  (10, [-1, 0, 0, 0])],
                 [(12, [0, 0, 0, 1]), (12, [0, 1, 0, 0]), (18, [0, 0, 0, -1]), (6, [0, -1, 0, 0])]):
        rec = _records(rows, d2)
        got, want = cpp_tables(tables_exe, rec, d2), O.cluster_summary(rec, d2)
        assert got == want, (rows, got, want)
