"""ctypes front end of the CPU oracle (libmag_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product (general_code_amd) never does.

Arrays follow the reference layout (lib/cpu/include/indexing.h:71-126):
SU(2) links are float64 [V*4, 4], SU(3) links float64 [V*4, 18] (row-major
re/im, lib/cpu/include/matrix.h:236-254), spins float64 [V, 3].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmag_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


_REF_LIME = os.path.join(_HERE, "_ref", "liblime_ref.so")
_lime = None


def build_ref():
    """Compile the reference's LIME library from its sources under
    /root/reference (oracle/Makefile target `ref`, output in oracle/_ref/);
    returns its path, or None where the reference is absent."""
    if not os.path.isdir("/root/reference/lib/cpu/src/c-lime"):
        return _REF_LIME if os.path.exists(_REF_LIME) else None
    subprocess.run(["make", "-C", _HERE, "ref"], check=True, stdout=subprocess.DEVNULL)
    return _REF_LIME


def lime_ref():
    """ctypes handle of oracle/_ref/liblime_ref.so (None if not built)."""
    global _lime
    if _lime is None:
        if not os.path.exists(_REF_LIME):
            return None
        L = ctypes.CDLL(_REF_LIME)
        vp, u64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)
        for name, res, args in [
            ("limeCreateWriter", vp, [vp]), ("limeDestroyWriter", ctypes.c_int, [vp]),
            ("limeCreateHeader", vp, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64]),
            ("limeDestroyHeader", None, [vp]),
            ("limeWriteRecordHeader", ctypes.c_int, [vp, vp]),
            ("limeWriteRecordData", ctypes.c_int, [vp, u64p, vp]),
            ("limeWriterCloseRecord", ctypes.c_int, [vp]),
            ("limeCreateReader", vp, [vp]), ("limeDestroyReader", None, [vp]),
            ("limeReaderNextRecord", ctypes.c_int, [vp]),
            ("limeReaderMBFlag", ctypes.c_int, [vp]), ("limeReaderMEFlag", ctypes.c_int, [vp]),
            ("limeReaderType", ctypes.c_char_p, [vp]), ("limeReaderBytes", ctypes.c_uint64, [vp]),
            ("limeReaderReadData", ctypes.c_int, [vp, u64p, vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lime = L
    return _lime


def _libc():
    c = ctypes.CDLL(None)
    c.fopen.restype = ctypes.c_void_p
    c.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    c.fclose.argtypes = [ctypes.c_void_p]
    return c


def lime_write(path, records):
    """Write [(type, bytes), ...] as one LIME message with the reference's
    writer (limeCreateHeader / limeWriteRecordHeader / limeWriteRecordData)."""
    L, c = lime_ref(), _libc()
    fp = c.fopen(str(path).encode(), b"wb")
    w = L.limeCreateWriter(fp)
    for i, (t, data) in enumerate(records):
        h = L.limeCreateHeader(int(i == 0), int(i == len(records) - 1), t.encode(), len(data))
        assert L.limeWriteRecordHeader(h, w) == 0
        n = ctypes.c_uint64(len(data))
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        assert L.limeWriteRecordData(buf, ctypes.byref(n), w) == 0
        L.limeWriterCloseRecord(w)
        L.limeDestroyHeader(h)
    L.limeDestroyWriter(w)
    c.fclose(fp)


def lime_read(path):
    """[(type, MB, ME, bytes), ...] of a LIME file, with the reference's reader."""
    L, c = lime_ref(), _libc()
    fp = c.fopen(str(path).encode(), b"rb")
    r = L.limeCreateReader(fp)
    out = []
    while True:
        st = L.limeReaderNextRecord(r)
        if st == -7:  # LIME_EOF
            break
        assert st == 0, st
        n = L.limeReaderBytes(r)
        buf = ctypes.create_string_buffer(int(n))
        nn = ctypes.c_uint64(n)
        assert L.limeReaderReadData(buf, ctypes.byref(nn), r) == 0
        out.append((L.limeReaderType(r).decode(), L.limeReaderMBFlag(r), L.limeReaderMEFlag(r),
                    buf.raw[:n]))
    L.limeDestroyReader(r)
    c.fclose(fp)
    return out


def _bind(L):
    """Set the ctypes signatures of every oracle entry point on handle L."""
    sig = {
        "or_su2_mul": (None, [_dp, _dp, _dp]),
        "or_su2_mul_dag": (None, [_dp, _dp, _dp]),
        "or_su2_dag_mul": (None, [_dp, _dp, _dp]),
        "or_su2_link_functional": (ctypes.c_double, [_dp]),
        "or_spin_contribution1": (None, [_dp, _dp, _dp]),
        "or_spin_contribution1_conj": (None, [_dp, _dp, _dp]),
        "or_spin_reflect_fast": (None, [_dp, _dp]),
        "or_spin_parallel": (ctypes.c_double, [_dp, _dp]),
        "or_spin_gauge_matrix": (None, [_dp, _dp]),
        "or_su2_contribution_site": (None, [_dp, _dp, _ip, ctypes.c_int64, _dp]),
        "or_su2_functional": (ctypes.c_double, [_dp, _ip]),
        "or_su2_functional_spin": (ctypes.c_double, [_dp, _dp, _ip]),
        "or_su2_gauge_transform_spins": (None, [_dp, _dp, _ip]),
        "or_su2_or_sweep_lexi": (None, [_dp, _dp, _ip]),
        "or_su2_or_sweep_lexi_omp": (None, [_dp, _dp, _ip]),
        "or_su2_relax_sweep_lexi": (None, [_dp, _dp, _ip, _dp, _dp]),
        "or_su2_maximize_final": (ctypes.c_int, [_dp, _dp, _ip, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_int]),
        "or_atan2_pos": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "or_sincos": (None, [ctypes.c_double, _dp, _dp]),
        "or_approx_same": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int]),
        "or_su2_maximize_approximate": (ctypes.c_int, [_dp, _dp, _ip, ctypes.c_int,
                                                       ctypes.c_int, ctypes.c_int]),
        "or_su2_spin_halfsweep": (None, [_dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, _dp, _dp]),
        "or_su2_heatbath_halfsweep": (None, [_dp, _dp, _ip, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_uint64,
                                             ctypes.c_uint32]),
        "or_su2_link_sweep": (None, [_dp, _ip, ctypes.c_double, _dp, _dp, _dp]),
        "or_su3_link_sweep": (None, [_dp, _ip, ctypes.c_double, _dp, _dp, _dp]),
        "or_omega_gauge": (None, [_dp, ctypes.c_double, _dp]),
        "or_block_field_fwd": (None, [_dp, _dp]),
        "or_block_field_bwd": (None, [_dp, _dp]),
        "or_su3_functional": (ctypes.c_double, [_dp, _ip]),
        "or_su3_reunitarize": (None, [_dp]),
        "or_su3_proj": (None, [_dp]),
        "or_su3_mul": (None, [_dp, _dp, _dp]),
        "or_su3_mul_dag": (None, [_dp, _dp, _dp]),
        "or_su2_plaquette": (ctypes.c_double, [_dp, _ip]),
        "or_su3_plaquette": (ctypes.c_double, [_dp, _ip]),
        "or_su2_wilson_loop": (ctypes.c_double, [_dp, _ip, ctypes.c_int, ctypes.c_int]),
        "or_su3_wilson_loop": (ctypes.c_double, [_dp, _ip, ctypes.c_int, ctypes.c_int]),
        "or_su2_abelian_angles": (None, [_dp, _ip, _dp]),
        "or_su3_abelian_angles": (None, [_dp, _ip, _dp]),
        "or_u1_monopole_plaquettes": (None, [_dp, _ip, _dp]),
        "or_su3_monopole_plaquettes": (None, [_dp, _ip, _dp]),
        "or_monopole_currents": (None, [_dp, _ip, _dp]),
        "or_monopole_clusters": (ctypes.c_int64, [_dp, _ip, ctypes.POINTER(ctypes.c_int32),
                                                  ctypes.c_int64]),
        "or_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
        "or_gen_su2_config": (None, [_dp, _ip, ctypes.c_double, ctypes.c_uint64]),
        "or_gen_su3_config": (None, [_dp, _ip, ctypes.c_double, ctypes.c_uint64]),
        "or_gen_copy_spins": (None, [_dp, _ip, ctypes.c_uint64, ctypes.c_uint32]),
        "or_gen_su2_config_copy": (None, [_dp, _ip, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.c_uint32]),
        "or_gen_su3_config_copy": (None, [_dp, _ip, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.c_uint32]),
        "or_gen_su3_exp": (None, [_dp, ctypes.c_double, _dp]),
        "or_gen_spins_uniform": (None, [_dp, _ip, ctypes.c_uint64]),
        "or_set_threads": (None, [ctypes.c_int]),
        "or_get_threads": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


_fast = None


def fast_lib():
    """The same C restatement compiled with the reference's own flags
    (-Ofast -march=native, /root/reference makefile:5) for the CPU-baseline
    timing only (never as the checker: -Ofast allows reassociation).  Built
    on first use for the machine it runs on (oracle/Makefile target `fast`)
    into a temporary directory, so no binary for another CPU is reused."""
    global _fast
    if _fast is None:
        import tempfile
        out = tempfile.mkdtemp(prefix="mag_oracle_fast_")
        subprocess.run(["make", "-C", _HERE, "fast", f"FAST_OUT={out}"], check=True,
                       stdout=subprocess.DEVNULL)
        _fast = _bind(ctypes.CDLL(os.path.join(out, "libmag_oracle_fast.so")))
    return _fast


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _dims(dims):
    arr = (ctypes.c_int * 4)(*[int(d) for d in dims])
    return arr


def volume(dims) -> int:
    return int(np.prod([int(d) for d in dims]))


# ---------------------------------------------------------------- algebra
def su2_mul(a, b):
    out = np.zeros(4)
    lib().or_su2_mul(_d(np.ascontiguousarray(a, np.float64)),
                     _d(np.ascontiguousarray(b, np.float64)), _d(out))
    return out


def su2_mul_dag(a, b):
    out = np.zeros(4)
    lib().or_su2_mul_dag(_d(np.ascontiguousarray(a, np.float64)),
                         _d(np.ascontiguousarray(b, np.float64)), _d(out))
    return out


def su2_dag_mul(a, b):
    out = np.zeros(4)
    lib().or_su2_dag_mul(_d(np.ascontiguousarray(a, np.float64)),
                         _d(np.ascontiguousarray(b, np.float64)), _d(out))
    return out


def su2_link_functional(a):
    return lib().or_su2_link_functional(_d(np.ascontiguousarray(a, np.float64)))


def contribution1(h, A, b):
    h = np.array(h, np.float64)
    lib().or_spin_contribution1(_d(h), _d(np.ascontiguousarray(A, np.float64)),
                                _d(np.ascontiguousarray(b, np.float64)))
    return h


def contribution1_conj(h, A, b):
    h = np.array(h, np.float64)
    lib().or_spin_contribution1_conj(_d(h), _d(np.ascontiguousarray(A, np.float64)),
                                     _d(np.ascontiguousarray(b, np.float64)))
    return h


def reflect_fast(s, V):
    s = np.array(s, np.float64)
    lib().or_spin_reflect_fast(_d(s), _d(np.ascontiguousarray(V, np.float64)))
    return s


def parallel(s, V):
    s = np.array(s, np.float64)
    d = lib().or_spin_parallel(_d(s), _d(np.ascontiguousarray(V, np.float64)))
    return s, d


def gauge_matrix(s):
    g = np.zeros(4)
    lib().or_spin_gauge_matrix(_d(np.ascontiguousarray(s, np.float64)), _d(g))
    return g


def omega_gauge(h, omega):
    g = np.zeros(4)
    lib().or_omega_gauge(_d(np.ascontiguousarray(h, np.float64)), float(omega), _d(g))
    return g


def block_field_fwd(B):
    h = np.zeros(3)
    lib().or_block_field_fwd(_d(np.ascontiguousarray(B, np.float64)), _d(h))
    return h


def block_field_bwd(B):
    h = np.zeros(3)
    lib().or_block_field_bwd(_d(np.ascontiguousarray(B, np.float64)), _d(h))
    return h


# ---------------------------------------------------------------- SU(2) MAG
def contribution_site(links, spins, dims, site):
    h = np.zeros(3)
    lib().or_su2_contribution_site(_d(links), _d(spins), _dims(dims), int(site), _d(h))
    return h


def su2_functional(links, dims):
    return lib().or_su2_functional(_d(links), _dims(dims))


def su2_functional_spin(links, spins, dims):
    return lib().or_su2_functional_spin(_d(links), _d(spins), _dims(dims))


def su2_gauge_transform_spins(links, spins, dims):
    out = links.copy()
    lib().or_su2_gauge_transform_spins(_d(out), _d(spins), _dims(dims))
    return out


def su2_or_sweep_lexi(links, spins, dims):
    lib().or_su2_or_sweep_lexi(_d(links), _d(spins), _dims(dims))


def su2_or_sweep_lexi_omp(links, spins, dims):
    """Timing baseline: the reference's in-place OpenMP sweep (racy)."""
    lib().or_su2_or_sweep_lexi_omp(_d(links), _d(spins), _dims(dims))


def su2_relax_sweep_lexi(links, spins, dims):
    mx = ctypes.c_double()
    av = ctypes.c_double()
    lib().or_su2_relax_sweep_lexi(_d(links), _d(spins), _dims(dims),
                                  ctypes.byref(mx), ctypes.byref(av))
    return mx.value, av.value


def su2_maximize_final(links, spins, dims, or_steps, tol_max, tol_avg, max_iter=100000):
    return lib().or_su2_maximize_final(_d(links), _d(spins), _dims(dims), int(or_steps),
                                       float(tol_max), float(tol_avg), int(max_iter))


def su2_maximize_approximate(links, spins, dims, or_steps, tol_digits, max_iter=100000):
    return lib().or_su2_maximize_approximate(_d(links), _d(spins), _dims(dims),
                                             int(or_steps), int(tol_digits), int(max_iter))


def atan2_pos(y, x) -> float:
    """The omega rule's atan2 (y > 0): a fixed sequence of IEEE operations."""
    return lib().or_atan2_pos(float(y), float(x))


def sincos(a):
    """The omega rule's (sin a, cos a) for a in [0, 2 pi]."""
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_sincos(float(a), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def approx_same(F, F_old, digits) -> bool:
    """The stopping test of make_maximization_approximate (mag.cpp:1744-1746)."""
    return bool(lib().or_approx_same(float(F), float(F_old), int(digits)))


def su2_spin_halfsweep(links, spins, dims, parity, mode, omega=1.0):
    mx = ctypes.c_double()
    sm = ctypes.c_double()
    lib().or_su2_spin_halfsweep(_d(links), _d(spins), _dims(dims), int(parity), int(mode),
                                float(omega), ctypes.byref(mx), ctypes.byref(sm))
    return mx.value, sm.value


def su2_heatbath_halfsweep(links, spins, dims, parity, temperature, seed, sweep):
    lib().or_su2_heatbath_halfsweep(_d(links), _d(spins), _dims(dims), int(parity),
                                    float(temperature), int(seed), int(sweep))


def su2_link_sweep(links, dims, omega):
    F = ctypes.c_double()
    mx = ctypes.c_double()
    av = ctypes.c_double()
    lib().or_su2_link_sweep(_d(links), _dims(dims), float(omega), ctypes.byref(F),
                            ctypes.byref(mx), ctypes.byref(av))
    return F.value, mx.value, av.value


def su3_link_sweep(links, dims, omega):
    F = ctypes.c_double()
    mx = ctypes.c_double()
    av = ctypes.c_double()
    lib().or_su3_link_sweep(_d(links), _dims(dims), float(omega), ctypes.byref(F),
                            ctypes.byref(mx), ctypes.byref(av))
    return F.value, mx.value, av.value


# ---------------------------------------------------------------- SU(3)
def su3_functional(links, dims):
    return lib().or_su3_functional(_d(links), _dims(dims))


def su3_reunitarize(m):
    m = np.array(m, np.float64)
    lib().or_su3_reunitarize(_d(m))
    return m


def su3_proj(m):
    m = np.array(m, np.float64)
    lib().or_su3_proj(_d(m))
    return m


def su3_mul(a, b):
    out = np.zeros(18)
    lib().or_su3_mul(_d(np.ascontiguousarray(a, np.float64)),
                     _d(np.ascontiguousarray(b, np.float64)), _d(out))
    return out


def su3_mul_dag(a, b):
    out = np.zeros(18)
    lib().or_su3_mul_dag(_d(np.ascontiguousarray(a, np.float64)),
                         _d(np.ascontiguousarray(b, np.float64)), _d(out))
    return out


# ---------------------------------------------------------------- checks
def su2_plaquette(links, dims):
    return lib().or_su2_plaquette(_d(links), _dims(dims))


def su3_plaquette(links, dims):
    return lib().or_su3_plaquette(_d(links), _dims(dims))


def su2_wilson_loop(links, dims, R, T):
    return lib().or_su2_wilson_loop(_d(links), _dims(dims), int(R), int(T))


def su3_wilson_loop(links, dims, R, T):
    return lib().or_su3_wilson_loop(_d(links), _dims(dims), int(R), int(T))


# ---------------------------------------------------------------- generators
def su2_abelian_angles(links, dims):
    """phi = atan2(a3, a0) per link, [4V] (matrix.h:1059-1062)."""
    out = np.empty(4 * volume(dims))
    lib().or_su2_abelian_angles(_d(links), _dims(dims), _d(out))
    return out


def su3_abelian_angles(links18, dims):
    """su3_angles of every link, [4V, 3] (matrix.h:1082-1100)."""
    out = np.empty((4 * volume(dims), 3))
    lib().or_su3_abelian_angles(_d(links18), _dims(dims), _d(out))
    return out


def u1_monopole_plaquettes(angles, dims):
    """[6, V] plaquette angles in [-pi, pi] (monopoles.cpp:415-457)."""
    out = np.empty((6, volume(dims)))
    lib().or_u1_monopole_plaquettes(_d(angles), _dims(dims), _d(out))
    return out


def su3_monopole_plaquettes(angles3, dims):
    """[3, 6, V] per-colour plaquette angles with the colour-sum rule
    (monopoles.cpp:368-413, 844-884)."""
    out = np.empty((3, 6, volume(dims)))
    lib().or_su3_monopole_plaquettes(_d(angles3), _dims(dims), _d(out))
    return out


def monopole_currents(plaq, dims):
    """[4V] currents J[4 site + mu] from [6, V] plaquettes (monopoles.cpp:44-122)."""
    out = np.empty(4 * volume(dims))
    lib().or_monopole_currents(_d(plaq), _dims(dims), _d(out))
    return out


CLUSTER_FIELDS = ("site", "length", "l0", "l1", "l2", "l3", "space", "time", "sites")


def monopole_clusters(J, dims):
    """Clusters of the currents J[4V] in the reference's order
    (calculate_clusters, monopoles.cpp:1277-1303): int32 [n, 9] rows
    {first site, length, length_mu[4], space, time, sites}."""
    J = np.ascontiguousarray(J, dtype=np.float64)
    V = volume(dims)
    assert J.shape == (4 * V,)
    rec = np.zeros((V, 9), dtype=np.int32)  # roots are distinct sites: n <= V
    n = lib().or_monopole_clusters(_d(J), _dims(dims),
                                   rec.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), V)
    return rec[:n].copy()


def group_percolating(wrappings):
    """The reference's search for a group of wrapped clusters whose winding
    vectors sum to zero (monopoles.cpp:1533-1590), restated with its
    enumeration: groups of 2, 3, ... clusters, the first cluster always in
    the group, the other positions advanced from the last one, an inner
    position advanced only while its successor is more than 2 ahead."""
    w = [list(x) for x in wrappings]
    if len(w) < 2:
        return []
    last = len(w) - 1
    for order in range(2, len(w) + 1):
        pos = list(range(order))
        while True:
            if all(sum(w[i][mu] for i in pos) == 0 for mu in range(4)):
                return pos
            if pos[-1] < last:
                pos[-1] += 1
                continue
            moved = False
            for i in range(order - 2, 0, -1):
                if pos[i + 1] - pos[i] > 2:
                    pos[i] += 1
                    for j in range(i + 1, order):
                        pos[j] = pos[i] + j - i
                    moved = True
                    break
            if not moved:
                break
    return []


def cluster_summary(rec, dims):
    """What apps/monopoles_su2/monopoles_su2.cpp:110-197 writes for one set of
    clusters: unwrapped {length: count}; wrapped clusters as (length,
    windings[4], percolating) sorted by length, longest first (ties in
    cluster order); asymmetry (space/3 - time)/(space/3 + time)."""
    unwrapped = {}
    wrapped = []
    space = int(rec[:, 6].sum()) if len(rec) else 0
    time = int(rec[:, 7].sum()) if len(rec) else 0
    for r in rec:
        lmu = [int(v) for v in r[2:6]]
        if all(v == 0 for v in lmu):
            unwrapped[int(r[1])] = unwrapped.get(int(r[1]), 0) + 1
        else:
            # C++ integer division truncates toward zero
            wrapped.append((int(r[1]), [int(np.fix(lmu[mu] / dims[mu])) for mu in range(4)]))
    wrapped.sort(key=lambda t: -t[0])
    perc = group_percolating([w for _, w in wrapped]) if len(wrapped) > 1 else []
    den = space / 3.0 + time
    asym = (space / 3.0 - time) / den if den != 0 else float("nan")
    return {"unwrapped": dict(sorted(unwrapped.items())),
            "wrapped": [(L, w, i in perc) for i, (L, w) in enumerate(wrapped)],
            "asymmetry": asym}


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def gen_su2_config(dims, eps, seed, copy=0):
    """Seeded synthetic SU(2) links; copy > 0 gives a gauge copy (same link
    draws, another random gauge)."""
    V = volume(dims)
    links = np.zeros((4 * V, 4))
    lib().or_gen_su2_config_copy(_d(links), _dims(dims), float(eps), int(seed), int(copy))
    return links


def gen_copy_spins(dims, seed, copy):
    """Starting spins (V, 3) of gauge copy `copy` in the spin form, on the
    links gen_su2_config(dims, eps, seed) (copy 0)."""
    out = np.zeros((volume(dims), 3))
    lib().or_gen_copy_spins(_d(out), _dims(dims), int(seed), int(copy))
    return out


def gen_su3_config(dims, eps, seed, copy=0):
    V = volume(dims)
    links = np.zeros((4 * V, 18))
    lib().or_gen_su3_config_copy(_d(links), _dims(dims), float(eps), int(seed), int(copy))
    return links


def gen_su3_exp(c8, eps):
    out = np.zeros(18)
    lib().or_gen_su3_exp(_d(np.ascontiguousarray(c8, np.float64)), float(eps), _d(out))
    return out


def gen_spins_uniform(dims, seed):
    V = volume(dims)
    spins = np.zeros((V, 3))
    lib().or_gen_spins_uniform(_d(spins), _dims(dims), int(seed))
    return spins


def set_threads(n: int):
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return lib().or_get_threads()
