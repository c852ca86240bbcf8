"""Python front end of libmag_gpu.so (ctypes over the C ABI in include/mag_gpu.h).

The HIP library is the only implementation: importing this module on a
machine without the built library, or calling into it without a GPU, raises.

Arrays follow the reference layout (ilyakudrov/general_code
lib/cpu/include/indexing.h:71-126): SU(2) links float64 [4V, 4]
(a0 + i a.sigma, matrix.h:8-13), SU(3) links float64 [4V, 18] (row-major
re/im, matrix.h:236-254), link index 4*site + mu with
site = x + Lx*(y + Ly*(z + Lz*t)).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmag_gpu.so")

MAG_OK = 0
MAG_ERR_ARG = -1
MAG_ERR_DIMS = -2
MAG_ERR_NOMEM = -3
MAG_ERR_NONFINITE = -4
MAG_ERR_NOT_CONVERGED = -5
MAG_ERR_HIP = -6
MAG_ERR_STATE = -7
MAG_ERR_CANCELLED = -8

SU2, SU3 = 2, 3
F64, MIXED = 0, 1
FORM_LINK, FORM_SPIN = 0, 1
MAX_COPIES = 8
TUNE_AUTO, TUNE_FIXED, TUNE_DEFAULT = 0, 1, 2
SHAPES = [(64, 1), (64, 2), (128, 1), (128, 2), (256, 1), (256, 2), (512, 1), (512, 2)]


class MagError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NotConverged(MagError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [
        ("group", ctypes.c_int),
        ("dims", ctypes.c_int * 4),
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
        ("omega", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("max_sweeps", ctypes.c_int),
        ("reunit_every", ctypes.c_int),
        ("tol_mixed", ctypes.c_double),
        ("form", ctypes.c_int),
        ("copies", ctypes.c_int),
        ("fp64_warmup", ctypes.c_int),
        ("track_gauge", ctypes.c_int),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class _Request(ctypes.Structure):
    _fields_ = [
        ("params", _Params),
        ("links", ctypes.POINTER(ctypes.c_double)),
        ("eps", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("copy", ctypes.c_uint32),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("status", ctypes.c_int),
        ("sweeps", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("best_copy", ctypes.c_int),
        ("fp64_start", ctypes.c_int),
        ("F", ctypes.c_double),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_done", ctypes.c_double),
    ]


class _ServerStats(ctypes.Structure):
    _fields_ = [
        ("budget", ctypes.c_uint64),
        ("peak_hbm", ctypes.c_uint64),
        ("submitted", ctypes.c_uint64),
        ("finished", ctypes.c_uint64),
        ("admitted", ctypes.c_uint64),
        ("replays", ctypes.c_uint64),
        ("peak_active", ctypes.c_int),
    ]

# name -> (restype, argtypes); every entry point of include/mag_gpu.h
SIGNATURES = {
    "mag_params_default": (None, [ctypes.POINTER(_Params)]),
    "mag_last_error": (ctypes.c_char_p, []),
    "mag_device_count": (ctypes.c_int, [_ip]),
    "mag_job_create": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.POINTER(_vp)]),
    "mag_job_destroy": (ctypes.c_int, [_vp]),
    "mag_job_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mag_job_footprint": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_uint64)]),
    "mag_job_upload_links": (ctypes.c_int, [_vp, _dp]),
    "mag_job_upload_links_device": (ctypes.c_int, [_vp, _vp]),
    "mag_job_generate": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_uint64]),
    "mag_job_generate_copy": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_uint64,
                                             ctypes.c_uint32]),
    "mag_job_sweeps": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mag_job_run": (ctypes.c_int, [_vp]),
    "mag_job_finalize": (ctypes.c_int, [_vp]),
    "mag_job_status": (ctypes.c_int, [_vp, _ip, _ip, _dp]),
    "mag_job_initial_functional": (ctypes.c_int, [_vp, _dp]),
    "mag_job_history": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, _ip]),
    "mag_job_fp64_start": (ctypes.c_int, [_vp, _ip]),
    "mag_job_set_copy_spins": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_job_get_copy_spins": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_job_copy_status": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _dp]),
    "mag_job_copy_history": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip]),
    "mag_job_best_copy": (ctypes.c_int, [_vp, _ip, _dp]),
    "mag_job_download_copy_links": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_tune_set": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mag_tune_clear": (ctypes.c_int, []),
    "mag_job_shape": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "mag_job_download_links": (ctypes.c_int, [_vp, _dp]),
    "mag_job_download_links_canonical": (ctypes.c_int, [_vp, _dp]),
    "mag_job_download_spins": (ctypes.c_int, [_vp, _dp]),
    "mag_job_functional": (ctypes.c_int, [_vp, _dp]),
    "mag_job_plaquette": (ctypes.c_int, [_vp, _dp]),
    "mag_job_sweeps_timed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_job_halfsweep_timed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_su2_functional": (ctypes.c_int, [_ip, _dp, _dp]),
    "mag_su3_functional": (ctypes.c_int, [_ip, _dp, _dp]),
    "mag_plaquette": (ctypes.c_int, [ctypes.c_int, _ip, _dp, _dp]),
    "mag_functional_noalloc": (ctypes.c_int, [ctypes.c_int, _ip, _vp, _vp, _vp]),
    "mag_hbm_copy_bandwidth": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _dp]),
    # batches
    "mag_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "mag_batch_destroy": (ctypes.c_int, [_vp]),
    "mag_batch_add": (ctypes.c_int, [_vp, _vp]),
    "mag_batch_remove": (ctypes.c_int, [_vp, _vp]),
    "mag_batch_size": (ctypes.c_int, [_vp, _ip]),
    "mag_batch_replay_sweeps": (ctypes.c_int, [_vp, _ip]),
    "mag_batch_replays": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mag_batch_poll": (ctypes.c_int, [_vp, _ip, _ip]),
    "mag_job_abelian_angles": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_job_monopole_currents": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_job_monopole_timed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_u1_monopole_currents": (ctypes.c_int, [_ip, _dp, _dp]),
    "mag_su3_monopole_currents": (ctypes.c_int, [_ip, _dp, _dp]),
    "mag_job_monopole_clusters": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp]),
    "mag_current_clusters": (ctypes.c_int, [_ip, ctypes.c_int, _dp, _vp, ctypes.c_int64, _vp]),
    "mag_job_clusters_timed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_batch_step": (ctypes.c_int, [_vp, _ip, _ip]),
    "mag_batch_run": (ctypes.c_int, [_vp]),
    "mag_batch_replays_timed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mag_batch_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mag_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.POINTER(_vp)]),
    "mag_server_destroy": (ctypes.c_int, [_vp]),
    "mag_server_submit": (ctypes.c_int, [_vp, ctypes.POINTER(_Request),
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "mag_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_Result)]),
    "mag_server_drain": (ctypes.c_int, [_vp]),
    "mag_server_cancel": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "mag_server_clock": (ctypes.c_int, [_vp, _dp]),
    "mag_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(_ServerStats)]),
    # spin form (reference mag.h surface)
    "mag_spin_session_create": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mag_spin_session_destroy": (ctypes.c_int, [_vp]),
    "mag_spin_session_upload_links": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_download_links": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_set_spins": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_get_spins": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_generate_spins": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "mag_spin_session_functional_spin": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_functional": (ctypes.c_int, [_vp, _dp]),
    "mag_spin_session_sweep": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_uint64, _dp, _dp]),
    "mag_spin_session_simulated_annealing": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                                            ctypes.c_double, ctypes.c_int,
                                                            ctypes.c_int, ctypes.c_uint64]),
    "mag_spin_session_maximize_final": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double,
                                                       ctypes.c_double, ctypes.c_int, _ip]),
    "mag_spin_session_maximize_approximate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                                             ctypes.c_int, _ip]),
    "mag_spin_session_gauge_transform": (ctypes.c_int, [_vp]),
    "mag_su2_generate_spins_uniform": (ctypes.c_int, [_ip, ctypes.c_uint64, _dp]),
    "mag_su2_functional_spin": (ctypes.c_int, [_ip, _dp, _dp, _dp]),
    "mag_su2_gauge_transform_spins": (ctypes.c_int, [_ip, _dp, _dp]),
    "mag_su2_maximize_final": (ctypes.c_int, [_ip, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, _ip]),
    "mag_su2_maximize_approximate": (ctypes.c_int, [_ip, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, _ip]),
    "mag_su2_simulated_annealing": (ctypes.c_int, [_ip, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_uint64]),
}

SPIN_OR, SPIN_RELAX, SPIN_OMEGA, SPIN_HEATBATH = 0, 1, 2, 3

_lib = None


def lib():
    """Load libmag_gpu.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m general_code_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != MAG_OK:
        msg = lib().mag_last_error().decode(errors="replace")
        if rc == MAG_ERR_NOT_CONVERGED:
            raise NotConverged(rc, msg)
        raise MagError(rc, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _dims(d):
    return (ctypes.c_int * 4)(*[int(x) for x in d])


def _volume(dims) -> int:
    d = [int(x) for x in dims]
    if len(d) != 4:
        raise ValueError(f"dims must hold 4 extents, got {len(d)}")
    return int(np.prod(d))


def _sized(a, dims, per_site: int, name: str) -> np.ndarray:
    """a as contiguous float64, checked to hold per_site reals per lattice
    site of dims (the C side reads and writes that many)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    V = _volume(dims)
    if a.size != per_site * V:
        raise ValueError(f"{name} holds {a.size} reals, expected {per_site} x V = {per_site * V}")
    return a


def hbm_copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20, device: int = 0) -> float:
    """GB/s (read + write) of a plain 16-byte streaming copy kernel."""
    g = ctypes.c_double()
    _check(lib().mag_hbm_copy_bandwidth(int(device), int(nbytes), int(iters), ctypes.byref(g)))
    return g.value


def device_count() -> int:
    n = ctypes.c_int()
    _check(lib().mag_device_count(ctypes.byref(n)))
    return n.value


@dataclass
class MagParams:
    """Job parameters (struct mag_params of include/mag_gpu.h)."""
    group: int = SU2
    dims: tuple = (8, 8, 8, 8)
    precision: int = F64
    device: int = 0
    omega: float = 1.7
    tol: float = 1e-12
    max_sweeps: int = 1000
    reunit_every: int = 10
    tol_mixed: float = 1e-9
    form: int = FORM_LINK
    copies: int = 1
    fp64_warmup: int = 0
    track_gauge: int = 0

    def to_c(self) -> _Params:
        p = _Params()
        lib().mag_params_default(ctypes.byref(p))
        p.group = int(self.group)
        for i in range(4):
            p.dims[i] = int(self.dims[i])
        p.precision = int(self.precision)
        p.device = int(self.device)
        p.omega = float(self.omega)
        p.tol = float(self.tol)
        p.max_sweeps = int(self.max_sweeps)
        p.reunit_every = int(self.reunit_every)
        p.tol_mixed = float(self.tol_mixed)
        p.form = int(self.form)
        p.copies = int(self.copies)
        p.fp64_warmup = int(self.fp64_warmup)
        p.track_gauge = int(self.track_gauge)
        return p

    @property
    def volume(self) -> int:
        return int(np.prod([int(d) for d in self.dims]))

    @property
    def reals_in(self) -> int:
        return 4 if self.group == SU2 else 18

    def footprint(self) -> int:
        b = ctypes.c_uint64()
        p = self.to_c()
        _check(lib().mag_job_footprint(ctypes.byref(p), ctypes.byref(b)))
        return b.value


class MagJob:
    """One configuration resident in HBM, gauge-fixed by checkerboard
    omega-overrelaxation on the GPU."""

    def __init__(self, params: MagParams | None = None, **kw):
        self.params = params if params is not None else MagParams(**kw)
        self._p = self.params.to_c()
        h = _vp()
        _check(lib().mag_job_create(ctypes.byref(self._p), ctypes.byref(h)))
        self._h = h

    # lifecycle ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().mag_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # data -----------------------------------------------------------------
    def upload(self, links: np.ndarray):
        a = np.ascontiguousarray(links, dtype=np.float64)
        V = self.params.volume
        if a.size != 4 * V * self.params.reals_in:
            raise ValueError(f"links must hold {4 * V * self.params.reals_in} reals, got {a.size}")
        _check(lib().mag_job_upload_links(self._h, _dptr(a)))

    def generate(self, eps: float, seed: int, copy: int = 0):
        """Seeded synthetic configuration on the GPU; copy > 0 is a gauge copy
        of it (same link draws, another random gauge)."""
        _check(lib().mag_job_generate_copy(self._h, float(eps), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                           int(copy)))

    def download(self) -> np.ndarray:
        V = self.params.volume
        out = np.empty((4 * V, self.params.reals_in), dtype=np.float64)
        _check(lib().mag_job_download_links(self._h, _dptr(out)))
        return out

    def download_canonical(self) -> np.ndarray:
        """SU(2) links in the reference's spin-form convention (track_gauge)."""
        out = np.empty((4 * self.params.volume, 4), dtype=np.float64)
        _check(lib().mag_job_download_links_canonical(self._h, _dptr(out)))
        return out

    def download_spins(self) -> np.ndarray:
        """Spins s(x) of the tracked gauge, [V, 3] (track_gauge)."""
        out = np.empty((self.params.volume, 3), dtype=np.float64)
        _check(lib().mag_job_download_spins(self._h, _dptr(out)))
        return out

    # work -----------------------------------------------------------------
    def sweeps(self, n: int):
        _check(lib().mag_job_sweeps(self._h, int(n)))

    def run(self, raise_on_cap: bool = False) -> bool:
        rc = lib().mag_job_run(self._h)
        if rc == MAG_ERR_NOT_CONVERGED and not raise_on_cap:
            return False
        _check(rc)
        return True

    def finalize(self):
        _check(lib().mag_job_finalize(self._h))

    def sweeps_timed(self, n: int) -> float:
        """ms for n sweeps (HIP events on the job's stream)."""
        tot = ctypes.c_double()
        _check(lib().mag_job_sweeps_timed(self._h, int(n), ctypes.byref(tot)))
        return tot.value

    def halfsweep_timed(self, n: int) -> float:
        """mean ms of one half-sweep kernel over n back-to-back launches."""
        avg = ctypes.c_double()
        _check(lib().mag_job_halfsweep_timed(self._h, int(n), ctypes.byref(avg)))
        return avg.value

    # results --------------------------------------------------------------
    def status(self):
        s = ctypes.c_int()
        c = ctypes.c_int()
        F = ctypes.c_double()
        _check(lib().mag_job_status(self._h, ctypes.byref(s), ctypes.byref(c), ctypes.byref(F)))
        return s.value, bool(c.value), F.value

    def initial_functional(self) -> float:
        F = ctypes.c_double()
        _check(lib().mag_job_initial_functional(self._h, ctypes.byref(F)))
        return F.value

    def history(self):
        n_max = self.params.max_sweeps
        F = np.zeros(n_max)
        R = np.zeros(n_max)
        A = np.zeros(n_max)
        n = ctypes.c_int()
        _check(lib().mag_job_history(self._h, _dptr(F), _dptr(R), _dptr(A), n_max,
                                     ctypes.byref(n)))
        m = n.value
        return F[:m].copy(), R[:m].copy(), A[:m].copy()

    def fp64_start(self) -> int:
        s = ctypes.c_int()
        _check(lib().mag_job_fp64_start(self._h, ctypes.byref(s)))
        return s.value

    def shape(self, fp32: bool = False):
        """(threads per workgroup, tiles per thread) of the job's FP64 or FP32
        link-form half-sweep; (0, 0) before its first sweep."""
        nt = ctypes.c_int()
        spt = ctypes.c_int()
        _check(lib().mag_job_shape(self._h, int(bool(fp32)), ctypes.byref(nt), ctypes.byref(spt)))
        return nt.value, spt.value

    # gauge copies (spin form) ---------------------------------------------
    def set_copy_spins(self, copy: int, spins: np.ndarray):
        a = np.ascontiguousarray(spins, dtype=np.float64)
        if a.size != 3 * self.params.volume:
            raise ValueError("spins must hold 3*V reals")
        _check(lib().mag_job_set_copy_spins(self._h, int(copy), _dptr(a)))

    def copy_spins(self, copy: int) -> np.ndarray:
        out = np.empty((self.params.volume, 3))
        _check(lib().mag_job_get_copy_spins(self._h, int(copy), _dptr(out)))
        return out

    def copy_status(self, copy: int):
        s = ctypes.c_int()
        c = ctypes.c_int()
        F = ctypes.c_double()
        _check(lib().mag_job_copy_status(self._h, int(copy), ctypes.byref(s), ctypes.byref(c),
                                         ctypes.byref(F)))
        return s.value, bool(c.value), F.value

    def copy_history(self, copy: int):
        n_max = self.params.max_sweeps
        F = np.zeros(n_max)
        R = np.zeros(n_max)
        A = np.zeros(n_max)
        n = ctypes.c_int()
        _check(lib().mag_job_copy_history(self._h, int(copy), _dptr(F), _dptr(R), _dptr(A),
                                          n_max, ctypes.byref(n)))
        m = n.value
        return F[:m].copy(), R[:m].copy(), A[:m].copy()

    def best_copy(self):
        c = ctypes.c_int()
        F = ctypes.c_double()
        _check(lib().mag_job_best_copy(self._h, ctypes.byref(c), ctypes.byref(F)))
        return c.value, F.value

    def download_copy(self, copy: int) -> np.ndarray:
        out = np.empty((4 * self.params.volume, self.params.reals_in), dtype=np.float64)
        _check(lib().mag_job_download_copy_links(self._h, int(copy), _dptr(out)))
        return out

    def functional(self) -> float:
        F = ctypes.c_double()
        _check(lib().mag_job_functional(self._h, ctypes.byref(F)))
        return F.value

    def abelian_angles(self, copy: int = 0) -> np.ndarray:
        """Abelian projection of the links (mag_job_abelian_angles): SU(2)
        [4V] angles, SU(3) [4V, 3] su3_angles, reference link order."""
        V = self.params.volume
        out = np.empty(4 * V if self.params.group == SU2 else (4 * V, 3), dtype=np.float64)
        _check(lib().mag_job_abelian_angles(self._h, int(copy), _dptr(out)))
        return out

    def monopole_currents(self, copy: int = 0) -> np.ndarray:
        """Monopole currents of the links (mag_job_monopole_currents): SU(2)
        [4V], SU(3) [3, 4V]."""
        V = self.params.volume
        out = np.empty(4 * V if self.params.group == SU2 else (3, 4 * V), dtype=np.float64)
        _check(lib().mag_job_monopole_currents(self._h, int(copy), _dptr(out)))
        return out

    def monopole_clusters(self, copy: int = 0):
        """Clusters of the copy's monopole currents, computed on the device
        (mag_job_monopole_clusters): int32 [n, 9] records {first site,
        length, length_mu[4], space, time, sites} in the reference's order;
        SU(3): a list of the three colours' records."""
        ncol = 1 if self.params.group == SU2 else 3
        recs = _clusters(lambda buf, cap, n: lib().mag_job_monopole_clusters(self._h, int(copy), buf, cap, n),
                         ncol, self.params.volume)
        return recs[0] if ncol == 1 else recs

    def clusters_timed(self, reps: int) -> float:
        """Mean device ms of the cluster kernels (copy 0, all colours)."""
        ms = ctypes.c_double()
        _check(lib().mag_job_clusters_timed(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def monopole_timed(self, reps: int) -> float:
        """Mean device ms of angles + plaquettes + currents (copy 0)."""
        ms = ctypes.c_double()
        _check(lib().mag_job_monopole_timed(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def plaquette(self) -> float:
        P = ctypes.c_double()
        _check(lib().mag_job_plaquette(self._h, ctypes.byref(P)))
        return P.value


# one-shot helpers ---------------------------------------------------------
def su2_functional(dims, links) -> float:
    """MAG_functional_su2 (lib/cpu/src/mag.cpp:993-1001) on the GPU."""
    a = _sized(links, dims, 16, "links")
    F = ctypes.c_double()
    _check(lib().mag_su2_functional(_dims(dims), _dptr(a), ctypes.byref(F)))
    return F.value


def su3_functional(dims, links) -> float:
    """mag_functional_su3 (lib/cpu/src/mag.cpp:1837-1862) on the GPU."""
    a = _sized(links, dims, 72, "links")
    F = ctypes.c_double()
    _check(lib().mag_su3_functional(_dims(dims), _dptr(a), ctypes.byref(F)))
    return F.value


def plaquette(group, dims, links) -> float:
    """plaket (lib/cpu/include/plaket.h:84-101) on the GPU."""
    a = _sized(links, dims, 16 if int(group) == SU2 else 72, "links")
    P = ctypes.c_double()
    _check(lib().mag_plaquette(int(group), _dims(dims), _dptr(a), ctypes.byref(P)))
    return P.value


# spin form -----------------------------------------------------------------
class SpinSession:
    """SU(2) links and spins resident on the GPU: the spin-form MAG of the
    reference's lib/cpu/include/mag.h (checkerboard sweep order)."""

    def __init__(self, dims, device: int = 0):
        self.dims = tuple(int(d) for d in dims)
        self.volume = int(np.prod(self.dims))
        h = _vp()
        _check(lib().mag_spin_session_create(_dims(self.dims), int(device), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mag_spin_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_links(self, links):
        a = np.ascontiguousarray(links, dtype=np.float64)
        if a.size != 16 * self.volume:
            raise ValueError("links must hold 16*V reals")
        _check(lib().mag_spin_session_upload_links(self._h, _dptr(a)))

    def download_links(self):
        out = np.empty((4 * self.volume, 4))
        _check(lib().mag_spin_session_download_links(self._h, _dptr(out)))
        return out

    def set_spins(self, spins):
        a = np.ascontiguousarray(spins, dtype=np.float64)
        if a.size != 3 * self.volume:
            raise ValueError("spins must hold 3*V reals")
        _check(lib().mag_spin_session_set_spins(self._h, _dptr(a)))

    def get_spins(self):
        out = np.empty((self.volume, 3))
        _check(lib().mag_spin_session_get_spins(self._h, _dptr(out)))
        return out

    def generate_spins(self, seed: int):
        _check(lib().mag_spin_session_generate_spins(self._h, int(seed)))

    def functional_spin(self) -> float:
        F = ctypes.c_double()
        _check(lib().mag_spin_session_functional_spin(self._h, ctypes.byref(F)))
        return F.value

    def functional(self) -> float:
        F = ctypes.c_double()
        _check(lib().mag_spin_session_functional(self._h, ctypes.byref(F)))
        return F.value

    def sweep(self, mode: int, param: float = 0.0, seed: int = 0):
        mx = ctypes.c_double()
        av = ctypes.c_double()
        _check(lib().mag_spin_session_sweep(self._h, int(mode), float(param), int(seed),
                                            ctypes.byref(mx), ctypes.byref(av)))
        return mx.value, av.value

    def simulated_annealing(self, T_init, T_final, T_step, OR_steps, thermalization_steps, seed):
        _check(lib().mag_spin_session_simulated_annealing(
            self._h, float(T_init), float(T_final), float(T_step), int(OR_steps),
            int(thermalization_steps), int(seed)))

    def maximize_final(self, OR_steps, tol_max, tol_avg, max_iter=100000) -> int:
        it = ctypes.c_int()
        _check(lib().mag_spin_session_maximize_final(self._h, int(OR_steps), float(tol_max),
                                                     float(tol_avg), int(max_iter),
                                                     ctypes.byref(it)))
        return it.value

    def maximize_approximate(self, OR_steps, tolerance_digits, max_iter=100000) -> int:
        it = ctypes.c_int()
        _check(lib().mag_spin_session_maximize_approximate(self._h, int(OR_steps),
                                                           int(tolerance_digits), int(max_iter),
                                                           ctypes.byref(it)))
        return it.value

    def gauge_transform(self):
        _check(lib().mag_spin_session_gauge_transform(self._h))


def generate_spins_uniform(dims, seed: int) -> np.ndarray:
    """generate_spins_uniform (lib/cpu/include/mag.h:179-199) on the GPU."""
    V = int(np.prod(dims))
    out = np.empty((V, 3))
    _check(lib().mag_su2_generate_spins_uniform(_dims(dims), int(seed), _dptr(out)))
    return out


def functional_spin(dims, links, spins) -> float:
    """MAG_functional_su2_spin (lib/cpu/src/mag.cpp:925-953) on the GPU."""
    F = ctypes.c_double()
    a = _sized(links, dims, 16, "links")
    s = _sized(spins, dims, 3, "spins")
    _check(lib().mag_su2_functional_spin(_dims(dims), _dptr(a), _dptr(s), ctypes.byref(F)))
    return F.value


def gauge_transform_spins(dims, links, spins) -> np.ndarray:
    """gauge_tranformation_spins (lib/cpu/src/mag.cpp:1045-1068); returns the
    transformed links."""
    out = _sized(links, dims, 16, "links").reshape(4 * _volume(dims), 4).copy()
    s = _sized(spins, dims, 3, "spins")
    _check(lib().mag_su2_gauge_transform_spins(_dims(dims), _dptr(out), _dptr(s)))
    return out


# monopoles -----------------------------------------------------------------
def u1_monopole_currents(dims, angles) -> np.ndarray:
    """calculate_current of an abelian configuration (monopoles.cpp:669-687):
    angles [4V] in reference link order -> J [4V]."""
    a = _sized(angles, dims, 4, "angles").reshape(-1)
    out = np.empty(4 * _volume(dims), dtype=np.float64)
    _check(lib().mag_u1_monopole_currents(_dims(dims), _dptr(a), _dptr(out)))
    return out


def su3_monopole_currents(dims, angles3) -> np.ndarray:
    """calculate_current of an su3_angles configuration (monopoles.cpp:646-667):
    angles [4V, 3] -> J [3, 4V]."""
    a = _sized(angles3, dims, 12, "angles3")
    out = np.empty((3, 4 * _volume(dims)), dtype=np.float64)
    _check(lib().mag_su3_monopole_currents(_dims(dims), _dptr(a), _dptr(out)))
    return out


CLUSTER_INTS = 9


def _clusters(call, ncol, V):
    """Run a cluster entry point with a first guess of the capacity and once
    more with the exact count if that was short."""
    cap = min(V, 1 << 16)
    for _ in range(2):
        buf = np.zeros((ncol, cap, CLUSTER_INTS), dtype=np.int32)
        n = np.zeros(ncol, dtype=np.int64)
        rc = call(buf.ctypes.data, cap, n.ctypes.data)
        if rc == MAG_OK:
            return [buf[k, :n[k]].copy() for k in range(ncol)]
        if rc == MAG_ERR_ARG and n.max() > cap:
            cap = int(n.max())
            continue
        _check(rc)
    raise MagError(MAG_ERR_ARG, "cluster count changed between calls")


def current_clusters(dims, J):
    """calculate_clusters of host currents (mag_current_clusters): J [4V]
    (one colour) or [3, 4V]; records as MagJob.monopole_clusters."""
    J = np.ascontiguousarray(J, dtype=np.float64)
    V = int(np.prod(dims))
    ncol = 1 if J.ndim == 1 else J.shape[0]
    if J.size != ncol * 4 * V:
        raise ValueError(f"J has {J.size} values, expected {ncol} x 4V = {ncol * 4 * V}")
    recs = _clusters(lambda buf, cap, n: lib().mag_current_clusters(_dims(dims), ncol, _dptr(J), buf, cap, n),
                     ncol, V)
    return recs[0] if J.ndim == 1 else recs


# batches -------------------------------------------------------------------
class MagBatch:
    """Many jobs swept together: batched launches per (group, storage) class,
    each class's replay captured in a hipGraph and run as an independent
    chain on its own stream (include/mag_gpu.h, mag_batch_*)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().mag_batch_create(int(device), ctypes.byref(h)))
        self._h = h
        self._jobs = []

    def close(self):
        if getattr(self, "_h", None):
            lib().mag_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, job: "MagJob"):
        _check(lib().mag_batch_add(self._h, job._h))
        self._jobs.append(job)

    def remove(self, job: "MagJob"):
        _check(lib().mag_batch_remove(self._h, job._h))
        self._jobs.remove(job)

    @property
    def replay_sweeps(self) -> int:
        n = ctypes.c_int()
        _check(lib().mag_batch_replay_sweeps(self._h, ctypes.byref(n)))
        return n.value

    def replays(self, n: int = 1):
        _check(lib().mag_batch_replays(self._h, int(n)))

    def poll(self):
        r = ctypes.c_int()
        f = ctypes.c_int()
        _check(lib().mag_batch_poll(self._h, ctypes.byref(r), ctypes.byref(f)))
        return r.value, f.value

    def step(self):
        """One scheduling step (mag_batch_step): returns (running, finished)."""
        r = ctypes.c_int()
        f = ctypes.c_int()
        _check(lib().mag_batch_step(self._h, ctypes.byref(r), ctypes.byref(f)))
        return r.value, f.value

    def run(self):
        _check(lib().mag_batch_run(self._h))

    def replays_timed(self, n: int) -> float:
        ms = ctypes.c_double()
        _check(lib().mag_batch_replays_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value


# serving -------------------------------------------------------------------
class MagServer:
    """Continuous-batching server (mag_server_* of include/mag_gpu.h): a
    native scheduler thread admits requests by HBM footprint and sweeps every
    admitted job in one graph-captured batch until it converges."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_active: int = 0):
        h = _vp()
        _check(lib().mag_server_create(int(device), int(hbm_budget), int(max_active),
                                       ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().mag_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, params: MagParams | None = None, eps: float = 0.0, seed: int = 0,
               copy: int = 0, links: np.ndarray | None = None, **kw) -> int:
        p = params if params is not None else MagParams(**kw)
        r = _Request()
        r.params = p.to_c()
        r.eps = float(eps)
        r.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        r.copy = int(copy)
        if links is not None:
            a = np.ascontiguousarray(links, dtype=np.float64)
            if a.size != 4 * p.volume * p.reals_in:
                raise ValueError(f"links must hold {4 * p.volume * p.reals_in} reals")
            r.links = _dptr(a)
        rid = ctypes.c_uint64()
        _check(lib().mag_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int) -> dict:
        r = _Result()
        _check(lib().mag_server_wait(self._h, int(rid), ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in _Result._fields_}

    def drain(self):
        _check(lib().mag_server_drain(self._h))

    def cancel(self, rid: int):
        """Cancel a queued or running request (its result: MAG_ERR_CANCELLED)."""
        _check(lib().mag_server_cancel(self._h, int(rid)))

    def clock(self) -> float:
        t = ctypes.c_double()
        _check(lib().mag_server_clock(self._h, ctypes.byref(t)))
        return t.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().mag_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}


# launch shapes --------------------------------------------------------------
def tune_set(mode: int, nt: int = 256, spt: int = 1):
    """TUNE_AUTO (time every shape on first use), TUNE_FIXED (nt, spt) or
    TUNE_DEFAULT (256, 1); every shape gives bit-identical results."""
    _check(lib().mag_tune_set(int(mode), int(nt), int(spt)))


def tune_clear():
    _check(lib().mag_tune_clear())
