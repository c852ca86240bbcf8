"""mag_functional_noalloc (the allocation-free functional of device-resident
links) against the allocating one-shot entry points mag_su2_functional /
mag_su3_functional (MAG_functional_su2, mag.cpp:993-1001;
mag_functional_su3, mag.cpp:1837-1862), called twice in a row on the same
device scratch array, the second call on a smaller lattice.

Device buffers for the inputs and the output come from the HIP runtime the
library itself links (hipMalloc through ctypes): the caller owns them, the
entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

H2D, D2H = 1, 2


def _hip():
    from general_code_amd import mag as G
    G.lib()  # loads libamdhip64.so.7 as its dependency; the CDLL below is that instance
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _links(group, dims, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(dims))
    if group == 2:
        q = rng.normal(size=(4 * V, 4))
        return q / np.linalg.norm(q, axis=1, keepdims=True)
    # SU(3): the one-shot entry point stores rows 0 and 1 and rebuilds row 2
    # as conj(row0 x row1), so the input must be special unitary
    z = rng.normal(size=(4 * V, 3, 3)) + 1j * rng.normal(size=(4 * V, 3, 3))
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    q = q / np.linalg.det(q)[:, None, None] ** (1.0 / 3.0)
    out = np.empty((4 * V, 18))
    out[:, 0::2] = q.real.reshape(4 * V, 9)
    out[:, 1::2] = q.imag.reshape(4 * V, 9)
    return out


def _noalloc(hip, group, dims, links):
    from general_code_amd import mag as G
    a = np.ascontiguousarray(links, dtype=np.float64)
    dl, dF = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dl), a.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dF), 8) == 0
    try:
        nan = np.array([np.nan])
        assert hip.hipMemcpy(dF, nan.ctypes.data, 8, H2D) == 0
        assert hip.hipMemcpy(dl, a.ctypes.data, a.nbytes, H2D) == 0
        G._check(G.lib().mag_functional_noalloc(group, G._dims(dims), dl, dF, None))
        out = np.zeros(1)
        assert hip.hipMemcpy(out.ctypes.data, dF, 8, D2H) == 0  # default stream: orders after
        return float(out[0])
    finally:
        hip.hipFree(dl)
        hip.hipFree(dF)


# (group, first dims, second dims): the second lattice is smaller; 16^4 SU(2)
# fills 32 partial sums, 6^4 one partly filled workgroup
CASES = [
    (2, (16, 16, 16, 16), (6, 6, 6, 6)),
    (3, (8, 8, 8, 16), (4, 4, 4, 4)),
    (2, (8, 8, 8, 8), (2, 4, 2, 4)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("group,first,second", CASES)
def test_noalloc_functional_matches_allocating_entry_point(mag, group, first, second):
    hip = _hip()
    ref = mag.su2_functional if group == 2 else mag.su3_functional
    for k, dims in enumerate((first, second)):
        x = _links(group, dims, 100 * group + k)
        want = ref(dims, x)
        got = _noalloc(hip, group, dims, x)
        assert np.isfinite(got)
        assert abs(got - want) <= 1e-13 * abs(want), (dims, got, want)


@pytest.mark.gpu
def test_noalloc_groups_alternate_on_one_scratch(mag):
    hip = _hip()
    seq = [(3, (8, 8, 8, 8)), (2, (4, 4, 4, 4)), (3, (2, 2, 2, 2))]
    for k, (group, dims) in enumerate(seq):
        x = _links(group, dims, 7 + k)
        want = (mag.su2_functional if group == 2 else mag.su3_functional)(dims, x)
        assert abs(_noalloc(hip, group, dims, x) - want) <= 1e-13 * abs(want)


def test_noalloc_refuses_oversized_and_bad_arguments():
    from general_code_amd import mag
    L = mag.lib()
    buf = ctypes.c_void_p(16)  # never dereferenced: every call below fails its host checks
    big = mag._dims((256, 256, 256, 2))  # 2^25 sites, beyond the 2^24 maximum
    assert L.mag_functional_noalloc(2, big, buf, buf, None) == mag.MAG_ERR_DIMS
    assert b"too large" in L.mag_last_error()
    assert L.mag_functional_noalloc(4, mag._dims((4,) * 4), buf, buf, None) == mag.MAG_ERR_ARG
    assert L.mag_functional_noalloc(2, mag._dims((4, 4, 3, 4)), buf, buf, None) == mag.MAG_ERR_DIMS
    assert L.mag_functional_noalloc(2, mag._dims((4,) * 4), None, buf, None) == mag.MAG_ERR_ARG
