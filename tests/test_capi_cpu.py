"""CPU-side checks of the C ABI library: it loads and exports every entry
point that include/mag_gpu.h declares (no GPU work is started)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mag_\w+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mag_job_create" in names and "mag_job_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(mag):
    lib = ctypes.CDLL(mag.mag.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_binds_every_symbol(mag):
    assert declared_functions() == set(mag.mag.SIGNATURES)


def test_params_default_and_footprint(mag):
    p = mag.MagParams(group=mag.SU3, dims=(24, 24, 24, 48), precision=mag.MIXED)
    b = p.footprint()
    V = 24 ** 3 * 48
    assert b > V * 4 * 12 * 12  # FP64 master + FP32 working links
    c = p.to_c()
    assert c.reunit_every == 10 and abs(c.omega - 1.7) < 1e-15


def test_bad_params_rejected_without_gpu(mag):
    p = mag.MagParams(group=mag.SU2, dims=(7, 8, 8, 8))
    try:
        p.footprint()
    except mag.MagError as e:
        assert e.code == -2
    else:
        raise AssertionError("odd extent accepted")
