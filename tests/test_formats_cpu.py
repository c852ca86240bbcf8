"""On-disk link formats of the reference (data.h:114-142, indexing.h:190-323)
read by the C++ drop-in header general_code_amd/cpp/mag.h (host only)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def reader(tmp_path_factory, mag):
    exe = str(tmp_path_factory.mktemp("bin") / "read_links")
    pkg = os.path.join(ROOT, "general_code_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(pkg, "cpp"), os.path.join(ROOT, "tests", "cpp", "read_links_main.cpp"),
                    "-L", pkg, "-lmag_gpu", f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    return exe


def _qcdstag_su2(links, dims):
    """Write links in the QCDSTAG su2 order: direction t first, element
    order (a0, a3, a2, a1) (FilePatternQCDSTAG<4, su2>, indexing.h:286-323)."""
    V = int(np.prod(dims))
    L = links.reshape(V, 4, 4)
    out = np.zeros((4, V, 4))
    for mu in range(4):
        mu1 = 0 if mu == 3 else mu + 1
        out[mu1] = L[:, mu, :][:, [0, 3, 2, 1]]
    return out.reshape(-1)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_su2_formats_roundtrip(oracle, reader, tmp_path, precision):
    dims = (4, 2, 4, 6)
    links = oracle.gen_su2_config(dims, 0.5, 3)
    if precision == "float":
        links = links.astype(np.float32).astype(np.float64)
    dt = np.float64 if precision == "double" else np.float32
    lex = tmp_path / "lex.bin"
    links.astype(dt).tofile(lex)
    stag = tmp_path / "stag.bin"
    _qcdstag_su2(links, dims).astype(dt).tofile(stag)
    for path, fmt in [(lex, "lexicographical"), (stag, "qcdstag")]:
        out = tmp_path / f"out_{fmt}.bin"
        subprocess.run([reader, "2", *map(str, dims), str(path), fmt, precision, str(out)], check=True)
        got = np.fromfile(out, dtype=np.float64).reshape(-1, 4)
        assert np.array_equal(got, links)


def test_su3_lexicographical_roundtrip(oracle, reader, tmp_path):
    dims = (2, 2, 2, 4)
    links = oracle.gen_su3_config(dims, 0.4, 5)
    p = tmp_path / "su3.bin"
    links.tofile(p)
    out = tmp_path / "out.bin"
    subprocess.run([reader, "3", *map(str, dims), str(p), "lexicographical", "double", str(out)],
                   check=True)
    assert np.array_equal(np.fromfile(out).reshape(-1, 18), links)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_su3_qcdstag_roundtrip(oracle, reader, tmp_path, precision):
    """The generic QCDSTAG pattern (indexing.h:249-284) for su3: direction
    t first (mu1 = (mu + 1) % 4), the 18 reals of a link in their own order
    (no reordering, unlike the su2 specialisation)."""
    dims = (2, 4, 2, 6)
    V = int(np.prod(dims))
    links = oracle.gen_su3_config(dims, 0.4, 9)
    if precision == "float":
        links = links.astype(np.float32).astype(np.float64)
    dt = np.float64 if precision == "double" else np.float32
    L = links.reshape(V, 4, 18)
    stag = np.zeros((4, V, 18))
    for mu in range(4):
        stag[(mu + 1) % 4] = L[:, mu, :]
    p = tmp_path / "stag3.bin"
    stag.astype(dt).tofile(p)
    out = tmp_path / "out3.bin"
    subprocess.run([reader, "3", *map(str, dims), str(p), "qcdstag", precision, str(out)],
                   check=True)
    assert np.array_equal(np.fromfile(out).reshape(-1, 18), links.reshape(-1, 18))


def test_bad_file_reports_error(reader, tmp_path):
    p = tmp_path / "short.bin"
    np.zeros(10).tofile(p)
    r = subprocess.run([reader, "2", "2", "2", "2", "2", str(p), "lexicographical", "double",
                        str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "short read" in r.stderr


def _be(links):
    return np.ascontiguousarray(links, dtype=">f8").tobytes()


@pytest.mark.parametrize("group", [2, 3])
def test_ildg_read_write(oracle, reader, tmp_path, group):
    """ILDG (LIME) files: the "ildg-binary-data" record holds big-endian
    doubles in the lexicographical order (data.h:144-176, FilePatternILDG
    indexing.h:325-351).  Files written by the reference's own LIME writer
    (c-lime, compiled from the reference sources into oracle/_ref) are read by
    read_links(..., "ildg"), and files from write_links_ildg are read back by
    the reference's LIME reader; without the reference checkout only the
    self round trip runs."""
    dims = (4, 2, 2, 4)
    links = (oracle.gen_su2_config(dims, 0.5, 8) if group == 2
             else oracle.gen_su3_config(dims, 0.4, 8))
    N = links.shape[1]
    have_ref = oracle.build_ref() is not None and oracle.lime_ref() is not None
    lex = tmp_path / "lex.bin"
    links.tofile(lex)
    mine = tmp_path / "mine.lime"
    subprocess.run([reader, str(group), *map(str, dims), str(lex), "lexicographical", "double",
                    str(mine), "ildg"], check=True)
    back = tmp_path / "back.bin"
    subprocess.run([reader, str(group), *map(str, dims), str(mine), "ildg", "double", str(back)],
                   check=True)
    assert np.array_equal(np.fromfile(back).reshape(-1, N), links)
    if not have_ref:
        pytest.skip("reference LIME library not built (no reference checkout)")
    recs = oracle.lime_read(mine)
    assert [r[0] for r in recs] == ["ildg-format", "ildg-binary-data"]
    assert (recs[0][1], recs[-1][2]) == (1, 1)  # message begin / end flags
    assert recs[1][3] == _be(links)
    theirs = tmp_path / "theirs.lime"
    xml = b"<?xml version=\"1.0\"?><ildgFormat><field>gauge</field></ildgFormat>"
    oracle.lime_write(theirs, [("ildg-format", xml), ("ildg-data-lfn", b"lfn://test"),
                               ("ildg-binary-data", _be(links))])
    out = tmp_path / "out.bin"
    subprocess.run([reader, str(group), *map(str, dims), str(theirs), "ildg", "double", str(out)],
                   check=True)
    assert np.array_equal(np.fromfile(out).reshape(-1, N), links)


def test_ildg_wrong_size_is_refused(oracle, reader, tmp_path):
    dims = (4, 2, 2, 4)
    links = oracle.gen_su2_config(dims, 0.5, 8)
    lex = tmp_path / "lex.bin"
    links.tofile(lex)
    f = tmp_path / "f.lime"
    subprocess.run([reader, "2", *map(str, dims), str(lex), "lexicographical", "double", str(f),
                    "ildg"], check=True)
    r = subprocess.run([reader, "2", "4", "2", "2", "2", str(f), "ildg", "double",
                        str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "ildg-binary-data holds" in r.stderr


def test_spin_files_roundtrip_and_short_file(mag, tmp_path):
    """read_spins / write_spins (mag.cpp:54-141): V spins of 3 raw doubles,
    lexicographic; a short file is an error (the reference prints one and
    goes on with garbage)."""
    exe = str(tmp_path / "spins")
    pkg = os.path.join(ROOT, "general_code_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(pkg, "cpp"), os.path.join(ROOT, "tests", "cpp", "spins_main.cpp"),
                    "-L", pkg, "-lmag_gpu", f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    dims = (4, 2, 4, 6)
    V = int(np.prod(dims))
    s = np.random.default_rng(3).normal(size=(V, 3))
    p, out = tmp_path / "s.bin", tmp_path / "o.bin"
    s.tofile(p)
    subprocess.run([exe, *map(str, dims), str(p), str(out)], check=True)
    assert np.array_equal(np.fromfile(out).reshape(V, 3), s)
    s[:-1].tofile(p)
    r = subprocess.run([exe, *map(str, dims), str(p), str(out)], capture_output=True, text=True)
    assert r.returncode == 1 and "read_spins" in r.stderr
