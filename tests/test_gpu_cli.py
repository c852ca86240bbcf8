"""The mag_fixation app (general_code_amd/cpp/mag_fixation.cpp) end to end on
the GPU, with the flags of apps/mag_su2/mag_fixation.cpp:60-117."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "general_code_amd", "mag_fixation")


def _run(args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _num(out, key):
    return [float(m) for m in re.findall(rf"{key}\s+([-+0-9.eE]+)", out)]


def test_spin_path_from_file(oracle, tmp_path):
    dims = (8, 8, 8, 8)
    links = oracle.gen_su2_config(dims, 0.3, 21)
    conf = tmp_path / "conf.bin"
    links.tofile(conf)
    outp = tmp_path / "fixed.bin"
    out = _run(["--path_conf", str(conf), "--conf_format", "lexicographical",
                "--file_precision", "double", "--x_size", "8", "--y_size", "8", "--z_size", "8",
                "--t_size", "8", "--T_init", "2.0", "--T_final", "0.5", "--T_step", "0.25",
                "--OR_steps", "3", "--thermalization_steps", "3", "--tolerance_maximal", "1e-9",
                "--tolerance_average", "1e-13", "--is_final", "1", "--seed", "4",
                "--path_conf_output", str(outp)])
    plaq = _num(out, "plaquette")
    assert abs(plaq[0] - oracle.su2_plaquette(links, dims)) < 1e-13
    assert abs(plaq[1] - plaq[0]) < 1e-12
    fixed = np.fromfile(outp).reshape(-1, 4)
    F = _num(out, "final functional is")[0]
    assert abs(oracle.su2_functional(fixed, dims) - F) < 1e-13
    assert abs(_num(out, "spin functional")[0] - F) < 1e-12
    assert F > oracle.su2_functional(links, dims) + 0.3


def test_link_form_synthetic(oracle, tmp_path):
    outp = tmp_path / "fixed.bin"
    out = _run(["--synthetic_eps", "0.3", "--seed", "42", "--x_size", "8", "--y_size", "8",
                "--z_size", "8", "--t_size", "8", "--link_form", "1", "--omega", "1.7",
                "--tol", "1e-12", "--max_sweeps", "1000", "--path_conf_output", str(outp)])
    dims = (8, 8, 8, 8)
    ref = oracle.gen_su2_config(dims, 0.3, 42)
    F_prev = oracle.su2_functional(ref, dims)
    for n in range(1000):
        F, _, _ = oracle.su2_link_sweep(ref, dims, 1.7)
        if abs(F - F_prev) < 1e-12 * abs(F):
            break
        F_prev = F
    assert abs(_num(out, "link-form sweeps")[0] - (n + 1)) <= 1
    fixed = np.fromfile(outp).reshape(-1, 4)
    assert abs(oracle.su2_functional(fixed, dims) - F) < 1e-12


@pytest.mark.parametrize("link_form", [1, 0])
def test_monopole_output(oracle, tmp_path, link_form):
    """--path_monopoles_output: link form from the device-resident fixed links
    (mag_job_monopole_currents), spin path from the host links through
    calculate_current; both against the oracle on the written links."""
    outp, mono = tmp_path / "fixed.bin", tmp_path / "J.bin"
    tabs = {k: tmp_path / f"{k}.csv" for k in ("unwrapped", "wrapped", "windings", "monopoles")}
    args = ["--synthetic_eps", "1.5", "--seed", "7", "--x_size", "8", "--y_size", "8",
            "--z_size", "8", "--t_size", "8", "--path_conf_output", str(outp),
            "--path_monopoles_output", str(mono),
            "--path_output_clusters_unwrapped", str(tabs["unwrapped"]),
            "--path_output_clusters_wrapped", str(tabs["wrapped"]),
            "--path_output_windings", str(tabs["windings"]),
            "--path_output_monopoles", str(tabs["monopoles"])]
    if link_form:
        args += ["--link_form", "1", "--omega", "1.7", "--tol", "1e-10", "--max_sweeps", "20000"]
    else:
        args += ["--T_init", "2.0", "--T_final", "0.5", "--T_step", "0.5", "--OR_steps", "2",
                 "--thermalization_steps", "2", "--tolerance_maximal", "1e-6",
                 "--tolerance_average", "1e-9", "--max_iter", "200"]
    out = _run(args)
    dims = (8, 8, 8, 8)
    fixed = np.fromfile(outp).reshape(-1, 4)
    J = np.fromfile(mono)
    ang = oracle.su2_abelian_angles(np.ascontiguousarray(fixed), dims)
    Jr = oracle.monopole_currents(oracle.u1_monopole_plaquettes(ang, dims), dims)
    assert np.abs(J - Jr).max() < 1e-12
    n = int(np.abs(np.round(J)).sum())
    assert n == _num(out, "monopole current links")[0]
    # the cluster tables, in the CSV formats of apps/monopoles_su2/monopoles_su2.cpp:97-197
    rec = oracle.monopole_clusters(J, dims)
    assert len(rec) == _num(out, "monopole clusters")[0]
    summ = oracle.cluster_summary(rec, dims)
    lines = tabs["unwrapped"].read_text().splitlines()
    assert lines[0] == "length,number"
    assert {int(a): int(b) for a, b in (l.split(",") for l in lines[1:])} == summ["unwrapped"]
    lines = tabs["wrapped"].read_text().splitlines()
    assert lines[0] == "length,x0_wrap,x1_wrap,x2_wrap,x3_wrap,percolating_group"
    got = [(int(f[0]), [int(v) for v in f[1:5]], f[5] == "percolating") for f in (l.split(",") for l in lines[1:])]
    assert got == summ["wrapped"]
    assert tabs["windings"].read_text().splitlines() == ["winding_number,cluster_number,direction"]
    lines = tabs["monopoles"].read_text().splitlines()
    assert lines[0] == "asymmetry"
    assert float(lines[1]) == pytest.approx(summ["asymmetry"], rel=1e-5)  # 6 digits, as the app


BASE = ["--x_size", "8", "--y_size", "8", "--z_size", "4", "--t_size", "8", "--T_init", "2.0",
        "--T_final", "0.5", "--T_step", "0.5", "--OR_steps", "2", "--thermalization_steps", "2"]
DIMS = (8, 8, 4, 8)


def _conf(oracle, tmp_path):
    links = oracle.gen_su2_config(DIMS, 0.4, 33)
    conf = tmp_path / "conf.bin"
    links.tofile(conf)
    return links, ["--path_conf", str(conf), "--conf_format", "lexicographical"]


def test_approximate_copy_search_writes_spins_and_functionals(oracle, tmp_path):
    """--is_final 0 (mag_fixation.cpp:205-249): three copies, each written
    to --path_spins_output (no comparison), the functional of every copy in
    the "copy,functional" CSV; the spin file holds V spins of the last copy."""
    links, conf = _conf(oracle, tmp_path)
    spins_out, csv = tmp_path / "spins.bin", tmp_path / "F.csv"
    out = _run(conf + BASE + ["--is_final", "0", "--tolerance_digits", "6", "--gauge_copies", "3",
                              "--is_functional_save", "1", "--path_functional_output", str(csv),
                              "--path_spins_output", str(spins_out), "--seed", "5"])
    F = _num(out, "spin functional")
    assert len(F) == 3 and len(set(F)) == 3  # three different copies
    lines = csv.read_text().splitlines()
    assert lines[0] == "copy,functional"
    assert [int(l.split(",")[0]) for l in lines[1:]] == [1, 2, 3]
    assert np.allclose([float(l.split(",")[1]) for l in lines[1:]], F, rtol=0, atol=1e-15)
    s = np.fromfile(spins_out).reshape(-1, 3)
    assert s.shape[0] == int(np.prod(DIMS))
    assert abs(oracle.su2_functional_spin(links, s, DIMS) - F[-1]) < 1e-13
    assert np.max(np.abs(np.linalg.norm(s, axis=1) - 1)) < 1e-12


def test_compare_keeps_the_better_spins(oracle, tmp_path):
    """--is_compare 1 --is_compare_spins 1: a copy's spins replace the file
    only when their functional beats that of the spins in --path_previous."""
    links, conf = _conf(oracle, tmp_path)
    best = tmp_path / "best.bin"
    _run(conf + BASE + ["--is_final", "0", "--tolerance_digits", "9", "--path_spins_output",
                        str(best), "--seed", "8"])
    F_best = oracle.su2_functional_spin(links, np.fromfile(best).reshape(-1, 3), DIMS)
    worse = tmp_path / "worse.bin"
    out = _run(conf + BASE + ["--is_final", "0", "--tolerance_digits", "1", "--OR_steps", "0",
                              "--is_compare", "1", "--is_compare_spins", "1", "--path_previous",
                              str(best), "--path_spins_output", str(worse), "--seed", "9"])
    assert _num(out, "old functional is")[0] == pytest.approx(F_best, abs=1e-13)
    assert "is not saved" in out and not worse.exists()
    # resuming from the best spins (--is_new_trial 0) does not lower them: saved
    again = tmp_path / "again.bin"
    out = _run(conf + BASE + ["--is_new_trial", "0", "--is_final", "0", "--tolerance_digits", "11",
                              "--is_compare", "1", "--is_compare_spins", "1", "--path_previous",
                              str(best), "--path_spins_output", str(again)])
    F_again = _num(out, "new functional is")[0]
    assert F_again >= F_best - 1e-15
    if F_again > F_best:
        assert again.exists()


def test_resume_from_spins_equals_the_library_path(mag, oracle, tmp_path):
    """--is_new_trial 0 --path_previous S --is_final 1: the final maximisation
    from the spins in S, the gauge transformation and the link output equal
    the same calls through the library, bit for bit."""
    links, conf = _conf(oracle, tmp_path)
    S = oracle.gen_spins_uniform(DIMS, 44)
    prev = tmp_path / "prev.bin"
    S.tofile(prev)
    outp = tmp_path / "fixed.bin"
    _run(conf + BASE + ["--is_new_trial", "0", "--path_previous", str(prev), "--is_final", "1",
                        "--tolerance_maximal", "1e-9", "--tolerance_average", "1e-13",
                        "--path_conf_output", str(outp)])
    with mag.SpinSession(DIMS) as s:
        s.upload_links(links)
        s.set_spins(S)
        s.maximize_final(2, 1e-9, 1e-13, 1000000)
        s.gauge_transform()
        ref = s.download_links()
    assert np.array_equal(np.fromfile(outp).reshape(-1, 4), ref)
