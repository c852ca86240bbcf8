"""Mixed precision against FP64 on the lattices the accuracy is stated for:
SU(2) 32^4 (eps 0.5, omega 1.7, tol 1e-12) and SU(3) 24^3x48 (eps 0.4,
omega 1.6, tol 1e-11), both runs to convergence (SU(3) 24^3x48 at omega 1.6
takes about 12 400 sweeps; cap 30 000).  The FP64 run's sweeps are the
oracle's, bit for bit (tests/test_gpu_parity.py: the same sweeps at these
sizes, the same stopping sweep on a run to tol), so its result stands for
the FP64 oracle's.

Mixed: 100 FP64 warm-up sweeps, then FP32 sweeps with the functional summed
in FP64 and an FP64 rebuild of the links from the FP64 master every 10
sweeps, the FP32 phase ending at a mean relative change of 1e-10 per sweep,
then FP64 sweeps down to tol.  Stated errors against the FP64 run of the
same configuration (asserted):
  functional at the end      |dF| / F < 1e-10  (the same maximum; the
                             stopping rule |F_k - F_k-1| < tol F leaves F
                             itself uncertain at ~1e-10)
  links, U(1)-invariant part max |d(a0^2 + a3^2)| (SU(2)), max |d|U_jj|^2|
                             (SU(3)) < 1e-3, rms < 1e-5 (measured: SU(2)
                             32^4 max 8e-5; the maximum is nearly flat in a
                             few directions, along which runs that stop at
                             the same |dF| differ); the diagonal U(1) gauge
                             left free by the MAG functional is not compared
  W(1x1), W(4x4)             < 1e-12 against the input configuration (both
                             runs are FP64 gauge transformations of it)
Run with -s to see the measured values (one JSON line per lattice)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    (2, (32, 32, 32, 32), 0.5, 1.7, 1e-12),
    (3, (24, 24, 24, 48), 0.4, 1.6, 1e-11),
]


def _diag_invariant(links, group):
    if group == 2:
        return links[:, 0] ** 2 + links[:, 3] ** 2
    m = links.reshape(-1, 3, 3, 2)
    return np.stack([m[:, j, j, 0] ** 2 + m[:, j, j, 1] ** 2 for j in range(3)], axis=1)


@pytest.mark.parametrize("group,dims,eps,omega,tol", CASES)
def test_mixed_against_fp64(mag, oracle, group, dims, eps, omega, tol):
    out = {}
    for name, prec in (("fp64", mag.F64), ("mixed", mag.MIXED)):
        with mag.MagJob(group=group, dims=dims, precision=prec, omega=omega, tol=tol,
                        max_sweeps=30000, reunit_every=10, tol_mixed=1e-10,
                        fp64_warmup=100) as j:
            j.generate(eps, 7)
            if name == "fp64":
                src = j.download()
            assert j.run(), name
            s, conv, F = j.status()
            out[name] = dict(sweeps=s, converged=conv, F=F, fp64_start=j.fp64_start(),
                             links=j.download())
    wl = oracle.su2_wilson_loop if group == 2 else oracle.su3_wilson_loop
    W = {k: (wl(v["links"], dims, 1, 1), wl(v["links"], dims, 4, 4)) for k, v in out.items()}
    W0 = (wl(src, dims, 1, 1), wl(src, dims, 4, 4))
    a, b = out["fp64"], out["mixed"]
    dF = abs(b["F"] - a["F"]) / a["F"]
    dd = np.abs(_diag_invariant(b["links"], group) - _diag_invariant(a["links"], group))
    dlink = float(np.max(dd))
    dlink_rms = float(np.sqrt(np.mean(dd ** 2)))
    dW = max(abs(W[k][i] - W0[i]) for k in W for i in range(2))
    print(json.dumps({"group": group, "dims": dims, "sweeps_fp64": a["sweeps"],
                      "sweeps_mixed": b["sweeps"], "fp32_sweeps": b["fp64_start"] - 100,
                      "F_fp64": a["F"], "rel_dF": dF, "max_d_u1_invariant": dlink,
                      "rms_d_u1_invariant": dlink_rms,
                      "W1x1": W0[0], "W4x4": W0[1], "max_dW": dW}))
    assert dF < 1e-10
    assert dlink < 1e-3 and dlink_rms < 1e-5
    assert dW < 1e-12
    assert b["fp64_start"] > 100


def test_su3_32_mixed_functional_within_1e10(mag):
    """SU(3) 32^4 (eps 0.4, omega 1.6, tol 1e-11), the lattice of the
    mixed-precision roofline target, both runs to convergence: the mixed
    run's functional within 1e-10 of the FP64 run's, plaquette of either
    run equal to the input's to 1e-12 (GPU FP64 plaquette)."""
    dims = (32, 32, 32, 32)
    out = {}
    for name, prec in (("fp64", mag.F64), ("mixed", mag.MIXED)):
        with mag.MagJob(group=mag.SU3, dims=dims, precision=prec, omega=1.6, tol=1e-11,
                        max_sweeps=30000, reunit_every=10, tol_mixed=1e-10, fp64_warmup=100) as j:
            j.generate(0.4, 7)
            P0 = j.plaquette()
            assert j.run(), name
            s, conv, F = j.status()
            out[name] = dict(sweeps=s, converged=conv, F=F, P=j.plaquette(), P0=P0)
    a, b = out["fp64"], out["mixed"]
    dF = abs(b["F"] - a["F"]) / a["F"]
    print(json.dumps({"dims": dims, "sweeps_fp64": a["sweeps"], "sweeps_mixed": b["sweeps"],
                      "F_fp64": a["F"], "rel_dF": dF, "dP_fp64": abs(a["P"] - a["P0"]),
                      "dP_mixed": abs(b["P"] - b["P0"])}))
    assert dF < 1e-10
    assert abs(a["P"] - a["P0"]) < 1e-12 and abs(b["P"] - b["P0"]) < 1e-12
