"""The measured batching and serving workloads at their real sizes.

Batched gauge copies: 16 configurations (4 x 16^4, 4 x 24^4, 4 x 32^4
SU(2), 4 x 16^3x32 SU(3)), 1-4 seeded gauge copies each (33 copies), run
plain (FP64 link form, each copy alone) and batched (mixed precision, one
graph-captured batch).  Served stream: the first 10 requests of the seeded
Poisson stream (mean 20/s; SU(2) and SU(3); 16^4, 24^4, 32^4, 24^3x48 and
SU(2) 48^4), served one at a time in FP64 and with continuous batching in
mixed precision.

Every copy / request must converge in both runs, and the mixed result's
functional must be within 1e-10 relative of the plain FP64 one (the
accuracy bar of the mixed-precision forms).  Plain and mixed runs both stop
when the functional changes by less than tol per sweep; they approach the
maximum from below along different paths, so their difference is bounded
by how far each still is from the maximum.
"""
import time

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

CAP = 30000  # sweep cap: every copy and request must converge well below it
BAR = 1e-10


def test_batched_gauge_copies_workload(mag):
    specs = bench.batch_workload()
    n = sum(sp["copies"] for sp in specs)
    assert n == 33
    t0 = time.perf_counter()
    t_plain, plain = bench.run_plain(mag, specs, CAP)
    t_batch, batched = bench.run_batched(mag, specs, CAP, mag.MIXED, bench.SU3_WARMUP)
    print(f"[batched copies] plain {t_plain:.2f} s, batched {t_batch:.2f} s "
          f"({t_plain / t_batch:.2f}x), wall {time.perf_counter() - t0:.1f} s", flush=True)
    dF = np.array([abs(b[2] - p[2]) / abs(p[2]) for p, b in zip(plain, batched)])
    sw_p = np.array([p[0] for p in plain])
    sw_b = np.array([b[0] for b in batched])
    print(f"[batched copies] max rel dF {dF.max():.2e}, sweeps plain {sw_p.min()}..{sw_p.max()}, "
          f"batched {sw_b.min()}..{sw_b.max()}", flush=True)
    assert all(p[1] for p in plain), "a plain copy hit the sweep cap"
    assert all(b[1] for b in batched), "a batched copy hit the sweep cap"
    assert dF.max() <= BAR


def test_served_poisson_stream_workload(mag):
    reqs = bench.serve_workload(10)
    sizes = {(r["group"], r["dims"]) for r in reqs}
    assert {d for _, d in sizes} == set(bench.SERVE_SIZES)  # every size of the stream
    plain_res, plain = bench.serve(mag, reqs, True, CAP)
    res, served = bench.serve(mag, reqs, False, CAP)
    print(f"[served stream] plain {plain['configs_per_s']:.2f} configs/s "
          f"p50 {plain['p50_ms']:.0f} ms p99 {plain['p99_ms']:.0f} ms; served "
          f"{served['configs_per_s']:.2f} configs/s p50 {served['p50_ms']:.0f} ms "
          f"p99 {served['p99_ms']:.0f} ms", flush=True)
    for r, a, b in zip(reqs, res, plain_res):
        print(f"  SU({r['group']}) {r['dims']}: sweeps plain {b['sweeps']} served {a['sweeps']} "
              f"rel dF {abs(a['F'] - b['F']) / abs(b['F']):.2e}", flush=True)
    assert all(r["status"] == 0 for r in res + plain_res)
    assert all(r["converged"] for r in plain_res), "a plain request hit the sweep cap"
    assert all(r["converged"] for r in res), "a served request hit the sweep cap"
    dF = max(abs(a["F"] - b["F"]) / abs(b["F"]) for a, b in zip(res, plain_res))
    assert dF <= BAR
    assert served["peak_active"] >= 2
