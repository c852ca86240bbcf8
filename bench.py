"""MAG gauge-fixing throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

A step is one checkerboard omega-overrelaxation sweep (V site updates) of a
synthetic configuration resident in HBM.  The headline workload is SU(2)
32^4 (eps 0.5, omega 1.7) in plain FP64; the same lattice in mixed
precision (FP32 sweeps, FP64 functional, FP64 master rebuild every 10
sweeps) and SU(3) 32^4 / 24^3x48 are reported beside it.  Rank 0 prints one
JSON line.

Roofline: the dominant kernel is the half-sweep kernel.  Its algorithmic
bytes per launch are (V/2) site updates x B_site, B_site = 8 links x
reals/link x bytes x 2 (read + write): 512 / 256 B (SU(2) FP64 / FP32),
1536 / 768 B (SU(3) FP64 / FP32).  The mixed form's red-site gauge (read +
write, red half only) is traffic beyond that and is not counted.  The mean
launch duration is measured with HIP events on the job's stream around
back-to-back half-sweep launches (mag_job_halfsweep_timed); a whole step
(both half-sweeps, the periodic FP64 rebuild, the reductions) is timed the
same way and priced with the same B_site (step_frac).  Peak HBM = 8.0 TB/s.
roofline.traffic is the HBM bytes per launch from a separate rocprofv3 PMC
run (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel of known size)
recorded in profiles/halfsweep_hbm_traffic_pmc.md (traffic_source).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def pmc_traffic(kernels, lattice: str):
    """HBM bytes per launch of half-sweep kernels from the committed PMC
    summary: the mean over `kernels` (the red and the black half of the
    mixed FP32 phase are different kernels launched alternately), or None
    when it holds no entry for one of them at that lattice."""
    path = os.path.join(ROOT, "profiles", "halfsweep_hbm_traffic_pmc.md")
    try:
        text = open(path).read()
        block = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        vals = [block.get(f"{k}|{lattice}", {}).get("traffic_bytes") for k in kernels]
    except (OSError, IndexError, ValueError):
        return None
    return None if None in vals else sum(vals) / len(vals)


def b_site(group: int, mixed: bool) -> int:
    """Algorithmic bytes per site update: 8 links read and written."""
    nr = 4 if group == 2 else 12
    return 8 * nr * (4 if mixed else 8) * 2


def b_site_moved(group: int, mixed: bool) -> float:
    """Bytes the half-sweeps move per site update: the links, plus in the
    mixed form the red sites' gauge (read + write in the red half)."""
    nr = 4 if group == 2 else 12
    return b_site(group, mixed) + (nr * 4 if mixed else 0)


def lattice_name(dims) -> str:
    if len(set(dims)) == 1:
        return f"{dims[0]}^4"
    if len(set(dims[:3])) == 1:
        return f"{dims[0]}^3x{dims[3]}"
    return "x".join(map(str, dims))


def measure(G, group, dims, precision, eps, omega, steps, warmup, seed=2024):
    V = 1
    for d in dims:
        V *= d
    job = G.MagJob(group=group, dims=dims, precision=precision, omega=omega, tol=0.0,
                   max_sweeps=2 * steps + warmup + 4, reunit_every=10, tol_mixed=0.0)
    try:
        job.generate(eps, seed)
        if warmup:
            job.sweeps_timed(warmup)
        else:
            job.sweeps(1)  # the autotuner times the launch shapes on a job's first sweep
        # whole steps (sweep = 2 half-sweeps + reduction), then the dominant
        # kernel's mean duration over back-to-back half-sweep launches
        total_ms = job.sweeps_timed(steps)
        kern_ms = job.halfsweep_timed(2 * steps)
        sweeps, _, F = job.status()
        shape = job.shape(fp32=precision == G.MIXED)  # chosen by the autotuner in the warm-up
    finally:
        job.close()
    mixed = precision == G.MIXED
    bytes_launch = (V // 2) * b_site(group, mixed)
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    step_ms = total_ms / steps
    # the mixed FP32 phase: red halves accumulate the site gauge, black halves
    # do not; lattices whose links exceed the 256 MB MALL run the streaming
    # kernels (k_su2_half_st: SU(2) in both precisions, k_su3_half_st: SU(3) FP32)
    link_bytes = 8 * (V // 2) * (4 if group == 2 else 12) * (4 if mixed else 8)
    st = "_st" if link_bytes > (256 << 20) and (group == 2 or mixed) else ""
    knames = ([f"void mag::k_su{group}_half{st}<float, true>", f"void mag::k_su{group}_half{st}<float, false>"]
              if mixed else [f"void mag::k_su{group}_half{st}<double, false>"])
    traffic = pmc_traffic(knames, lattice_name(dims))
    return {
        "value": V * steps / (total_ms * 1e-3),
        "ms_per_step": step_ms,
        "kernel_ms": kern_ms,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": None if traffic is None else int(round(traffic)),
            "bytes_per_launch": bytes_launch,
            "b_site": b_site(group, mixed),
        },
        "step_frac": round(V * b_site(group, mixed) / (step_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "moved_frac": round((V // 2) * b_site_moved(group, mixed) / (kern_ms * 1e-3) / 1e9
                            / PEAK_HBM_GBS, 4),
        "F_after": F,
        "sweeps": sweeps,
        "shape": list(shape),
    }


# SU(3) site form (read-only links, one site gauge per configuration): per
# site update the 8 links are read (8 x 12 FP32 reals, 384 B) and the
# site's gauge read and written (2 x 12 x 4 B); the neighbours' gauges are
# read as well but shared by up to 8 sites through the caches
SITE3_B_SITE_MIN = 8 * 12 * 4 + 2 * 12 * 4


def measure_copies(G, group, dims, precision, copies, eps, omega, steps, warmup, seed=2024):
    """The site forms' half-sweep kernel alone (mag_job_halfsweep_timed):
    SU(2) spin form or SU(3) site form, `copies` gauge copies of one
    configuration, FP32 phase (MIXED) or FP64.  Rate in site updates of all
    copies per second; bytes moved at least per site update and copy: the
    links once for all copies, each copy's own site variable read and
    written (SU(2) spins 3 reals, SU(3) gauges 12)."""
    V = 1
    for d in dims:
        V *= d
    job = G.MagJob(group=group, dims=dims, precision=precision, omega=omega, tol=0.0,
                   max_sweeps=2 * steps + warmup + 4, form=G.FORM_SPIN, copies=copies,
                   tol_mixed=0.0, fp64_warmup=0)
    try:
        job.generate(eps, seed)
        job.sweeps_timed(max(warmup, 1))
        kern_ms = job.halfsweep_timed(2 * steps)
    finally:
        job.close()
    w = 4 if precision == G.MIXED else 8
    nr, sr = (4, 3) if group == 2 else (12, 12)
    b_min = 8 * nr * w + copies * 2 * sr * w
    moved = (V // 2) * b_min / (kern_ms * 1e-3) / 1e9
    return {"copies": copies, "kernel_ms": kern_ms,
            "site_updates_per_s": copies * (V // 2) / (kern_ms * 1e-3),
            "b_min_per_site": b_min, "min_moved_GBs": round(moved, 1),
            "min_moved_frac": round(moved / PEAK_HBM_GBS, 4)}


def measure_site3(G, dims, eps, omega, steps, warmup, seed=2024):
    """Sweep rate of the SU(3) site form, one configuration, FP32 phase (the
    served form of an SU(3) request), priced two ways: against the link-form
    roofline of the same site updates (B_site 768, the SU(3) FP32 line of
    the roofline table) and by the bytes it moves at least (links read once,
    site gauge read and written)."""
    V = 1
    for d in dims:
        V *= d
    job = G.MagJob(group=3, dims=dims, precision=G.MIXED, omega=omega, tol=0.0,
                   max_sweeps=2 * steps + warmup + 4, form=G.FORM_SPIN, copies=1, tol_mixed=0.0,
                   fp64_warmup=0)
    try:
        job.generate(eps, seed)
        job.sweeps_timed(max(warmup, 1))
        total_ms = job.sweeps_timed(steps)
        kern_ms = job.halfsweep_timed(2 * steps)
    finally:
        job.close()
    step_ms = total_ms / steps
    rate = V / (step_ms * 1e-3)
    link_peak = PEAK_HBM_GBS * 1e9 / b_site(3, True)
    moved = (V // 2) * SITE3_B_SITE_MIN / (kern_ms * 1e-3) / 1e9
    return {"site_updates_per_s": rate, "ms_per_step": step_ms, "kernel_ms": kern_ms,
            "link_roofline_frac": round(rate / link_peak, 4),
            "b_site_min": SITE3_B_SITE_MIN, "min_moved_GBs": round(moved, 1),
            "min_moved_frac": round(moved / PEAK_HBM_GBS, 4),
            "form": "site form, one site gauge over read-only FP32 links, FP32 phase"}


# algorithmic bytes per site of abelian projection + plaquettes + currents:
# SU(2) angles 4 links x 4 doubles read + 4 angles written, plaquettes 4
# angles read + 6 written, currents 6 read + 4 written; SU(3) 12-real links
# and three colours
MONO_B_SITE = {2: (16 + 4) * 8 + (4 + 6) * 8 + (6 + 4) * 8,
               3: (48 + 12) * 8 + (12 + 18) * 8 + (18 + 12) * 8}


def monopole_measure(G, group, dims, eps, reps=20, cpu=False):
    """Device time of the abelian projection and monopole currents of a
    job's links (mag_job_monopole_timed: angles + plaquettes + currents,
    no host transfer) and its HBM rate against MONO_B_SITE."""
    V = 1
    for d in dims:
        V *= d
    with G.MagJob(group=group, dims=dims) as job:
        job.generate(eps, 5)
        ms = job.monopole_timed(reps)
        J = job.monopole_currents()
        nmono = int(abs(J.round()).sum())
        cl_ms = job.clusters_timed(reps)
        recs = job.monopole_clusters()
    recs = [recs] if group == 2 else recs
    gbs = V * MONO_B_SITE[group] / (ms * 1e-3) / 1e9
    out = {"ms": ms, "sites_per_s": V / (ms * 1e-3), "b_site": MONO_B_SITE[group],
           "achieved_GBs": round(gbs, 1), "roofline_frac": round(gbs / PEAK_HBM_GBS, 4),
           "current_links": nmono,
           "workload": f"generated links eps {eps} (unfixed), 3 kernels per call",
           "clusters": {"ms": cl_ms, "count": [len(r) for r in recs],
                        "largest": [int(r[:, 1].max()) if len(r) else 0 for r in recs],
                        "kernels": "5 per colour (init, hook, flatten, sums, collect)"}}
    if cpu:
        out["clusters"]["cpu"] = cpu_clusters(J.reshape(-1, 4 * V), dims)
    return out


def cpu_clusters(J, dims):
    """The oracle's restatement of the reference's cluster walk
    (calculate_clusters, one host thread as in the reference) on the same
    currents, every colour."""
    import numpy as np
    from oracle import oracle as O
    O.build()
    t0 = time.perf_counter()
    for k in range(J.shape[0]):
        O.monopole_clusters(np.ascontiguousarray(J[k]), dims)
    return {"ms": (time.perf_counter() - t0) * 1e3, "threads": 1,
            "sample": "the same currents, all colours, one run"}


# the annealing schedule of the reference app's defaults
# (apps/mag_su2/mag_fixation.cpp: T_init 2.5, T_final 0.01, T_step 0.01,
# OR_steps 6, thermalization_steps 50)
ANNEAL = dict(T_init=2.5, T_final=0.01, T_step=0.01, OR_steps=6, thermalization_steps=50)


def anneal_sweeps(T_init, T_final, T_step, OR_steps, thermalization_steps):
    """Sweeps of make_simulated_annealing (mag.cpp:1585-1603): the
    thermalisation heat-bath sweeps, then per temperature one heat-bath and
    OR_steps overrelaxation sweeps, steps of T_step / 4 in [1.2, 1.6]."""
    n, T = thermalization_steps, T_init
    while T > T_final / 2:
        n += 1 + OR_steps
        T -= T_step / 4 if (1.2 <= T <= 1.6 + T_step) else T_step
    return n


def annealing_measure(G, dims, seed=11):
    """Simulated annealing of the spin form on the GPU (the reference app's
    first and longest phase) with the app's default schedule."""
    V = 1
    for d in dims:
        V *= d
    with G.MagJob(group=2, dims=dims) as job:
        job.generate(0.5, seed)
        links = job.download()
    with G.SpinSession(dims) as s:
        s.upload_links(links)
        s.generate_spins(seed)
        s.simulated_annealing(1.0, 0.9, 0.1, 1, 0, seed)  # warm-up
        t0 = time.perf_counter()
        s.simulated_annealing(seed=seed, **ANNEAL)
        t = time.perf_counter() - t0
    n = anneal_sweeps(**ANNEAL)
    return {"s": round(t, 4), "sweeps": n, "site_updates_per_s": V * n / t, "schedule": ANNEAL}


def cpu_annealing(dims, seconds):
    """The oracle's checkerboard heat-bath + 6 OR sweeps per temperature on
    the host's threads (the algorithm the GPU runs), for about `seconds`."""
    from oracle import oracle as O
    O.build()
    V = 1
    for d in dims:
        V *= d
    links = O.gen_su2_config(dims, 0.5, 11)
    spins = O.gen_spins_uniform(dims, 11)
    steps, k, t0 = 0, 0, time.perf_counter()
    while steps < 2 or time.perf_counter() - t0 < seconds:
        for p in (0, 1):
            O.su2_heatbath_halfsweep(links, spins, dims, p, 1.0, 11, k)
        k += 1
        for _ in range(ANNEAL["OR_steps"]):
            for p in (0, 1):
                O.su2_spin_halfsweep(links, spins, dims, p, 0)
        steps += 1
    t = time.perf_counter() - t0
    return {"site_updates_per_s": V * (1 + ANNEAL["OR_steps"]) * steps / t,
            "threads": O.get_threads(), "temperature_steps": steps,
            "sample": "checkerboard heat bath + 6 OR sweeps per step (oracle), same lattice"}


def _median_rate(fn, V, seconds, min_runs=5):
    """site-updates/s of fn() (one sweep): median over runs after one
    warm-up, for about `seconds` (at least min_runs runs)."""
    fn()
    ts = []
    t_end = time.perf_counter() + seconds
    while len(ts) < min_runs or (time.perf_counter() < t_end and len(ts) < 2000):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return V / ts[len(ts) // 2], len(ts)


def cpu_baseline(dims, eps, omega, seconds):
    """CPU timings on the host's cores, same lattice and configuration:
    (i) the reference's own sweep as it runs there -- the spin form,
    lexicographic, in place, OpenMP collapse(4) (mag.cpp:1375-1401, racy
    as in the reference) -- from the C restatement compiled with the
    reference's flags (-Ofast -march=native, makefile:5) on this host: the
    reported value; the same with the oracle's own build (-O3, no
    contraction) beside it; and (ii) the oracle's checkerboard omega link
    sweep, the algorithm the GPU runs.  Each with all threads and with one
    thread; median of >= 5 sweeps after a warm-up."""
    from oracle import oracle as O
    O.build()
    Fl = O.fast_lib()
    threads = O.get_threads()
    links = O.gen_su2_config(dims, eps, 2024)
    spins = O.gen_spins_uniform(dims, 2024)
    V = 1
    for d in dims:
        V *= d
    q = seconds / 6.0
    fast_sweep = lambda: Fl.or_su2_or_sweep_lexi_omp(O._d(links), O._d(spins), O._dims(dims))  # noqa: E731
    fast_all, n_fast = _median_rate(fast_sweep, V, q)
    spin_all, n_spin = _median_rate(lambda: O.su2_or_sweep_lexi_omp(links, spins, dims), V, q)
    link_all, n_link = _median_rate(lambda: O.su2_link_sweep(links, dims, omega), V, q)
    O.set_threads(1)
    Fl.or_set_threads(1)
    try:
        fast_one, _ = _median_rate(fast_sweep, V, q)
        spin_one, _ = _median_rate(lambda: O.su2_or_sweep_lexi_omp(links, spins, dims), V, q)
        link_one, _ = _median_rate(lambda: O.su2_link_sweep(links, dims, omega), V, q)
    finally:
        O.set_threads(threads)
        Fl.or_set_threads(threads)
    lat = "x".join(map(str, dims))
    return {
        "value": fast_all,
        "unit": "site-updates/s",
        "cores": threads,
        "cpu_model": O.cpu_model(),
        "kind": "port",
        "sample": f"SU(2) {lat} FP64: the reference's spin overrelaxation sweep as it runs "
                  "there (lexicographic, in place, OpenMP collapse(4), mag.cpp:1375-1401), "
                  "restated in C and compiled with the reference's flags (gcc -Ofast "
                  f"-march=native -fopenmp) on this host, {threads} threads, median of "
                  f"{n_fast} sweeps after a warm-up",
        "single_thread": fast_one,
        "oracle_build": {"value": spin_all, "single_thread": spin_one,
                         "flags": "gcc -O3 -march=x86-64-v2 -ffp-contract=off -fopenmp",
                         "sweeps": n_spin},
        "algorithm_matched": {
            "value": link_all,
            "single_thread": link_one,
            "sample": f"SU(2) {lat} checkerboard omega={omega} link sweeps of the oracle "
                      f"(the GPU's algorithm), {n_link} sweeps",
        },
    }


def batch_workload(seed=3):
    """16 configurations: 4 x 16^4, 4 x 24^4, 4 x 32^4 SU(2) (eps 0.5,
    omega 1.7, tol 1e-12) and 4 x 16^3x32 SU(3) (eps 0.4, omega 1.6, tol
    1e-11); configuration i has 1-4 random gauge copies (count from a
    seeded numpy Generator), copy c generated with gauge key c."""
    import numpy as np
    rng = np.random.default_rng(seed)
    copies = rng.integers(1, 5, size=16)
    specs = []
    for i in range(16):
        if i < 12:
            dims = [(16,) * 4, (24,) * 4, (32,) * 4][i // 4]
            specs.append(dict(group=2, dims=dims, eps=0.5, omega=1.7, tol=1e-12,
                              seed=1000 + i, copies=int(copies[i])))
        else:
            specs.append(dict(group=3, dims=(16, 16, 16, 32), eps=0.4, omega=1.6, tol=1e-11,
                              seed=1000 + i, copies=int(copies[i])))
    return specs


def run_plain(G, specs, max_sweeps):
    """Every gauge copy alone, one after the other, plain FP64 link form."""
    out = []
    total = 0.0
    for sp in specs:
        for c in range(sp["copies"]):
            job = G.MagJob(group=sp["group"], dims=sp["dims"], omega=sp["omega"], tol=sp["tol"],
                           max_sweeps=max_sweeps)
            try:
                job.generate(sp["eps"], sp["seed"], copy=c)
                t0 = time.perf_counter()
                job.run()
                total += time.perf_counter() - t0
                sw, conv, F = job.status()
            finally:
                job.close()
            out.append((sw, conv, F))
    return total, out


# FP64 warm-up sweeps before the mixed phase: the spin form's FP32-rounded
# links are a fixed perturbation of the whole problem, and a trajectory
# that passes near a saddle during the FP32 phase can end on a neighbouring
# maximum (another Gribov copy).  Measured on the batched and served
# workloads: 100 sweeps moved one 48^4 request, 300 moved one 24^4 copy of
# the batch, 600 moved none; FP64 spin sweeps throughout never move any
# and cost 7% more time on the batch.
SU2_WARMUP, SU3_WARMUP = 600, 100

# Mixed link form: the FP32 phase hands over to FP64 sweeps while the
# functional still changes by 10 tol per sweep, so the FP64 stopping rule --
# the plain run's -- decides where the run ends.  Measured on the eight
# SU(3) 16^3x32 copies of the batched workload: with tol_mixed = tol the
# FP32 phase runs past the plain run's stopping point (max dF/F 1.3e-10,
# F above the plain result); with 10 tol max dF/F 2.3e-13, same sweep counts.
LINK_TOL_MIXED = 10.0

# SU(3) site form: its FP32 phase tests F, evaluated in FP64 every 10
# sweeps, on the mean change per sweep since the previous evaluation and
# hands over to FP64 at 1.2 tol; the FP64 sweeps then stop on the plain
# run's rule.  Measured on the 13 SU(3) requests of the serving workload
# (16^4 - 32^4, omega 1.6, tol 1e-11), each alone: hand-over at 2 / 1.5 /
# 1.2 tol: 6.12 / 5.93 / 5.78 s, FP64 polishes of up to 2648 / 1458 / 618
# sweeps, sweep counts the plain run's (+-1), max dF/F 9.7e-12 / 9.8e-12 /
# 9.8e-12; at 1.0 tol the FP32 phase itself meets the plain rule and runs
# to the end of its 10-sweep period past the plain stopping point (5-15
# sweeps more, max dF/F 1.5e-10).
SITE_TOL_MIXED = 1.2


def run_batched(G, specs, max_sweeps, su3_precision, su3_warmup, su3_site=True,
                su2_warmup=SU2_WARMUP, su2_precision=None):
    """All configurations in one graph-captured batch: each SU(2)
    configuration is one spin-form copy set (links read once for all its
    copies; mixed: FP32-rounded links and FP32 spins after the FP64
    warm-up, field and functional summed in FP64), each SU(3) configuration
    one site-form copy set (read-only links, one site gauge per copy;
    su3_site=False: each SU(3) copy a link-form job); `su3_precision`
    applies to both unless `su2_precision` is given.  Measured on the
    config-3 workload: SU(3) in the site form 1.57 s, in the link form
    1.94 s."""
    su2_precision = su3_precision if su2_precision is None else su2_precision
    jobs = []
    for sp in specs:
        if sp["group"] == 2:
            j = G.MagJob(group=2, dims=sp["dims"], omega=sp["omega"], tol=sp["tol"],
                         max_sweeps=max_sweeps, form=G.FORM_SPIN, copies=sp["copies"],
                         precision=su2_precision, tol_mixed=sp["tol"],
                         fp64_warmup=su2_warmup if su2_precision == G.MIXED else 0)
            j.generate(sp["eps"], sp["seed"])
            jobs.append((j, list(range(sp["copies"]))))
        elif su3_site:
            j = G.MagJob(group=3, dims=sp["dims"], omega=sp["omega"], tol=sp["tol"],
                         max_sweeps=max_sweeps, form=G.FORM_SPIN, copies=sp["copies"],
                         precision=su3_precision, tol_mixed=SITE_TOL_MIXED * sp["tol"],
                         fp64_warmup=su3_warmup)
            j.generate(sp["eps"], sp["seed"])
            jobs.append((j, list(range(sp["copies"]))))
        else:
            for c in range(sp["copies"]):
                j = G.MagJob(group=3, dims=sp["dims"], omega=sp["omega"], tol=sp["tol"],
                             max_sweeps=max_sweeps, precision=su3_precision, reunit_every=10,
                             tol_mixed=LINK_TOL_MIXED * sp["tol"], fp64_warmup=su3_warmup)
                j.generate(sp["eps"], sp["seed"], copy=c)
                jobs.append((j, [None]))
    batch = G.MagBatch()
    try:
        for j, _ in jobs:
            batch.add(j)
        t0 = time.perf_counter()
        batch.run()
        el = time.perf_counter() - t0
        out = []
        for j, cs in jobs:
            for c in cs:
                out.append(j.status() if c is None else j.copy_status(c))
    finally:
        batch.close()
        for j, _ in jobs:
            j.close()
    return el, out


# Sweep cap of the batched and served workloads: every copy and request
# converges below it (SU(3) 24^3x48 at omega 1.6 needs about 12 400 sweeps,
# 32^4 about 14 200).
WORKLOAD_CAP = 30000


def batch_bench(G, max_sweeps=WORKLOAD_CAP, su3_precision=None, su3_warmup=SU3_WARMUP):
    specs = batch_workload()
    n = sum(sp["copies"] for sp in specs)
    su3_precision = G.MIXED if su3_precision is None else su3_precision
    t_plain, plain = run_plain(G, specs, max_sweeps)
    t_batch, batched = run_batched(G, specs, max_sweeps, su3_precision, su3_warmup)
    dF = [abs(b[2] - p[2]) / abs(p[2]) for p, b in zip(plain, batched) if p[1] and b[1]]
    both = sum(1 for p, b in zip(plain, batched) if p[1] and b[1])
    return {
        "workload": "16 configurations (4x16^4, 4x24^4, 4x32^4 SU(2), 4x16^3x32 SU(3)), "
                    f"{n} gauge copies, each run to convergence",
        "copies": n,
        "plain_fp64_s": round(t_plain, 4),
        "plain_configs_per_s": n / t_plain,
        "batched_s": round(t_batch, 4),
        "batched_configs_per_s": n / t_batch,
        "speedup": t_plain / t_batch,
        "su3_precision": "mixed" if su3_precision == G.MIXED else "fp64",
        "su3_form": "site form (copy sets over read-only links)",
        "fp64_warmup": {"su2": SU2_WARMUP, "su3": su3_warmup},
        "converged_plain": sum(1 for p in plain if p[1]),
        "converged_batched": sum(1 for b in batched if b[1]),
        "max_rel_dF_vs_plain_converged": max(dF) if dF else None,
        "copies_rel_dF_above_1e-10": sum(1 for d in dF if d > 1e-10),
        "converged_both": both,
        "speedup_converged_only": (t_plain / t_batch) if both == n else None,
        "tol_mixed_su3_link": f"{LINK_TOL_MIXED:g} x tol",
        "sweeps_plain": [p[0] for p in plain],
        "sweeps_batched": [b[0] for b in batched],
    }


SERVE_SIZES = [(16,) * 4, (24,) * 4, (32,) * 4, (24, 24, 24, 48), (48,) * 4]


def serve_workload(n=40, rate=20.0, seed=4):
    """Poisson arrivals (mean `rate` per second, seeded): 60% SU(2) (eps 0.5,
    omega 1.7, tol 1e-12) over the five sizes, 40% SU(3) (eps 0.4, omega
    1.6, tol 1e-11) over the four that are not SU(2)-only (48^4)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for i in range(n):
        su2 = bool(rng.random() < 0.6)
        dims = SERVE_SIZES[int(rng.integers(0, 5 if su2 else 4))]
        reqs.append(dict(t=float(t[i]), group=2 if su2 else 3, dims=dims, eps=0.5 if su2 else 0.4,
                         omega=1.7 if su2 else 1.6, tol=1e-12 if su2 else 1e-11, seed=5000 + i))
    return reqs


def serve_params(G, r, plain, max_sweeps, su3_tol_mixed=SITE_TOL_MIXED * 1e-11,
                 su3_warmup=SU3_WARMUP, su2_warmup=SU2_WARMUP, su3_site=True):
    """Job parameters of request r: plain FP64 link form, or the served
    form (see serve)."""
    p = G.MagParams(group=r["group"], dims=r["dims"], omega=r["omega"], tol=r["tol"],
                    max_sweeps=max_sweeps)
    if not plain:
        p.precision = G.MIXED
        if r["group"] == 2:
            p.form, p.copies, p.tol_mixed = G.FORM_SPIN, 1, r["tol"]
            p.fp64_warmup = su2_warmup
        else:
            p.tol_mixed, p.fp64_warmup = su3_tol_mixed, su3_warmup
            if su3_site:
                p.form, p.copies = G.FORM_SPIN, 1
    return p


def serve(G, reqs, plain, max_sweeps, su3_tol_mixed=SITE_TOL_MIXED * 1e-11, su3_warmup=SU3_WARMUP,
          su2_warmup=SU2_WARMUP, su3_site=True):
    """Replay the arrival stream against a MagServer.  plain: one job at a
    time, FP64 link form (the reference's one-configuration-per-run use);
    else continuous batching, mixed precision after an FP64 warm-up: SU(2)
    in the spin form (FP32-rounded links and FP32 spins, field and
    functional summed in FP64), SU(3) in the site form (read-only links,
    one site gauge; su3_site=False: the link form)."""
    def params(r):
        return serve_params(G, r, plain, max_sweeps, su3_tol_mixed, su3_warmup, su2_warmup,
                            su3_site)

    with G.MagServer(max_active=1 if plain else 0) as srv:
        t0 = time.perf_counter()
        ids = []
        for r in reqs:
            dt = r["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(params(r), eps=r["eps"], seed=r["seed"]))
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    lat = sorted(r["t_done"] - r["t_submit"] for r in res)
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]  # noqa: E731
    return res, {
        "p50_ms": round(1e3 * pct(0.50), 2),
        "p99_ms": round(1e3 * pct(0.99), 2),
        "configs_per_s": len(res) / span,
        "configs_per_s_converged": sum(1 for r in res if r["converged"]) / span,
        "makespan_s": round(span, 3),
        "converged": sum(1 for r in res if r["converged"]),
        "peak_active": st["peak_active"],
        "peak_hbm_MiB": round(st["peak_hbm"] / 2 ** 20, 1),
    }


def serve_alone(G, reqs, max_sweeps):
    """Every request run alone, one after the other, in the served form:
    the time the served run would take if concurrency bought nothing."""
    out = {}
    for r in reqs:
        with G.MagJob(params=serve_params(G, r, False, max_sweeps)) as j:
            j.generate(r["eps"], r["seed"])
            t0 = time.perf_counter()
            j.run()
            dt = time.perf_counter() - t0
        key = f"su{r['group']}_{lattice_name(r['dims'])}"
        e = out.setdefault(key, {"requests": 0, "s": 0.0})
        e["requests"] += 1
        e["s"] = round(e["s"] + dt, 4)
    tot = lambda g: round(sum(v["s"] for k, v in out.items() if k.startswith(g)), 3)  # noqa: E731
    return {"by_group_and_size": out, "su2_s": tot("su2"), "su3_s": tot("su3"),
            "total_s": tot("su")}


def serve_bench(G, n=40, max_sweeps=WORKLOAD_CAP, alone=False):
    reqs = serve_workload(n)
    plain_res, plain = serve(G, reqs, True, max_sweeps)
    res, served = serve(G, reqs, False, max_sweeps)
    # functional agreement over the requests that converged in both runs
    dF = [abs(a["F"] - b["F"]) / abs(b["F"]) for a, b in zip(res, plain_res)
          if a["converged"] and b["converged"]]
    off = [dict(group=r["group"], dims=r["dims"], F_plain=b["F"], F_served=a["F"],
                sweeps_plain=b["sweeps"], sweeps_served=a["sweeps"])
           for r, a, b in zip(reqs, res, plain_res)
           if a["converged"] and b["converged"] and abs(a["F"] - b["F"]) > 1e-10 * abs(b["F"])]
    return {
        "workload": f"{n} requests, Poisson 20/s (seeded), 60% SU(2) / 40% SU(3), sizes "
                    "16^4, 24^4, 32^4, 24^3x48, 48^4 (SU(2) only); sweep cap "
                    f"{max_sweeps}",
        "plain_fp64_one_at_a_time": plain,
        "continuous_batching": served,
        "throughput_ratio": served["configs_per_s"] / plain["configs_per_s"],
        "max_rel_dF_vs_plain_converged": max(dF) if dF else None,
        "requests_compared": len(dF),
        "requests_rel_dF_above_1e-10": sum(1 for d in dF if d > 1e-10),
        "requests_off": off,
        "converged_both": sum(1 for a, b in zip(res, plain_res) if a["converged"] and b["converged"]),
        "throughput_ratio_converged_only":
            (served["configs_per_s_converged"] / plain["configs_per_s_converged"])
            if plain["configs_per_s_converged"] else None,
        **({"served_form_alone": serve_alone(G, reqs, max_sweeps)} if alone else {}),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="extra lines: the sweep and site-form kernels only (profiling runs)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-alone", action="store_true",
                    help="also run every served request alone in its served form "
                         "(serving.served_form_alone)")
    ap.add_argument("--serve-requests", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent MAG jobs shard one per GPU with no communication)", file=sys.stderr)
        sys.exit(2)

    import general_code_amd as G

    dims = (32, 32, 32, 32)
    eps, omega = 0.5, 1.7
    copy_gbs = G.hbm_copy_bandwidth(1 << 30, 20)
    print(f"[bench] copy kernel {copy_gbs:.0f} GB/s; SU(2) {dims} FP64 steps={args.steps} "
          f"warmup={args.warmup}", file=sys.stderr, flush=True)
    head = measure(G, 2, dims, G.F64, eps, omega, args.steps, args.warmup)

    extra = {}
    if not args.no_extra:
        for name, group, d, prec, e, om in [
            ("su2_32^4_mixed", 2, dims, G.MIXED, eps, omega),
            ("su3_32^4_mixed", 3, dims, G.MIXED, 0.4, 1.6),
            ("su3_32^4_fp64", 3, dims, G.F64, 0.4, 1.6),
            ("su3_24^3x48_mixed", 3, (24, 24, 24, 48), G.MIXED, 0.4, 1.6),
            ("su3_24^3x48_fp64", 3, (24, 24, 24, 48), G.F64, 0.4, 1.6),
            # working sets beyond the 256 MB MALL (links 648 / 324 MiB, 486 MiB)
            ("su2_48^4_fp64", 2, (48,) * 4, G.F64, eps, omega),
            ("su2_48^4_mixed", 2, (48,) * 4, G.MIXED, eps, omega),
            ("su3_48^3x24_mixed", 3, (48, 48, 48, 24), G.MIXED, 0.4, 1.6),
        ]:
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            r = measure(G, group, d, prec, e, om, args.steps, args.warmup)
            extra[name] = {
                "site_updates_per_s": r["value"],
                "ms_per_step": r["ms_per_step"],
                "kernel_ms": r["kernel_ms"],
                "roofline_frac": r["roofline"]["frac"],
                "step_frac": r["step_frac"],
                "moved_frac": r["moved_frac"],
                "achieved_GBs": r["roofline"]["achieved"],
                "traffic_bytes_per_launch": r["roofline"]["traffic"],
                "alg_bytes_per_launch": r["roofline"]["bytes_per_launch"],
                "b_site": r["roofline"]["b_site"],
                "launch_shape": r["shape"],
            }
        for name, d in [("su3_32^4_mixed_site", dims), ("su3_24^3x48_mixed_site", (24, 24, 24, 48))]:
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            extra[name] = measure_site3(G, d, 0.4, 1.6, args.steps, args.warmup)
        # the half-sweep kernels of the batched and served forms
        print("[bench] site forms", file=sys.stderr, flush=True)
        extra["site_form_kernels"] = {
            name: measure_copies(G, group, d, prec, k, e, om, args.steps, args.warmup)
            for name, group, d, prec, k, e, om in [
                ("su2_32^4_spin_fp32_1copy", 2, dims, G.MIXED, 1, eps, omega),
                ("su2_32^4_spin_fp32_4copies", 2, dims, G.MIXED, 4, eps, omega),
                ("su2_48^4_spin_fp32_1copy", 2, (48,) * 4, G.MIXED, 1, eps, omega),
                ("su2_48^4_spin_fp64_1copy", 2, (48,) * 4, G.F64, 1, eps, omega),
                ("su3_32^4_site_fp32_1copy", 3, dims, G.MIXED, 1, 0.4, 1.6),
                ("su3_32^4_site_fp32_4copies", 3, dims, G.MIXED, 4, 0.4, 1.6),
                ("su3_32^4_site_fp64_1copy", 3, dims, G.F64, 1, 0.4, 1.6),
                ("su3_48^3x24_site_fp32_1copy", 3, (48, 48, 48, 24), G.MIXED, 1, 0.4, 1.6),
            ]}

    if not args.no_extra and not args.kernels_only:
        for name, group, d, e in [("monopoles_su2_32^4", 2, dims, 1.5),
                                  ("monopoles_su3_24^3x48", 3, (24, 24, 24, 48), 1.2)]:
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            extra[name] = monopole_measure(G, group, d, e, cpu=not args.no_cpu)
        print("[bench] su2_32^4_annealing", file=sys.stderr, flush=True)
        extra["su2_32^4_annealing"] = annealing_measure(G, dims)
        if not args.no_cpu:
            extra["su2_32^4_annealing"]["cpu"] = cpu_annealing(dims, 8.0)

    batched = None
    if not args.no_batch:
        print("[bench] batched gauge copies", file=sys.stderr, flush=True)
        batched = batch_bench(G)
    served = None
    if not args.no_serve:
        print("[bench] serving", file=sys.stderr, flush=True)
        served = serve_bench(G, args.serve_requests, alone=args.serve_alone)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(dims, eps, omega, args.cpu_seconds)

    line = {
        "metric": "MAG site-updates/s",
        "value": head["value"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox exp-map links + Haar gauge transform, generated on GPU)",
        "config": {
            "workload": "SU(2) MAG 32^4, eps 0.5, checkerboard omega-OR omega=1.7, plain FP64; "
                        "one step = one sweep",
            "lattice": list(dims),
            "group": "SU(2)",
            "omega": omega,
        },
        "roofline": {k: head["roofline"][k] for k in
                     ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "traffic_source": "profiles/halfsweep_hbm_traffic_pmc.md: rocprofv3 --pmc FETCH_SIZE "
                          "and WRITE_SIZE in separate runs of this kernel and lattice",
        "step_frac": head["step_frac"],
        "kernel": "k_su2_half<double> (one checkerboard half-sweep)",
        "kernel_ms": head["kernel_ms"],
        "launch_shape": head["shape"],
        "b_site": head["roofline"]["b_site"],
        "copy_kernel_GBs": round(copy_gbs, 1),
        "frac_of_copy": round(head["roofline"]["achieved"] / copy_gbs, 4),
        "cpu_baseline": cpu,
        "extra": extra,
        "batched": batched,
        "serving": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
