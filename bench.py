"""Benchmark of the LBM time step on one MI355X.

Workload (default): D3Q19 MRT + Smagorinsky LES (Cs 0.17, tau0 0.505)
flow past a sphere, 256^3 FP32 SoA, sphere R = 24 at (64, 128, 128) with
BFL curved links, Zou-He velocity inlet u = 0.05 with seeded 1% turbulence,
convective outflow, periodic y/z.  One step = one fused pull-stream-collide
launch over the whole domain; MLUPS counts every node, as the reference does
(sphere_3d_wake.cpp:81-92).

By default the line also nests the metric's batched-serve half: a
`batched` record (the 512-domain batched wakes below at 2000 steps) and a
`serve` record (the request stream below at 0.2x its step counts), each
against its plain one-at-a-time path (`--sub-records none` skips them).

`--workload batched-wakes` runs the batched multi-domain path alone: 512
D2Q9 MRT cylinder-wake domains of uneven size, one launch per step over all
of them, and compares simulations per second with the plain
one-domain-at-a-time path.  `--workload serve` replays a Poisson stream of
mixed 3D bluff-body and 2D thermal-cavity requests through the
continuous-batching serving loop and then through the plain sequential
loop, and reports simulations per second and p50/p99 time-to-result.
`--workload slab` times one GPU's z-slab share of the 256^3 sphere workload.
`--workload wing` times the D3Q27 NACA-0012 wing channel (384x192x192)
with FP32/FP16/BF16 storage (tests/precision_wing.py states each
storage's error against the FP64 oracle on the same flow at reduced size).

Prints ONE JSON line.  The roofline entry divides the algorithmic bytes of
a step (one read and one write of every population: 152 B/node for D3Q19
FP32) by the average duration of the step kernel, measured with HIP events
on the domain's stream over the timed launches.  The CPU baseline times the
oracle's unfused restatement of the reference step on the host cores on a
short sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "D3Q19 MLUPS & %HBM peak; batched-serve vs plain sims/s + p99 at stated err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(n, storage):
    """HBM bytes per k_step launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (256^3, FP32 or FP16), else None."""
    names = {0: "k_step_d3q19_sphere256_pmc.md", 1: "k_step_d3q19_sphere256_fp16_pmc.md"}
    if n != 256 or storage not in names:
        return None
    path = os.path.join(ROOT, "profiles", names[storage])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def sphere_workload(grid_cls, n=256, R=24.0, storage=0, lattice=19, plane_pad=-1):
    g = grid_cls(n, n, n, 0.505, lattice=lattice, storage=storage, plane_pad=plane_pad)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.05, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(0.05, 0.0)
    g.set_inflow_turbulence(0.01, 2)
    g.set_convective_outflow_east(0.05)
    g.set_solid_sphere(64.0, 128.0, 128.0, R)
    nl = g.build_curved_sphere_links(64.0, 128.0, 128.0, R)
    return g, nl


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, steps, R=24.0, n_1t=128, steps_1t=10):
    """The oracle's unfused restatement of the reference step sequence on the
    host cores, built here for this host (-O3 -march=native with fused
    multiply-adds, as the reference's ISPC path is built with
    --opt=fast-math), timed on every thread on the n^3 sphere workload and on
    one thread on an n_1t^3 one.  Falls back to the prebuilt x86-64-v3 FMA
    build if the host build fails."""
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    out_dir = tempfile.mkdtemp(prefix="wake_oracle_native_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_DIR={out_dir}"],
                       capture_output=True, text=True)
    if r.returncode == 0:
        key, build = "32native", "gcc -O3 -march=native -ffp-contract=fast -fopenmp (built on this host)"
        L = ob.lib(key, path=os.path.join(out_dir, "liboracle32native.so"))
    else:
        log(f"host-native oracle build failed, timing the x86-64-v3 FMA build: {r.stderr.strip()[:200]}")
        key, build = "32fma", "gcc -O3 -march=x86-64-v3 -ffp-contract=fast -fopenmp (prebuilt)"
        L = ob.lib(key)
    all_threads = L.oracle_threads()

    def timed(m, k, threads):
        L.oracle_set_threads(threads)
        o = ob.OracleGrid(3, m, m, m, tau=0.505, bits=key)
        o.set_periodic_axes(True, True)
        o.initialize(1.0, 0.05, 0.0, 0.0)
        o.set_les_coupled(True)
        o.set_inflow_west(0.05, 0.0)
        o.set_inflow_turbulence(0.01, 2)
        o.set_convective_outflow_east(0.05)
        c = 64.0 * m / 256
        o.set_solid_sphere(c, m / 2, m / 2, R * m / 256)
        o.build_curved_sphere_links(c, m / 2, m / 2, R * m / 256)
        o.step("MRT", use_les=True, smag=0.17)  # warm-up
        t0 = time.perf_counter()
        o.run(k, "MRT", use_les=True, smag=0.17)
        dt = time.perf_counter() - t0
        return m ** 3 * k / dt / 1e6, dt

    v_all, dt_all = timed(n, steps, all_threads)
    v_one, dt_one = timed(n_1t, steps_1t, 1)
    L.oracle_set_threads(all_threads)
    return {"value": round(v_all, 2), "unit": "MLUPS", "cores": all_threads, "kind": "port",
            "sample": f"oracle (unfused reference step sequence, FP32) on the {n}^3 sphere workload, {steps} timed "
                      f"steps after 1 warm-up on {all_threads} threads ({dt_all:.1f} s); one thread: {n_1t}^3, "
                      f"{steps_1t} steps ({dt_one:.1f} s)",
            "single_thread_mlups": round(v_one, 3), "build": build, "host": _cpu_model()}


def batched_wakes(wa, ndom, seed=4, umax=0.05, storage=0):
    """512 (ndom) independent D2Q9 MRT channel flows past a cylinder: NX =
    128 + 32 U{0..28}, NY = NX/4, D = NY/8 at (NX/5, NY/2 + U{-2..2}), solid
    walls at y = 0 and NY - 1, parabolic Zou-He inlet with peak umax, Zou-He
    pressure outlet, Re log-uniform in [50, 300] on the mean inlet velocity
    U = 2/3 umax (tau = 3 U D / Re + 1/2), flow started from the inlet
    profile.  At D = 4 and Re = 300 the grid Reynolds number U / nu is 75:
    plain MRT diverges there, so the collision carries the coupled
    Smagorinsky closure (Cs = 0.17).  Returns the grids and their
    (cx, cy, D, U, Re) geometry."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nxs = 128 + 32 * rng.integers(0, 29, ndom)
    offs = rng.integers(-2, 3, ndom)
    res = np.exp(rng.uniform(math.log(50.0), math.log(300.0), ndom))
    U = 2.0 / 3.0 * umax
    grids, geo = [], []
    for nx, off, re in zip(nxs, offs, res):
        nx = int(nx)
        ny = nx // 4
        D = ny / 8.0
        g = wa.LatticeGrid2D(nx, ny, 3.0 * U * D / float(re) + 0.5, storage=storage)
        H = ny - 2
        s = np.arange(ny, dtype=np.float64) - 0.5
        prof = np.clip(4.0 * umax * s * (H - s) / (H * H), 0.0, None)
        g.initialize_fields(np.ones(nx * ny, np.float32), np.repeat(prof, nx).astype(np.float32),
                            np.zeros(nx * ny, np.float32))
        g.set_les_coupled(True)
        g.set_inflow_parabolic(umax)
        g.set_outflow_east(1.0)
        g.set_solid_box(0, 0, nx - 1, 0)
        g.set_solid_box(0, ny - 1, nx - 1, ny - 1)
        cx, cy = nx / 5.0, ny / 2.0 + float(off)
        g.set_solid_circle(cx, cy, D / 2)
        grids.append(g)
        geo.append((cx, cy, D, U, float(re)))
    return grids, geo


def batched_record(args):
    """Batched cylinder wakes: every domain stepped `steps` times by one
    launch per step (graph-captured chunks of 100, autotuned launch shape),
    force history per step, C_D and Strouhal per domain; the plain
    one-domain-at-a-time path (the single-domain step with the same per-step
    force, taken inside its step launch) is timed per domain on a short run
    and scaled to the same step count.  Returns the record as a dict."""
    import numpy as np

    import wake_amd
    from wake_amd import analysis

    storage = {"fp32": 0, "fp16": 1, "bf16": 2}[args.storage]
    ndom, S = args.domains, args.steps
    log(f"setup {ndom} D2Q9 cylinder-wake domains ({args.storage})")
    t0 = time.perf_counter()
    grids, geo = batched_wakes(wake_amd, ndom, storage=storage)
    nodes = sum(g.total_nodes() for g in grids)
    log(f"setup done in {time.perf_counter() - t0:.1f}s, {nodes / 1e6:.1f} M nodes")

    # plain path: each domain alone with its per-step force, wall time of a
    # short run scaled to S steps (small domains are launch-bound there)
    t_plain = 0.0
    probe = args.plain_probe
    for g, (cx, cy, D, U, re) in zip(grids, geo):
        g.set_force_probe(cx, cy, 0.0, 1.5 * D)
        g.run(2, "MRT", use_les=True, smag=0.17)
        g.synchronize()
        t0 = time.perf_counter()
        g.run(probe, "MRT", use_les=True, smag=0.17)
        g.force_history()
        t_plain += (time.perf_counter() - t0) / probe * S
        g.clear_force_probe()
    log(f"plain path with the per-step force: {t_plain:.2f} s projected for {S} steps of every domain "
        f"({probe}-step probe each)")

    b = wake_amd.Batch(lattice=9, storage=storage, collision="MRT", use_les=True, chunk=100)
    slots = [b.add(g, S - 2 - probe, smag=0.17, forces=(cx, cy, 0.0, 1.5 * D))
             for g, (cx, cy, D, U, re) in zip(grids, geo)]
    t0 = time.perf_counter()
    shape = b.tune(probe_steps=4)
    tune_s = time.perf_counter() - t0
    left = max(b.remaining(s) for s in slots)
    log(f"autotuned launch shape block={shape[0]} grid_factor={shape[1]} in {tune_s:.2f}s; {left} steps left")
    t0 = time.perf_counter()
    ms = b.run(left, use_graph=not getattr(args, "no_graph", False))
    wall = time.perf_counter() - t0
    mlups = nodes * left / wall / 1e6
    log(f"batched: {left} steps in {wall:.2f} s wall ({ms / 1e3:.2f} s device), {mlups:.0f} MLUPS")
    # the same S steps per domain as the plain projection
    t_batch = wall * S / left
    cds, sts, bad = [], [], 0
    for s, (cx, cy, D, U, re) in zip(slots, geo):
        F = b.forces(s)
        tail = F[len(F) // 2:]
        q = 0.5 * U * U * D
        if not np.all(np.isfinite(tail)):
            bad += 1
            continue
        cds.append(float(np.mean(tail[:, 0])) / q)
        f = analysis.analyze_frequency((F[-4096:, 1] / q).astype(np.float32), 1.0)
        sts.append(analysis.strouhal_number(f, D, U))
    from wake_amd.hbm_probe import ceiling_d2q9

    pull9 = ceiling_d2q9()
    bpn9 = 72.0 if storage == 0 else 36.0
    gbps = bpn9 * nodes * left / (ms * 1e-3) / 1e9  # over the device time of the chunks (forces included)
    log(f"D2Q9 pull copy on this GPU: {pull9:.0f} GB/s (FP32); batched steps {gbps:.0f} GB/s over the device time")
    log(f"C_D median {np.median(cds):.3f} [{np.min(cds):.3f}, {np.max(cds):.3f}], "
        f"St median {np.median(sts):.3f}, non-finite domains {bad}; {ndom / t_batch:.3f} sims/s batched vs "
        f"{ndom / t_plain:.3f} plain ({t_plain / t_batch:.2f}x)")
    return {
        "metric": METRIC, "value": round(ndom / t_batch, 4), "unit": "sims/s", "n_gpus": 1, "steps": S,
        "warmup": 2 + probe, "ms_per_step": round(wall * 1e3 / left, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if storage == 0 else ("fp16-storage/fp32" if storage == 1 else "bf16-storage/fp32"),
        "data": "synthetic",
        "config": {
            "workload": f"{ndom} D2Q9 MRT+Smagorinsky(Cs=0.17) cylinder wakes (NX 128-1024 step 32, NY=NX/4, D=NY/8, "
                        f"parabolic Zou-He inlet umax=0.05, Re 50-300 log-uniform, seed 4), {S} steps each with the "
                        f"force of every step, one launch per step over all domains, graph chunks of 100 steps",
            "nodes": nodes, "launch_shape": {"block": shape[0], "grid_factor": shape[1]},
            "batched_mlups": round(mlups, 1), "plain_sims_per_s": round(ndom / t_plain, 4),
            "batched_gbps": round(gbps, 1), "pull_d2q9_gbps": round(pull9, 1),
            "frac_of_pull_same_width": round(gbps / pull9, 4) if storage == 0 else None,
            "speedup_vs_plain": round(t_plain / t_batch, 3),
            "plain": "single-domain step, one domain at a time, per-step force inside its step launch, wall time "
                     f"of a {probe}-step run per domain scaled to {S} steps",
            "error": "arithmetic identical to the plain path per node (bit-identical batch of one); members vs "
                     "the FP32 oracle within the FP32-floor criterion (tests/test_gpu_batch.py, test_gpu_configs.py)",
            "C_D_median": round(float(np.median(cds)), 4) if cds else None,
            "St_median": round(float(np.median(sts)), 4) if sts else None,
            "non_finite_domains": bad, "parallelism": "single GPU",
        },
    }


def main_batched(args):
    print(json.dumps(batched_record(args)), flush=True)


def serve_record(args):
    """Serving stream: Poisson arrivals of mixed requests under an HBM
    budget, served by the continuous-batching loop, then the same stream
    through the plain one-request-at-a-time FP32 loop.  Returns the record."""
    import numpy as np

    import wake_amd  # noqa: F401
    from wake_amd import serve

    reqs = serve.request_stream(args.requests, rate=args.rate, seed=5, step_scale=args.step_scale)
    work = sum((r.n ** 3 if r.kind == "bluff3d" else r.n ** 2) * r.steps for r in reqs)
    log(f"{len(reqs)} requests over {reqs[-1].arrival:.2f} s, {work / 1e9:.1f} G node-steps, "
        f"{sum(r.kind == 'bluff3d' for r in reqs)} bluff3d / {sum(r.kind == 'cavity2d' for r in reqs)} cavity2d, "
        f"{sum(r.tol >= 1e-3 for r in reqs)} with FP16 storage")
    s = serve.Server(budget_bytes=args.budget_gb * 1e9, chunk=100)
    t_srv = s.serve(reqs, realtime=True)
    lat = serve.latency_stats(s.done)
    log(f"served: {t_srv:.2f} s makespan, p50 {lat['p50_s']:.2f} s, p99 {lat['p99_s']:.2f} s, "
        f"peak HBM {s.peak_used / 1e9:.1f} GB, {s.ticks} ticks; host s waiting {s.stats['wait_s']:.2f}, "
        f"at boundaries {s.stats['boundary_s']:.2f}, building domains {s.stats['build_s']:.2f}")
    if args.results:
        serve.write_results_jsonl(s.done, args.results)
        log(f"per-request results written to {args.results}")
    plain, t_plain = serve.plain_loop(reqs, realtime=True)
    lat_p = serve.latency_stats(plain)
    log(f"plain:  {t_plain:.2f} s makespan, p50 {lat_p['p50_s']:.2f} s, p99 {lat_p['p99_s']:.2f} s; "
        f"{t_plain / t_srv:.2f}x")
    # agreement of served (FP16 where allowed) and plain FP32 results
    by = {r.rid: r for r in plain}
    dev = []
    for r in s.done:
        for k, v in r.result.items():
            ref = by[r.rid].result.get(k)
            if k == "C_L" or not isinstance(v, float) or not isinstance(ref, float):
                continue
            if not (math.isfinite(v) and math.isfinite(ref)):
                continue
            dev.append(abs(v / ref - 1) if ref != 0 else abs(v))
    return {
        "metric": METRIC, "value": round(len(reqs) / t_srv, 4), "unit": "sims/s", "n_gpus": 1,
        "steps": int(sum(r.steps for r in reqs)), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp16/fp32-storage, fp32 arithmetic", "data": "synthetic",
        "config": {
            "workload": f"serving stream: {len(reqs)} Poisson arrivals at {args.rate}/s (seed 5), 60% D3Q19 MRT-LES "
                        f"rotated-ellipsoid bluff bodies 64^3-192^3, 40% D2Q9+D2Q5 Boussinesq cavities 128^2-512^2, "
                        f"steps U[2k,20k] x {args.step_scale}, HBM budget {args.budget_gb} GB",
            "served_makespan_s": round(t_srv, 3), "plain_makespan_s": round(t_plain, 3),
            "plain_sims_per_s": round(len(reqs) / t_plain, 4), "speedup_vs_plain": round(t_plain / t_srv, 3),
            "p50_s": round(lat["p50_s"], 3), "p99_s": round(lat["p99_s"], 3),
            "plain_p50_s": round(lat_p["p50_s"], 3), "plain_p99_s": round(lat_p["p99_s"], 3),
            "aggregate_mlups": round(work / t_srv / 1e6, 1), "plain_mlups": round(work / t_plain / 1e6, 1),
            "peak_hbm_gb": round(s.peak_used / 1e9, 2),
            "diverged_requests": len(s.diverged),
            "host_seconds": {k: round(v, 3) for k, v in s.stats.items()},
            "result_rel_dev_vs_plain_fp32": {"median": float(np.median(dev)) if dev else None,
                                             "max": float(np.max(dev)) if dev else None},
            "parallelism": "single GPU",
        },
    }


def main_serve(args):
    print(json.dumps(serve_record(args)), flush=True)


def wing_setup(g, nx, ny, chord, u=0.04, aoa=8.0):
    """D3Q27 TRT + coupled LES channel: walls at y = 0, ny - 1 (halfway
    bounce-back at the domain edges), periodic z, Zou-He inlet u, convective
    outlet, NACA 0012 extruded over the full span, leading edge at
    (nx/4, ny/2), angle of attack `aoa`."""
    g.set_periodic_axes(False, True)
    g.initialize(1.0, u, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(u, 0.0)
    g.set_convective_outflow_east(u)
    g.set_solid_naca(12, chord, aoa, nx / 4.0, ny / 2.0)
    return g


def main_wing(args):
    """NACA 0012 wing at 8 deg, D3Q27 TRT + LES, Re 2000 on the chord: the
    full 384x192x192 domain timed with FP32, FP16 and BF16 population
    storage.  Each storage's error against the FP64 oracle on the same flow
    at reduced size comes from tests/precision_wing.py."""
    import wake_amd

    u, re = 0.04, 2000.0
    names = {0: "fp32", 1: "fp16", 2: "bf16"}
    perf = {}
    nx, ny, nz, chord = 384, 192, 192, 96.0
    tau = 3.0 * u * chord / re + 0.5
    for st in (0, 1, 2):
        g = wing_setup(wake_amd.LatticeGrid3D(nx, ny, nz, tau, lattice=27, storage=st), nx, ny, chord)
        g.run(10, "TRT", use_les=True, smag=0.17)
        ms = g.run(args.steps, "TRT", use_les=True, smag=0.17, timed=True)
        kernel_ms = g.kernel_time()
        bpn = 216 if st == 0 else 108
        perf[names[st]] = {"MLUPS": round(nx * ny * nz * args.steps / (ms * 1e-3) / 1e6, 1),
                           "kernel_us": round(kernel_ms * 1e3, 1),
                           "kernel_gbps": round(bpn * nx * ny * nz / (kernel_ms * 1e-3) / 1e9, 1)}
        log(f"full {nx}x{ny}x{nz} {names[st]}: {perf[names[st]]}")
        del g
    from wake_amd.hbm_probe import ceilings_d3q27
    ceil = ceilings_d3q27(nx, ny, nz)
    for nm, key in (("fp32", "pull_d3q27_gbps"), ("fp16", "pull_d3q27_16bit_gbps"), ("bf16", "pull_d3q27_16bit_gbps")):
        perf[nm]["frac_of_pull_same_width"] = round(perf[nm]["kernel_gbps"] / ceil[key], 4)
    log(f"D3Q27 pull copy on this GPU: {ceil['pull_d3q27_gbps']:.0f} GB/s (FP32), "
        f"{ceil['pull_d3q27_16bit_gbps']:.0f} GB/s (16-bit)")
    out = {
        "metric": METRIC, "value": perf["fp32"]["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": 10, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (and fp16/bf16 storage)", "data": "synthetic",
        "config": {"workload": f"D3Q27 TRT(Lambda=1/4)+LES channel, NACA 0012 extruded at 8 deg, chord {chord:g}, "
                               f"{nx}x{ny}x{nz}, inlet u=0.04, Re=2000",
                   "storage_perf": perf,
                   "bytes_per_node": {"fp32": 216, "fp16": 108, "bf16": 108},
                   "pull_ceiling_gbps": {k: round(v, 1) for k, v in ceil.items()},
                   "error_study": "tests/precision_wing.py (FP64 oracle at 1/4 size)",
                   "parallelism": "single GPU"},
    }
    print(json.dumps(out), flush=True)


def main_tg2d(args):
    """D2Q9 BGK Taylor-Green vortex, 256^2 fully periodic, tau 0.8, u0 0.04,
    equilibrium plus seeded uniform noise of 1e-6, 1000 steps: the domain's
    run() (8 steps per launch from an LDS tile, as a fully periodic domain of
    this size takes), the same with one launch per step, and the domain as a
    batch of one (graph-captured chunks of 100 launches), against the oracle
    on the host.  The working set (2 x 2.36 MB) stays in the on-chip caches,
    so this is not an HBM-roofline figure."""
    import numpy as np

    import wake_amd

    n, tau, u0, S = 256, 0.8, 0.04, args.steps
    k = 2 * math.pi / n
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ux = (u0 * np.sin(k * x) * np.cos(k * y)).astype(np.float32).ravel()
    uy = (-u0 * np.cos(k * x) * np.sin(k * y)).astype(np.float32).ravel()
    rho = np.ones(n * n, np.float32)

    def grid():
        g = wake_amd.LatticeGrid2D(n, n, tau)
        g.set_periodic(True)
        g.initialize_fields(rho, ux, uy)
        g.add_noise(1e-6, 1)
        return g

    def timed_run(g):
        g.run(16, "BGK")
        g.synchronize()
        t0 = time.perf_counter()
        g.run(S, "BGK")
        g.synchronize()
        return time.perf_counter() - t0

    s1 = grid()
    s1.set_multistep(False)
    wall_1 = timed_run(s1)
    mlups_1 = n * n * S / wall_1 / 1e6
    del s1
    g = grid()
    e0 = g.compute_total_kinetic_energy()
    wall = timed_run(g)
    mlups = n * n * S / wall / 1e6
    nu = (tau - 0.5) / 3.0
    ratio = g.compute_total_kinetic_energy() / e0
    expect = math.exp(-4 * nu * k * k * (S + 16))
    h = grid()
    b = wake_amd.Batch(lattice=9, collision="BGK", chunk=100)
    b.add(h, S + 10)
    b.run(10)
    t0 = time.perf_counter()
    b.run(S)
    wall_b = time.perf_counter() - t0
    mlups_b = n * n * S / wall_b / 1e6
    log(f"TG 256^2 BGK: run() {mlups:.0f} MLUPS ({wall / S * 1e6:.2f} us/step), one launch per step {mlups_1:.0f} "
        f"MLUPS ({wall_1 / S * 1e6:.2f} us/step), batch of one (graph) "
        f"{mlups_b:.0f} MLUPS ({wall_b / S * 1e6:.2f} us/step); KE ratio {ratio:.5f} vs exp(-4 nu k^2 t) {expect:.5f}")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    o = ob.OracleGrid(2, n, n, tau=tau)
    o.set_periodic(True)
    o.initialize_fields(rho, ux, uy)
    o.add_noise(1e-6, 1)
    o.run(16, "BGK")
    t0 = time.perf_counter()
    o.run(S, "BGK")
    cpu = n * n * S / (time.perf_counter() - t0) / 1e6
    err = float(np.linalg.norm(np.concatenate([g.ux_data(), g.uy_data()]) - np.concatenate([o.ux(), o.uy()]))
                / np.linalg.norm(np.concatenate([o.ux(), o.uy()])))
    log(f"oracle on the host: {cpu:.0f} MLUPS on {ob.lib(32).oracle_threads()} threads; GPU vs oracle u rel L2 {err:.2e}")
    best, best_wall, path = max((mlups, wall, "run_8_steps_per_launch"), (mlups_1, wall_1, "one_launch_per_step"),
                                (mlups_b, wall_b, "batch_of_one_graph"))
    out = {
        "metric": METRIC, "value": round(best, 1), "unit": "MLUPS", "n_gpus": 1, "steps": S, "warmup": 16,
        "ms_per_step": round(best_wall * 1e3 / S, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "D2Q9 BGK Taylor-Green vortex 256^2 periodic, tau 0.8, u0 0.04, noise 1e-6 (seed 1)",
                   "value_path": path,
                   "run_mlups": round(mlups, 1), "one_launch_per_step_mlups": round(mlups_1, 1),
                   "batch_of_one_graph_mlups": round(mlups_b, 1),
                   "on_chip": "2 x 2.36 MB working set: cache-resident, not an HBM-roofline figure",
                   "ke_ratio": round(ratio, 6), "ke_ratio_analytic": round(expect, 6),
                   "u_rel_l2_vs_oracle": err, "cpu_oracle_mlups": round(cpu, 1),
                   "parallelism": "single GPU"},
    }
    print(json.dumps(out), flush=True)


def main_slab(args):
    """One GPU's share of the 256^3 sphere workload split into 8 z-slabs:
    256 x 256 x 32 owned planes plus 2 halo planes (z 96..127 and its
    neighbours, the slab that holds the sphere's lower half), stepped by the
    same fused kernel.  Periodic z stands in for the halo exchange (a halo
    plane is pulled like an interior plane).  Reported beside it: the bytes
    each face would send per step (the 5 populations with c_z = +-1 of one
    256 x 256 plane) and their time over one xGMI link at 153 GB/s, an
    estimate: nothing here uses a second GPU."""
    import wake_amd

    storage = {"fp32": 0, "fp16": 1, "bf16": 2}[args.storage]
    n, owned, z0, R = 256, 32, 96, 24.0
    nz = owned + 2
    cz = 128.0 - z0 + 1.0  # the sphere centre in slab coordinates (plane 0 is the lower halo)
    g = wake_amd.LatticeGrid3D(n, n, nz, 0.505, storage=storage)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.05, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(0.05, 0.0)
    g.set_inflow_turbulence(0.01, 2)
    g.set_convective_outflow_east(0.05)
    g.set_solid_sphere(64.0, 128.0, cz, R)
    nl = g.build_curved_sphere_links(64.0, 128.0, cz, R)
    g.run(args.warmup, "MRT", use_les=True, smag=0.17)
    g.synchronize()
    ms = g.run(args.steps, "MRT", use_les=True, smag=0.17, timed=True)
    kernel_ms = g.kernel_time()
    owned_nodes = n * n * owned
    elem = 4 if storage == 0 else 2
    face = 5 * n * n * elem
    mlups = owned_nodes * args.steps / (ms * 1e-3) / 1e6
    log(f"slab 256x256x{owned}+2: {ms / args.steps * 1e3:.1f} us/step (kernel {kernel_ms * 1e3:.1f} us), "
        f"{mlups:.0f} MLUPS on the owned nodes; face {face / 1e6:.3f} MB each way, "
        f"{face / 153e9 * 1e6:.1f} us per link (estimate)")
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.storage, "data": "synthetic",
        "config": {"workload": f"one GPU's z-slab of the 256^3 D3Q19 MRT+LES sphere workload split 8 ways: "
                               f"256x256x{owned} owned + 2 halo planes (z {z0}..{z0 + owned - 1})",
                   "owned_nodes": owned_nodes, "bfl_links": nl, "kernel_ms": round(kernel_ms, 5),
                   "face_bytes_per_direction": face,
                   "xgmi_us_per_face_estimate": round(face / 153e9 * 1e6, 2),
                   "parallelism": "single GPU (one slab)"},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["sphere", "batched-wakes", "serve", "wing", "slab", "tg2d"],
                    default="sphere")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", choices=["fp32", "fp16", "bf16"], default="fp32")
    ap.add_argument("--plane-pad", type=int, default=-1)
    ap.add_argument("--block-x", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=30,
                    help="oracle steps timed for the CPU baseline, about 10 s on 16 host threads (0: skip)")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--domains", type=int, default=512)
    ap.add_argument("--plain-probe", type=int, default=20)
    ap.add_argument("--no-graph", action="store_true", help="batched-wakes: launch each step directly (no HIP graph)")
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--step-scale", type=float, default=1.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--results", default=None, help="serve: write per-request results (JSON lines) here")
    ap.add_argument("--sub-records", default="batched,serve",
                    help="sphere workload: also run these (comma-separated: batched, serve; 'none') and nest their "
                         "records in the line -- the metric's batched-serve half")
    ap.add_argument("--batched-steps", type=int, default=2000, help="steps per domain of the nested batched record")
    ap.add_argument("--serve-step-scale", type=float, default=0.2,
                    help="step scale of the nested serve record's request stream")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "wing":
        if args.steps is None:
            args.steps = 500
        main_wing(args)
        return
    if args.workload == "serve":
        main_serve(args)
        return
    if args.workload == "tg2d":
        if args.steps is None:
            args.steps = 1000
        main_tg2d(args)
        return
    if args.workload == "slab":
        if args.steps is None:
            args.steps = 2000
        main_slab(args)
        return
    if args.workload == "batched-wakes":
        if args.steps is None:
            args.steps = 20000
        main_batched(args)
        return
    if args.steps is None:
        args.steps = 2000

    import numpy as np

    import wake_amd

    storage = {"fp32": 0, "fp16": 1, "bf16": 2}[args.storage]
    bpn = {0: 152, 1: 76, 2: 76}[storage]
    n = args.n
    N = n ** 3
    log(f"setup {n}^3 sphere workload ({args.storage})")
    t0 = time.perf_counter()
    g, nlinks = sphere_workload(wake_amd.LatticeGrid3D, n, 24.0 * n / 256, storage, plane_pad=args.plane_pad)
    if args.block_x:
        g.set_block_x(args.block_x)
    log(f"setup done in {time.perf_counter() - t0:.1f}s, {nlinks} BFL links, device bytes {g.device_bytes() / 1e9:.2f} GB")
    g.run(args.warmup, "MRT", use_les=True, smag=0.17)
    g.synchronize()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    ms = g.run(args.steps, "MRT", use_les=True, smag=0.17, timed=True)
    g.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = g.kernel_time()
    log(f"timed {args.steps} steps: events {ms:.1f} ms, wall {wall * 1e3:.1f} ms, "
        f"step kernel {kernel_ms * 1e3:.1f} us/launch")
    mlups_off = N * args.steps / wall / 1e6
    R = 24.0 * n / 256
    centre = (64.0 * n / 256, n / 2, n / 2)
    q = 0.5 * 0.05 ** 2 * math.pi * R * R
    # The same steps with the momentum exchange of every step (the
    # reference's scenario takes C_D / C_L after each step): computed inside
    # the step launch, folded on the device, read once at the end.
    g.set_force_probe(*centre, 2 * R)
    g.synchronize()
    t0 = time.perf_counter()
    ms_f = g.run(args.steps, "MRT", use_les=True, smag=0.17, timed=True)
    first, hist = g.force_history()
    wall_f = time.perf_counter() - t0
    kernel_ms_f = g.kernel_time()
    mlups = N * args.steps / wall_f / 1e6
    log(f"timed {args.steps} steps with per-step forces: events {ms_f:.1f} ms, wall {wall_f * 1e3:.1f} ms "
        f"(history read included), step kernel {kernel_ms_f * 1e3:.1f} us/launch; "
        f"{mlups:.0f} MLUPS vs {mlups_off:.0f} without forces ({mlups / mlups_off:.4f})")
    achieved = bpn * N / (kernel_ms_f * 1e-3) / 1e9
    # the per-step force history against one direct call on the final state
    F = g.momentum_exchange(*centre, 2 * R)
    cd, cl = F[0] / q, F[1] / q
    assert hist.shape[0] == args.steps + 1 and np.array_equal(hist[-1], F), "force history != direct call"
    cd_hist = hist[:, 0] / q
    ok = bool(np.all(np.isfinite(hist))) and math.isfinite(g.compute_total_mass())
    log(f"C_D {cd:.4f} C_L {cl:.5f}; C_D over the timed steps mean {cd_hist[1:].mean():.4f} "
        f"[{cd_hist[1:].min():.4f}, {cd_hist[1:].max():.4f}] finite={ok}")

    from wake_amd.hbm_probe import ceilings
    ceil = ceilings(n, n, n)
    pull_key = "pull_d3q19_gbps" if storage == 0 else "pull_d3q19_16bit_gbps"
    log(f"HBM ceilings on this GPU: contiguous copy {ceil['copy_contiguous_gbps']:.0f} GB/s, "
        f"D3Q19 pull copy {ceil['pull_d3q19_gbps']:.0f} GB/s (16-bit {ceil['pull_d3q19_16bit_gbps']:.0f} GB/s); "
        f"step kernel {achieved:.0f} GB/s")

    cpu = None
    if args.cpu_steps > 0:
        log(f"CPU baseline: oracle {args.cpu_n}^3, {args.cpu_steps} steps")
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps, 24.0)
        log(f"CPU baseline {cpu['value']:.1f} MLUPS on {cpu['cores']} threads, {cpu['single_thread_mlups']:.2f} on one "
            f"({cpu['host']}, {cpu['build']})")

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_f * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if storage == 0 else ("fp16-storage/fp32" if storage == 1 else "bf16-storage/fp32"),
        "data": "synthetic",
        "config": {
            "workload": f"D3Q19 MRT + coupled Smagorinsky LES flow past a sphere, {n}^3, R={R:g} at "
                        f"({64 * n // 256},{n // 2},{n // 2}), BFL links, Zou-He inlet u=0.05 + 1% seeded "
                        f"turbulence, convective outflow, periodic y/z",
            "lattice": "D3Q19", "collision": "MRT+LES(Cs=0.17,tau0=0.505)", "nodes": N,
            "bfl_links": nlinks, "storage": args.storage, "C_D": round(cd, 5), "C_L": round(cl, 6),
            "forces": "C_D/C_L of every step by FP64 momentum exchange inside the step launch (timed)",
            "mlups_without_forces": round(mlups_off, 1),
            "forces_on_vs_off": round(mlups / mlups_off, 4),
            "kernel_ms_without_forces": round(kernel_ms, 5),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n, storage),
            "traffic_source": (None if measured_traffic(n, storage) is None else
                               ("profiles/" + ("k_step_d3q19_sphere256_pmc.md" if storage == 0 else
                                               "k_step_d3q19_sphere256_fp16_pmc.md") +
                                " (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-calibrated)")),
            "bytes_per_node": bpn,
            "kernel": "k_step (fused pull + MRT/LES collide + per-step momentum exchange), HIP events around "
                      "every 16th launch of the timed steps",
            "kernel_ms": round(kernel_ms_f, 5),
            "step_ms_events": round(ms_f / args.steps, 5),
            "practical_ceiling": {
                "copy_contiguous_gbps": round(ceil["copy_contiguous_gbps"], 1),
                "pull_d3q19_gbps": round(ceil["pull_d3q19_gbps"], 1),
                "pull_d3q19_16bit_gbps": round(ceil["pull_d3q19_16bit_gbps"], 1),
                "frac_of_pull_same_width": round(achieved / ceil[pull_key], 4),
                "source": f"wake_hip_hbm_probe on this GPU, {n}^3 nodes, 20 launches each (152 B/node; "
                          f"76 B/node for the 16-bit pull)",
            },
        },
        "cpu_baseline": cpu,
    }
    del g
    subs = [x for x in args.sub_records.split(",") if x and x != "none"]
    if "batched" in subs:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.storage = args.batched_steps, "fp32"
        out["batched"] = batched_record(sub)
    if "serve" in subs:
        sub = argparse.Namespace(**vars(args))
        sub.step_scale = args.serve_step_scale
        out["serve"] = serve_record(sub)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
