"""Throughput of the nuosc RK4 time-stepper core on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision 64|32]
                    [--nn 8] [--zmax 12.8] [--dz 0.05] [--no-cpu]

Workload (default): the 512 x 512 periodic box with 642 icosahedral
directions (--zmax 12.8 --dz 0.05 --nv 8), FP64, mu=1, ko=0.1, cfl=0.4,
alpha=0.9, sigma_nu=0.6, sigma_nubar=0.5, seeded eps=1e-6 complex
perturbation.  A step is one full RK4 step (four stage launches, replayed
from a hipGraph) over the whole state, which stays resident in HBM.

Prints one JSON line:
  value          site*dir RK4 updates/s = nx*nz*nv*K / wall(K steps)
  roofline       for the dominant kernel (RK4 stage 2): algorithmic bytes per
                 launch (4 state passes x 8 fields x sizeof(real) per update)
                 / its average duration from HIP events, against 8 TB/s
  step_roofline  the same for a whole step (12 passes: 768 B per update in
                 FP64), plus vs_16pass_ceiling = updates/s over 8 TB/s / 1024 B
                 (FP64; 512 B FP32), the ceiling of the 16-pass RK4 streaming
                 in which the per-update targets are stated
  fp32           (default box run) the same workload in the FP32 form: value,
                 speed-up over FP64, its stage-2 roofline
  batched        (default box run) the 64-run batched scan (FP32), 20 steps:
                 aggregate updates/s together and one after another, the
                 speed-up, and each mode's HBM fraction at 384 B per update
  served         (default box run) the run server, FP32: 20 s of Poisson
                 0.35/s arrivals of runs drawn like the batched scan's (grids
                 64..384 squared, steps U[100,1000]): p50/p99 request latency
                 and queueing delay, aggregate updates/s
  cpu_baseline   the CPU oracle on a 128 x 128 x 642 crop of the same physics,
                 timed on this host's cores, median of 3 samples (a reported
                 baseline), with its reference-equivalent rate

--shape column runs the 1-D open column (1 x 4096 cells x 642 directions,
gz=3, an ipt-0 pulse entering through z0) instead of the box.
--nflavor 3 runs the same box with 3 flavours (18 fields, pmo=1): the same
12 passes per step, of 18 fields; the roofline line is the stage-2 kernel
(4 passes of 18 fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# state passes per update of each stage (read + write of the 8 fields):
# S1 u -> Y2; S2 Y2, u -> Y3, A; S3 Y3, u -> Y4; S4 Y4, A -> u'
STAGE_PASSES = (2, 4, 3, 3)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", type=int, default=64, choices=(64, 32))
    p.add_argument("--nn", type=int, default=8)
    p.add_argument("--zmax", type=float, default=12.8)
    p.add_argument("--dz", type=float, default=0.05)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-extras", action="store_true",
                   help="box FP64: skip the FP32 and batched legs of the default line")
    p.add_argument("--no-autotune", action="store_true", help="box: keep the default z-segment plan")
    p.add_argument("--zlen", type=int, default=0,
                   help="box, 2 flavours: this z-segment length instead of the autotuned one (reproducible profiles)")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--nflavor", type=int, default=2, choices=(2, 3), help="box: 2 (8 fields) or 3 (18 fields)")
    p.add_argument("--workload", default="box", choices=("box", "batched", "served", "precision"),
                   help="box: one large run (default); batched: 64 uneven runs stepped together vs one by one; "
                        "served: Poisson arrivals into the run server; precision: the FP32 form against FP64 "
                        "from the same initial state (--shape box|column)")
    p.add_argument("--shape", default="box", choices=("box", "column"),
                   help="box and precision workloads: the 512x512x642 periodic box (seeded eps=1e-6) or the 1-D "
                        "open column (4096 cells in z, 642 directions, gz=3, an ipt-0 pulse entering through z0)")
    p.add_argument("--runs", type=int, default=64)
    p.add_argument("--grids", default=None,
                   help="batched/served: comma-separated square grid sizes the runs are drawn from "
                        "(default 64,128,192,256 batched; 64,...,384 served)")
    p.add_argument("--rate", type=float, default=2.0, help="served: arrivals per second")
    p.add_argument("--duration", type=float, default=60.0, help="served: seconds of arrivals")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain 1 GiB -> 1 GiB device copy (for scaling rocprofv3 byte counters)")
    return p.parse_args()


def copy_calibration():
    """A plain device copy of known size: reads 1 GiB and writes 1 GiB per
    launch (torch elementwise kernel), three launches.  Its rocprofv3
    FETCH_SIZE / WRITE_SIZE readings calibrate those counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"read_bytes": 4 * n, "write_bytes": 4 * n, "launches": 3}}), flush=True)


TRAFFIC_FILE = "profiles/stage_traffic.json"


def traffic_from_profiles(precision):
    """HBM bytes per stage-2 launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration), if recorded."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(f"fp{precision}", {}).get("stage2_bytes")
    except Exception:
        return None


def scan_run(rng, k, grids=(64, 128, 192, 256), dz=0.05):
    """One parameter-scan run: grid and direction count drawn from the scan
    sets, mu ~ U[0.5,2], alpha ~ U[0.7,1], sigma_nu ~ U[0.4,0.8],
    sigma_nubar = sigma_nu - 0.1, eps = 1e-6, own seed."""
    n = int(rng.choice(grids))
    nn = int(rng.choice((4, 8)))
    half = n * dz / 2
    sig = float(rng.uniform(0.4, 0.8))
    geo = dict(nv=nn, x0=-half, x1=half, z0=-half, z1=half, dx=dz, dz=dz, CFL=0.4, ko=0.1,
               mu=float(rng.uniform(0.5, 2.0)), pmo=0.0)
    init = dict(ipt=5, alpha=float(rng.uniform(0.7, 1.0)), eps0=1e-6, sigma=1.0, lnue=sig, lnueb=sig - 0.1,
                seed=1000 + k)
    return geo, init


def bench_batched(a):
    """The same runs stepped one after another (each replaying its own graph)
    and all together in one batch (one launch per stage for all runs)."""
    import numpy as np

    import nuosc_amd as N
    prec = N.PREC_FP64 if a.precision == 64 else N.PREC_FP32
    rng = np.random.default_rng(7)
    grids = tuple(int(x) for x in (a.grids or "64,128,192,256").split(","))
    specs = [scan_run(rng, k, grids=grids) for k in range(a.runs)]
    sims = []
    for geo, init in specs:
        g = N.NuOsc(device=0, precision=prec, use_graph=True, **geo)
        g.fillInitValue(**init)
        sims.append(g)
    upd = sum(g.nx * g.nz * g.nv for g in sims)
    # one by one, each run with its own autotuned z-segment length
    for g in sims:
        g.autotune(trials=2)
        g.step_rk4(1)
        g.synchronize()
    t0 = time.perf_counter()
    for g in sims:
        g.step_rk4(a.steps)
        g.synchronize()
    t_seq = time.perf_counter() - t0
    # together
    b = N.Batch(device=0, precision=prec)
    for g in sims:
        b.add(g)
    zbest, _ = b.autotune(trials=2)
    chosen = b.launch
    t_form = {}
    for lf in (chosen, 1 - chosen):   # the autotuned pair, then the other launch form
        b.set_launch(lf)
        b.step(1)
        t0 = time.perf_counter()
        b.step(a.steps)
        t_form[lf] = time.perf_counter() - t0
    t_bat = t_form[chosen]
    ok = all(g.analysis().status == 0 for g in sims)
    out = {
        "metric": "site*dir RK4 updates/s (batched scan)",
        "value": upd * a.steps / t_bat,
        "unit": "updates/s",
        "sequential_value": upd * a.steps / t_seq,
        "speedup_vs_sequential": t_seq / t_bat,
        "n_gpus": 1, "steps": a.steps, "runs": a.runs,
        "dtype": "f64" if a.precision == 64 else "f32",
        "config": {"workload": f"{a.runs} runs, grids drawn from {list(grids)} squared, 162/642 directions, "
                               f"uneven physics", "updates_per_step": upd, "batch_zlen": zbest,
                   "batch_launch": "persistent" if chosen == N.BATCH_PERSISTENT else "grid"},
        "grid_value": upd * a.steps / t_form[N.BATCH_GRID],
        "persistent_value": upd * a.steps / t_form[N.BATCH_PERSISTENT],
        "state_ok": ok,
    }
    b.close()
    for g in sims:
        g.close()
    print(json.dumps(out), flush=True)


def bench_served(a):
    """Poisson arrivals into the run server; latency percentiles."""
    print(json.dumps(served_leg(a.rate, a.duration, a.steps, a.precision, a.grids)), flush=True)


def served_leg(rate, duration, steps, precision, grids=None):
    """Poisson arrivals (rate per second, for duration seconds) of runs with
    grids drawn from `grids` squared, 162/642 directions and U[steps,
    10 steps] RK4 steps into the run server; returns request latency and
    queueing percentiles and the aggregate rate."""
    import numpy as np

    import nuosc_amd as N
    prec = N.PREC_FP64 if precision == 64 else N.PREC_FP32
    rng = np.random.default_rng(11)
    grids = tuple(int(x) for x in (grids or "64,128,192,256,320,384").split(","))
    srv = N.Server(device=0, precision=prec, hbm_budget_bytes=200 << 30, max_tick_steps=10)
    t_start = time.perf_counter()
    n, k = 0, 0
    next_t = rng.exponential(1.0 / rate)
    submitted = {}
    last_note = 0.0
    while next_t < duration:
        now = time.perf_counter() - t_start
        if now - last_note > 30:
            print(f"served: {k} requests submitted at {now:.1f} s", file=sys.stderr, flush=True)
            last_note = now
        if now < next_t:
            time.sleep(min(0.05, next_t - now))
            continue
        geo, init = scan_run(rng, k, grids=grids)
        nsteps = int(rng.integers(steps, 10 * steps + 1))
        rid = srv.submit(nsteps=nsteps, ana_every=max(1, nsteps // 20), **geo, **init)
        submitted[rid] = nsteps
        k += 1
        next_t += rng.exponential(1.0 / rate)
    print(f"served: {len(submitted)} requests submitted in {time.perf_counter() - t_start:.1f} s", file=sys.stderr,
          flush=True)
    lat, queue_wait, bad = [], [], 0
    for n_done in range(1, len(submitted) + 1):
        r = srv.take(timeout=3600)
        info, series = r
        bad += info.status != 0
        lat.append(info.t_done - info.t_submit)
        queue_wait.append(info.t_admit - info.t_submit)
        if n_done % 10 == 0:
            print(f"served: {n_done}/{len(submitted)} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr,
                  flush=True)
    st = srv.stats()
    srv.close()
    wall = time.perf_counter() - t_start
    out = {
        "metric": "served scan: request latency and aggregate updates/s",
        "value": st.updates / wall,
        "unit": "updates/s",
        "requests": len(submitted), "failed": bad, "rate_per_s": rate, "arrival_window_s": duration,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "queue_p50_s": float(np.percentile(queue_wait, 50)), "queue_p99_s": float(np.percentile(queue_wait, 99)),
        "gpu_busy_frac": st.busy_s / wall,
        "dtype": "f64" if precision == 64 else "f32",
        "config": {"workload": f"Poisson {rate}/s for {duration}s, grids drawn from {list(grids)} squared, "
                               f"162/642 directions, "
                               f"steps U[{steps},{10 * steps}], 200 GB budget"},
    }
    return out


def bench_precision(a):
    """The FP32 form against FP64 from the same initial state: every ANA_EVERY
    = 10 steps the Pee and |rho_ex| series of both, at the end max|dfield| per
    field, and the rates of change of log|rho_ex| of both."""
    import numpy as np

    import nuosc_amd as N
    if a.shape == "box":
        half = a.zmax
        geo = dict(nv=a.nn, x0=-half, x1=half, z0=-half, z1=half, dx=a.dz, dz=a.dz, CFL=0.4, ko=0.1, mu=1.0,
                   pmo=0.0)
        init = dict(ipt=5, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5, seed=12345)
        label = "512x512 periodic box x 642 directions, seeded eps=1e-6"
    else:
        geo = dict(nv=a.nn, x0=-0.025, x1=0.025, z0=-102.4, z1=102.4, dx=0.05, dz=0.05, CFL=0.4, ko=0.1, mu=1.0,
                   pmo=0.0, gz=3, bc=N.BC_OPEN_Z, zc=-105.4)
        init = dict(ipt=0, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5)
        label = "1-D open column, 4096 cells x 642 directions, gz=3, ipt-0 pulse entering through z0"
    sims = {}
    for prec in (64, 32):
        g = N.NuOsc(device=0, precision=N.PREC_FP64 if prec == 64 else N.PREC_FP32, use_graph=True, **geo)
        g.fillInitValue(**init)
        sims[prec] = g
    ana_every = 10
    series = {64: [], 32: []}
    for step in range(0, a.steps + 1, ana_every):
        for prec, g in sims.items():
            if step:
                g.step_rk4(ana_every)
            r = g.analysis()
            series[prec].append((r.phy_time, r.surv, r.maxdP, np.sqrt(max(r.rho_ex2, 0.0))))
        if step % 100 == 0:
            print(f"precision: step {step}/{a.steps}", file=sys.stderr, flush=True)
    s64, s32 = np.array(series[64]), np.array(series[32])
    dfield = []
    for f in range(8):
        f64, f32 = sims[64].get_field(f), sims[32].get_field(f)
        dfield.append(float(np.max(np.abs(f32 - f64)) / max(np.max(np.abs(f64)), 1e-300)))
        del f64, f32
    t, r64, r32 = s64[:, 0], s64[:, 3], s32[:, 3]
    # mean rate of change of log|rho_ex| over the run, and the fitted rate
    # over the longest stretch in which the FP64 amplitude keeps growing by
    # at least 1% per sample (a linear-growth window), if there is one
    mean64 = float((np.log(r64[-1]) - np.log(r64[0])) / (t[-1] - t[0]))
    mean32 = float((np.log(r32[-1]) - np.log(r32[0])) / (t[-1] - t[0]))
    grow = np.concatenate([[False], r64[1:] > 1.01 * r64[:-1]])
    best, cur, start = (0, 0), 0, 0
    for k, gk in enumerate(grow):
        cur = cur + 1 if gk else 0
        if cur == 1:
            start = k - 1
        if cur > best[1] - best[0]:
            best = (start, k + 1)
    growth = {"mean_rate_fp64": mean64, "mean_rate_fp32": mean32,
              "mean_rate_rel_error": abs(mean32 - mean64) / max(abs(mean64), 1e-300)}
    if best[1] - best[0] >= 6:
        sl = slice(best[0], best[1])
        g64 = float(np.polyfit(t[sl], np.log(r64[sl]), 1)[0])
        g32 = float(np.polyfit(t[sl], np.log(r32[sl]), 1)[0])
        growth.update({"window_rate_fp64": g64, "window_rate_fp32": g32,
                       "window_rate_rel_error": abs(g32 - g64) / abs(g64),
                       "window_t": [float(t[sl][0]), float(t[sl][-1])]})
    for g in sims.values():
        g.close()
    out = {
        "metric": "FP32 form vs FP64, same initial state",
        "steps": a.steps,
        "config": {"workload": label},
        "max_rel_dfield": dfield,
        "pee_max_abs_error": float(np.max(np.abs(s32[:, 1] - s64[:, 1]))),
        "max_abs_P_minus_1": {"fp64": float(np.max(s64[:, 2])), "fp32": float(np.max(s32[:, 2]))},
        "rho_ex": {"initial": float(r64[0]), "final_fp64": float(r64[-1]), "final_fp32": float(r32[-1])},
        "growth_rate": growth,
    }
    print(json.dumps(out), flush=True)


# The oracle on this container's 8 cores runs the 128x128x642 crop at
# 1.21e7 updates/s; the reference's own OpenMP build (nuosc_class.cpp, FP64),
# timed on the same crop and the same 8 cores when this path was surveyed,
# ran it at 1.60e7: the oracle is 0.76x the reference's speed.
ORACLE_VS_REFERENCE = 1.21e7 / 1.60e7


def cpu_baseline(nn, dz, seconds, nflavor=2, pmo=0.0, reps=3):
    """Oracle (OpenMP C restatement of the reference) on a 128x128 crop:
    reps timed samples of about seconds/reps each, the median reported."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    half = 64 * dz
    o = O.OracleSim(nn=nn, x0=-half, x1=half, z0=-half, z1=half, dx=dz, dz=dz, cfl=0.4, ko=0.1, mu=1.0, pmo=pmo,
                    nflavor=nflavor)
    o.fill_init(ipt=5, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5, seed=12345)
    nx, nz, nv = o.shape
    o.step(1)  # warm caches / page in
    rates, nsteps = [], []
    for _ in range(reps):
        n, t0 = 0, time.perf_counter()
        while True:
            o.step(1)
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds / reps or n >= 20:
                break
        rates.append(nx * nz * nv * n / el)
        nsteps.append(n)
    o.close()
    med = float(np.median(rates))
    return {
        "value": med,
        "unit": "site*dir RK4 updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{nx}x{nz}x{nv} crop of the same physics ({nflavor} flavours, FP64), median of {reps} samples of "
                  f"{nsteps} RK4 steps ({', '.join('%.3g' % r for r in rates)} updates/s), "
                  f"OMP_NUM_THREADS={threads}",
        "reference_equivalent": med / ORACLE_VS_REFERENCE,
        "reference_equivalent_note": "value / 0.76: the oracle ran at 0.76x the reference's own OpenMP build on the "
                                     "same crop and 8 cores of the build container",
    }


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {a.gpus}); "
              "the periodic box would shard as an (x,z) domain decomposition with radius-2 halos.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "batched":
        bench_batched(a)
        return
    if a.workload == "served":
        bench_served(a)
        return
    if a.workload == "precision":
        bench_precision(a)
        return
    import nuosc_amd as N

    out, box = measure_box(N, a, a.precision)
    default_box = not (a.shape == "column" or a.nflavor == 3 or a.no_extras)
    if default_box and a.precision == 64:
        # the FP32 form on the same workload, and the batched scan, each in a
        # short leg of its own (the FP64 handle is closed by then)
        a32 = argparse.Namespace(**vars(a))
        a32.precision = 32
        o32, box32 = measure_box(N, a32, 32)
        out["fp32"] = {
            "value": o32["value"], "ms_per_step": o32["ms_per_step"],
            "speedup_vs_fp64": o32["value"] / out["value"],
            "roofline": {k: o32["roofline"][k] for k in ("kernel", "achieved", "frac", "traffic", "launch_ms")},
            "step_roofline_frac": o32["step_roofline"]["frac"], "state_ok": o32["state_ok"],
        }
        out["batched"] = batched_leg(N, a)
        # p50/p99 run latency of the served scan in a short window (20 s of
        # Poisson 0.35/s arrivals, about what the GPU sustains with this draw;
        # steps U[100,1000] so the leg ends within about a minute); the
        # 600 s run at the full step range is `--workload served --rate 0.35
        # --duration 600 --precision 32 --steps 200`
        sv = served_leg(0.35, 20.0, 100, 32)
        out["served"] = {k: sv[k] for k in ("value", "unit", "requests", "failed", "latency_p50_s",
                                            "latency_p99_s", "queue_p50_s", "queue_p99_s", "gpu_busy_frac",
                                            "dtype")}
        out["served"]["workload"] = sv["config"]["workload"]
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.nn, a.dz, a.cpu_seconds, a.nflavor, 1.0 if a.nflavor == 3 else 0.0)
    print(json.dumps(out), flush=True)


def measure_box(N, a, precision):
    """One large run (the box, or the column with --shape column): autotune,
    warm-up, per-stage kernel times from HIP events, then K timed steps."""
    prec = N.PREC_FP64 if precision == 64 else N.PREC_FP32
    esize = 8 if precision == 64 else 4
    three = a.nflavor == 3
    nvar = 18 if three else 8
    passes = STAGE_PASSES
    pmo = 1.0 if three else 0.0
    column = a.shape == "column"
    if column:   # the 1-D open column of the precision workload
        cfg = dict(nv=a.nn, x0=-0.025, x1=0.025, z0=-102.4, z1=102.4, dx=0.05, dz=0.05, CFL=0.4, ko=0.1, mu=1.0,
                   pmo=pmo, gz=3, bc=N.BC_OPEN_Z, zc=-105.4, precision=prec, use_graph=True, nflavor=a.nflavor)
    else:
        cfg = dict(nv=a.nn, x0=-a.zmax, x1=a.zmax, z0=-a.zmax, z1=a.zmax, dx=a.dz, dz=a.dz, CFL=0.4, ko=0.1,
                   mu=1.0, pmo=pmo, precision=prec, use_graph=True, nflavor=a.nflavor)
    g = N.NuOsc(device=0, **cfg)
    if column:
        g.fillInitValue(ipt=0, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5)
    else:
        g.fillInitValue(ipt=5, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5, seed=12345)
    nx, nz, nv = g.shape
    upd = nx * nz * nv
    if a.zlen and not three:
        g.set_zlen(a.zlen)
        zlen = a.zlen
    else:
        zlen = g.plan()[0] if a.no_autotune else g.autotune(trials=2)[0]

    g.step_rk4(a.warmup)
    g.synchronize()
    # per-stage kernel durations from HIP events around each stage kernel on
    # the library's stream, right before the timed steps
    stage_ms = g.profile_stages(3)
    t0 = time.perf_counter()
    g.step_rk4(a.steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = g.last_step_ms()
    ana = g.analysis()
    ok = ana.status == 0 and ana.nonfinite == 0
    g.close()

    value = upd * a.steps / wall
    s2_bytes = STAGE_PASSES[1] * nvar * esize * upd
    s2_gbs = s2_bytes / (stage_ms[1] * 1e-3) / 1e9
    step_bytes = sum(passes) * nvar * esize * upd
    step_gbs = step_bytes / (wall / a.steps) / 1e9
    traffic = None if (three or column) else traffic_from_profiles(precision)
    out = {
        "metric": "site*dir RK4 updates/s",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == 64 else "f32",
        "data": "synthetic (ipt-0 pulse entering through z0)" if column
                else "synthetic (seeded eps=1e-6 complex perturbation)",
        "config": {
            "workload": (f"{nx}x{nz} 1-D open column (gz=3) x {nv} icosahedral directions, " if column else
                         f"{nx}x{nz} periodic box x {nv} icosahedral directions, ") +
                        f"{'3 flavours (18 fields), ' if three else ''}"
                        f"{'FP64' if precision == 64 else 'FP32 + FP64 moments'}, mu=1 ko=0.1 cfl=0.4 pmo={pmo:g}",
            "nx": nx, "nz": nz, "nv": nv, "dz": a.dz, "zlen": zlen,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_stage3z stage 2" if three else "k_stage stage 2",
            "achieved": s2_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": s2_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"{TRAFFIC_FILE}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, "
                               "scaled by a plain copy's readings; not measured in this run") if traffic else None,
            "algorithmic_bytes_per_launch": s2_bytes,
            "launch_ms": stage_ms[1],
        },
        "step_roofline": {
            "achieved": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
            "bytes_per_update": sum(passes) * nvar * esize,
            "vs_16pass_ceiling": value / (HBM_PEAK_GBS * 1e9 / (16 * nvar * esize)),
            "stage_ms": [round(x, 4) for x in stage_ms],
            "device_ms_per_step": dev_ms / a.steps,
        },
        "state_ok": ok,
    }
    return out, g


def batched_leg(N, a, runs=64, steps=20):
    """The batched-scan workload (64 uneven runs, FP32) for a few steps:
    aggregate updates/s stepped all together (at the autotuned z-segment
    length and launch form; the other launch form is timed too) and one
    after another, and each mode's HBM fraction at 384 B per FP32 update
    (12 passes x 8 fields)."""
    import numpy as np
    rng = np.random.default_rng(7)
    specs = [scan_run(rng, k) for k in range(runs)]
    sims = []
    for geo, init in specs:
        g = N.NuOsc(device=0, precision=N.PREC_FP32, use_graph=True, **geo)
        g.fillInitValue(**init)
        sims.append(g)
    upd = sum(g.nx * g.nz * g.nv for g in sims)
    for g in sims:
        g.autotune(trials=2)
        g.step_rk4(1)
        g.synchronize()
    t0 = time.perf_counter()
    for g in sims:
        g.step_rk4(steps)
        g.synchronize()
    t_seq = time.perf_counter() - t0
    b = N.Batch(device=0, precision=N.PREC_FP32)
    for g in sims:
        b.add(g)
    zbest, _ = b.autotune(trials=2)
    chosen = b.launch
    # the autotuned pair, then the other launch form at the same zlen
    t_form = {}
    for lf in (chosen, 1 - chosen):
        b.set_launch(lf)
        b.step(1)
        t0 = time.perf_counter()
        b.step(steps)
        t_form[lf] = time.perf_counter() - t0
    t_bat = t_form[chosen]
    ok = all(g.analysis().status == 0 for g in sims)
    b.close()
    for g in sims:
        g.close()
    bpu = 12 * 8 * 4
    v_bat, v_seq = upd * steps / t_bat, upd * steps / t_seq
    return {
        "value": v_bat, "unit": "updates/s", "sequential_value": v_seq, "speedup_vs_sequential": t_seq / t_bat,
        "hbm_frac": v_bat * bpu / (HBM_PEAK_GBS * 1e9), "sequential_hbm_frac": v_seq * bpu / (HBM_PEAK_GBS * 1e9),
        "bytes_per_update": bpu, "steps": steps, "runs": runs, "dtype": "f32", "batch_zlen": zbest,
        "batch_launch": "persistent" if chosen == N.BATCH_PERSISTENT else "grid",
        "grid_value": upd * steps / t_form[N.BATCH_GRID],
        "persistent_value": upd * steps / t_form[N.BATCH_PERSISTENT],
        "workload": f"{runs} runs, grids drawn from 64..256 squared, 162/642 directions, uneven physics",
        "updates_per_step": upd, "state_ok": ok,
    }


if __name__ == "__main__":
    main()
