"""Throughput of the fused D3Q27 BGK step on one MI355X.

Workload: Taylor-Green vortex, 512^3 periodic, U0 = 0.05, Re = 1600, AA-pattern
in-place storage -- the large TGV precision sweep.  A step is one time step
over the whole 512^3 domain (one fused kernel launch).  The headline value is
the FP64 (drop-in) throughput in GLUPS; the same run also steps FP32- and
FP16-shifted storage from the same initial state and reports their GLUPS and
their velocity L2 error against the FP64 run after the same steps.

    python bench.py [--steps K] [--warmup W] [--n 512] [--storage aa|pull] [--placement window|allocations|none]

Prints one JSON line.  Inputs are resident in HBM before the timed region; the
timed region is K launches on the library's stream, bracketed by HIP events.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLUPS + % HBM roofline; FP64 vs FP16-shifted storage at stated TGV error"
PEAK_GBS = 8000.0
PI_R = 3.1415926


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tgv_tau(n, re, u0):
    return u0 * n / re / (1.0 / 3.0) / 2 / PI_R


def run_gpu(n, steps, warmup, storage, dtype, u0, re, tune=True, launch=None, placement=None):
    """placement: None (the first allocation), ("allocations", trials) or
    ("window", slack_gib, step_gib)."""
    from imexlb_amd import LBM
    kw = {}
    if launch:
        kw = dict(block=launch[0], cells_per_thread=launch[1], xcd_remap=launch[2],
                  nt_loads=launch[3] if len(launch) > 3 else -1)
        tune = False
    g = LBM(n, n, n, tgv_tau(n, re, u0), u0=u0, case="tgv", engine="fused", dtype=dtype,
            storage=storage, **kw)
    g.Initialize()
    info_p = None
    lat_gb = 27 * n ** 3 * {"f64": 8, "f32": 4, "f16": 2}[dtype] * (2 if storage == "pull" else 1) / 1e9
    if placement and placement[0] == "allocations" and placement[1] > 1:
        # the lattice's HBM placement moves the step rate by up to 20 %; the
        # library times up to `trials` allocations and keeps the fastest
        kept, ms_each = g.tune_placement(placement[1], steps=10)
        rates = [round(n ** 3 / (m * 1e-3) / 1e9, 3) if m > 0 else None for m in ms_each]
        info_p = {"method": "separate allocations", "trials": placement[1], "glups_each": rates, "kept": kept,
                  "held_gb_during_search": round(placement[1] * lat_gb, 1), "held_gb_after": round(lat_gb, 1)}
        log(f"[bench] {dtype}: placement GLUPS {rates}, kept {kept}")
    elif placement and placement[0] == "window":
        # the rate follows the lattice's base address smoothly: slide it
        # through one address range of lattice + slack and keep the fastest
        # offset (the library then gives back the chunks the lattice does not
        # touch)
        from imexlb_amd import ImexlbError
        _, slack, step = placement
        try:
            kept, ms_each = g.tune_placement_window(slack_gib=slack, step_gib=step, steps=10)
        except ImexlbError as err:
            # the library keeps the lattice it had: time the first allocation
            log(f"[bench] {dtype}: window search failed ({err}); timing the first allocation")
            kept = None
        if kept is not None:
            rates = [round(n ** 3 / (m * 1e-3) / 1e9, 3) if m > 0 else None for m in ms_each]
            held = g.info()["device_bytes"] / 1e9
            info_p = {"method": "window", "slack_gib": slack, "step_gib": step, "glups_each": rates, "kept": kept,
                      "held_gb_during_search": round(lat_gb + slack * 2 ** 30 / 1e9, 1),
                      "held_gb_after": round(held, 1)}
            log(f"[bench] {dtype}: window GLUPS {rates}, kept offset {kept} ({rates[kept]})")
    if tune:
        # launch shape from the in-project autotuner, on the placement kept
        # above (re-initialises the state)
        choice = g.autotune(os.environ.get("IMEXLB_TUNE_CACHE"), steps=8)
        log(f"[bench] {dtype}: autotuned block {choice[0]}, cells/thread {choice[1]}, xcd order {choice[2]}, "
            f"nt loads {choice[3]}")
    g.step(warmup, moments_every=0)
    g.sync()
    ms = g.step_timed(steps, moments_every=0)
    # one more step writes the moments used for the error report
    g.step(1, moments_every=1)
    info = g.info()
    info["placement"] = info_p
    return g, ms, info


def placement_summary(placement, bytes_lup):
    """First, median and kept rates of a placement search, as GLUPS and as
    the fraction of the 8 TB/s peak.  "First" is the first allocation, or the
    window's first offset (where a fresh allocation of the lattice lands)."""
    if not placement:
        return {}
    rates = [r for r in placement["glups_each"] if r]
    frac = lambda g: g * bytes_lup / PEAK_GBS
    first, med, kept = rates[0], float(np.median(rates)), placement["glups_each"][placement["kept"]]
    return {"glups_first_alloc": first, "glups_median": med, "glups_kept": kept,
            "frac_first_alloc": frac(first), "frac_median": frac(med), "frac_kept": frac(kept)}


def value_basis(p):
    if not p:
        return "value and frac: the timed run on the first lattice allocation"
    if p["method"] == "window":
        return (f"value and frac: the timed run with the lattice at the fastest of {len(p['glups_each'])} offsets "
                f"({p['step_gib']} GiB apart) in one window of lattice + {p['slack_gib']} GiB "
                f"({p['held_gb_during_search']} GB held during the search, {p['held_gb_after']} GB after); "
                "glups/frac_first_alloc is offset 0, glups/frac_median the median offset")
    return (f"value and frac: the timed run on the fastest of {p['trials']} lattice allocations "
            f"({p['held_gb_during_search']} GB held during the search); glups/frac_first_alloc is the first "
            "allocation, glups/frac_median the median candidate")


def trials_for_budget(n, dtype, storage, default, budget_gb):
    """Placement candidates that fit in `budget_gb` of device memory held at
    once (every candidate keeps its own lattice until the search ends)."""
    elem = {"f64": 8, "f32": 4, "f16": 2}[dtype]
    lat_gb = 27 * n ** 3 * elem * (2 if storage == "pull" else 1) / 1e9
    return max(1, min(default, int(budget_gb // lat_gb))), lat_gb


def tgv_analytic_velocity(n, u0, re, steps):
    """The Taylor-Green velocity of the initial field (tgv/src/lbm.cpp:197-200,
    pi = 3.1415926) decayed viscously: |k|^2 = 3 (2 pi / n)^2, u(t) = u(0)
    exp(-nu |k|^2 t), nu = tau0 / 3.  Valid while the nonlinear transfer is
    small (t U0 k << 1)."""
    nu = tgv_tau(n, re, u0) / 3.0
    k = 2.0 * np.pi / n
    decay = np.exp(-nu * 3.0 * k * k * steps)
    a = np.arange(n) / n * 2.0 * PI_R
    s, c = np.sin(a), np.cos(a)
    # arrays indexed [z][y][x]
    u = u0 * decay * s[None, None, :] * c[None, :, None] * c[:, None, None]
    v = -u0 * decay * c[None, None, :] * s[None, :, None] * c[:, None, None]
    return u, v


def vel_l2_analytic(m, n, u0, re, steps):
    u, v = tgv_analytic_velocity(n, u0, re, steps)
    num = float(np.sum((m[1] - u) ** 2) + np.sum((m[2] - v) ** 2) + np.sum(m[3] ** 2))
    den = float(np.sum(u ** 2) + np.sum(v ** 2))
    return (num / den) ** 0.5


def vel_l2(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a[1:], b[1:]))
    den = sum(float(np.sum(y ** 2)) for y in b[1:])
    return (num / den) ** 0.5


def traffic_from_profiles(workload):
    """HBM bytes per launch of the workload's kernel, from the calibrated
    FETCH_SIZE/WRITE_SIZE passes recorded under profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f).get(workload)
        return (t["bytes_per_launch"], t["ratio"], t["source"]) if t else None
    except (OSError, ValueError, KeyError):
        return None


def host_cpu():
    """CPU model, logical CPUs this process may run on, and host RAM."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    ram_gb = None
    try:
        with open("/proc/meminfo") as f:
            ram_gb = round(int(f.readline().split()[1]) / 2 ** 20, 1)
    except (OSError, ValueError, IndexError):
        pass
    return model, avail, ram_gb


def cpu_baseline(budget_s):
    """The oracle -- the reference's phases (Collision, the 26-message ghost-3
    periodic self-copy, the two streaming passes, Update) on its ghost-3
    f(q, x, y, z) layout in FP64 with OpenMP -- on a bounded sample of the
    same workload: TGV 128^3, same U0 and Re.  It is compiled for this host
    (-march=native) into a temporary directory, runs on every CPU this process
    may use (OMP_PROC_BIND=spread, OMP_PLACES=cores), and is timed phase by
    phase after one warm-up step, as the reference's driver times them."""
    import tempfile
    model, avail, ram_gb = host_cpu()
    threads = min(avail, int(os.environ.get("OMP_NUM_THREADS", "0")) or avail)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as orc
    build = "generic x86-64-v3 build (native build failed)"
    try:
        path = orc.build_native(tempfile.mkdtemp(prefix="imexlb_oracle_native_"))
        orc.use_library(path)
        build = "-march=native build"
    except Exception as exc:  # the generic in-tree build still times the same code
        log(f"[bench] CPU baseline: native oracle build failed ({exc}); timing the generic build")
    n = 128
    o = orc.Oracle(orc.TGV, n, n, n, tau0=tgv_tau(n, 1600, 0.05), u0=0.05, threads=threads)
    o.Initialize()
    o.run(1)  # warm-up
    phases = ("Collision", "pack", "exchange", "unpack", "Streaming", "Update")
    spent = dict.fromkeys(phases, 0.0)
    t0 = time.perf_counter()
    k = 0
    while True:
        for name in phases:
            ts = time.perf_counter()
            getattr(o, name)()
            spent[name] += time.perf_counter() - ts
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s or k >= 200:
            break
    glups = n ** 3 * k / el / 1e9
    return {"value": glups, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle (reference phases, ghost 3, FP64, OpenMP, {build}) TGV {n}^3 Re 1600, "
                      f"{k} steps in {el:.1f} s after 1 warm-up step",
            "cpu_model": model, "nproc": avail, "host_ram_gb": ram_gb,
            "omp": {"threads": threads, "proc_bind": os.environ["OMP_PROC_BIND"],
                    "places": os.environ["OMP_PLACES"]},
            "gb_per_s": glups * 432.0,
            "ms_per_step_by_phase": {name: 1e3 * t / k for name, t in spent.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--launch", default=None,
                    help="block,cells_per_thread,xcd_order,nt_loads for FP64 (skips the autotuner)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--storage", default="aa", choices=["aa", "pull"])
    ap.add_argument("--re", type=float, default=1600.0)
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lowp", action="store_true")
    ap.add_argument("--placement", default="window", choices=["window", "allocations", "none"],
                    help="lattice placement search: offsets in one window of device memory chunks (default), "
                         "separate allocations, or the first allocation")
    ap.add_argument("--window-slack-gib", type=float, default=96.0,
                    help="window beyond the lattice for --placement window; the fast offsets recur irregularly, "
                         "about 30-60 GiB apart (held during the search only)")
    ap.add_argument("--placements", type=int, default=None,
                    help="time this many separate allocations instead (1: the first allocation only)")
    ap.add_argument("--placement-budget-gb", type=float, default=128.0,
                    help="--placement allocations: device memory the candidates may hold at once")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the multi-GPU halo over RCCL is a design "
              "(see the sharding section of the design notes), not a measured path.", file=sys.stderr)
        sys.exit(2)

    import imexlb_amd  # noqa: F401  (fails loudly without libimexlb.so)
    from imexlb_amd import BANDWIDTH_PROBES, bandwidth_probe, copy_bandwidth

    n, K, W = a.n, a.steps, a.warmup
    cells = n ** 3
    log(f"[bench] TGV {n}^3 {a.storage} FP64: warmup {W}, timed {K}")
    launch = tuple(int(v) for v in a.launch.split(",")) if a.launch else None
    placement = {}
    for dt in ("f64", "f32", "f16"):
        if a.placements is not None:
            placement[dt] = ("allocations", a.placements) if a.placements > 1 else None
        elif a.placement == "window":
            placement[dt] = ("window", a.window_slack_gib, {"f64": 2, "f32": 1, "f16": 1}[dt])
        elif a.placement == "allocations":
            t, _ = trials_for_budget(n, dt, a.storage, {"f64": 8, "f32": 8, "f16": 12}[dt], a.placement_budget_gb)
            placement[dt] = ("allocations", t)
        else:
            placement[dt] = None
    g64, ms64, info64 = run_gpu(n, K, W, a.storage, "f64", a.u0, a.re, not a.no_tune, launch, placement["f64"])
    glups64 = cells * K / (ms64 * 1e-3) / 1e9
    per_launch_ms = ms64 / K
    bytes_lup = info64["bytes_per_lup"]
    achieved = bytes_lup * cells / (per_launch_ms * 1e-3) / 1e9
    m64 = g64.moments()
    g64.close()  # its lattice is not needed again; frees room for the other storage types
    an64 = vel_l2_analytic(m64, n, a.u0, a.re, W + K + 1)
    log(f"[bench] FP64 velocity L2 vs the analytic decay after {W + K + 1} steps: {an64:.3e}")
    log(f"[bench] FP64 {glups64:.2f} GLUPS, {achieved:.0f} GB/s")

    extra = {}
    if not a.no_lowp:
        for dt in ("f32", "f16"):
            log(f"[bench] TGV {n}^3 {a.storage} {dt}")
            g, ms, info = run_gpu(n, K, W, a.storage, dt, a.u0, a.re, not a.no_tune, None, placement[dt])
            gl = cells * K / (ms * 1e-3) / 1e9
            if info["placement"]:
                info["placement"].update(placement_summary(info["placement"], info["bytes_per_lup"]))
            m = g.moments()
            err = vel_l2(m, m64)
            an = vel_l2_analytic(m, n, a.u0, a.re, W + K + 1)
            bw = info["bytes_per_lup"] * cells / (ms / K * 1e-3) / 1e9
            extra[f"{dt}_shifted"] = {"glups": gl, "ms_per_step": ms / K, "achieved_GBps": bw,
                                      "frac": bw / PEAK_GBS, "vs_f64": gl / glups64,
                                      "vel_l2_vs_f64": err, "vel_l2_vs_analytic": an, "after_steps": W + K + 1,
                                      "placement": info["placement"]}
            log(f"[bench] {dt} {gl:.2f} GLUPS ({gl / glups64:.2f}x FP64), L2 vs FP64 {err:.2e}")
            del g, m
    del m64

    # bandwidth ceiling: the plain copy and the tuned stream probes (copies,
    # read-only, write-only) over 4 GiB per stream; the best copy is the
    # measured ceiling for a kernel with one read and one write per value
    gbps_plain, _ = copy_bandwidth(4 << 30, reps=10)
    probes = {"copy_plain": gbps_plain}
    for kind in BANDWIDTH_PROBES:
        probes[kind] = bandwidth_probe(kind, 4 << 30, reps=10)[0]
    best_copy = max((k for k in probes if k.startswith("copy")), key=probes.get)
    gbps_copy = probes[best_copy]
    log(f"[bench] device copy {gbps_plain:.0f} GB/s plain, {gbps_copy:.0f} GB/s best ({best_copy}); probes "
        + ", ".join(f"{k} {v:.0f}" for k, v in probes.items()))
    workload = f"tgv{n}_{a.storage}_f64"
    traffic = traffic_from_profiles(workload)
    cpu = None
    if not a.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(a.cpu_budget)

    out = {
        "metric": METRIC,
        "value": glups64,
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": per_launch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex D3Q27 BGK {n}^3 periodic, U0 {a.u0}, Re {a.re:g}, "
                               f"{a.storage.upper()} storage, FP64 populations and collision",
                   "cells": cells, "storage": a.storage, "re": a.re, "u0": a.u0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_GBS, "traffic": traffic[0] if traffic else None,
                     "traffic_over_algorithmic": traffic[1] if traffic else None,
                     "traffic_source": traffic[2] if traffic else None,
                     "frac_of_measured_copy": achieved / gbps_copy,
                     "bytes_per_lup": bytes_lup, "launch": {k: info64[k] for k in ("block", "cells_per_thread", "xcd_remap", "nt_loads")},
                     "kernel": "k_fused (FP64, periodic, "
                     + ("AA" if a.storage == "aa" else "pull") + ")",
                     "measured_copy_GBps": gbps_copy, "measured_copy_kernel": best_copy,
                     "bandwidth_probes_GBps": {k: round(v, 1) for k, v in probes.items()},
                     "placement": info64["placement"],
                     **placement_summary(info64["placement"], bytes_lup),
                     "value_basis": value_basis(info64["placement"])},
        "cpu_baseline": cpu,
        "tgv_error": {"after_steps": W + K + 1, "f64_vel_l2_vs_analytic": an64,
                      "note": "analytic = initial TGV field decayed by exp(-3 nu k^2 t), the linear solution; "
                              "the departure grows as the nonlinear transfer, about 0.25 t U0 k (t U0 k = "
                              f"{(W + K + 1) * a.u0 * 2 * np.pi / n:.3f} here). The FP32/FP16 lines carry their "
                              "L2 against FP64 and against the same analytic field"},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
