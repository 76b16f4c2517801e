"""Long runs of the benchmark configurations on one GPU; each prints one JSON
line.

    python -m imexlb_amd.runs precision [--n 512] [--steps 5000]
    python -m imexlb_amd.runs cavity    [--n 128] [--steps 20000] [--placements N]
    python -m imexlb_amd.runs sphere    [--steps 50000] [--placements N]
    python -m imexlb_amd.runs ensemble  [--steps 2000] [--domains 256] [--storage aa|pull] [--placements N]
    python -m imexlb_amd.runs timing    [--n 512 | --shape NX,NY,NZ] [--steps 20] [--dtype f16] [--window]
                                        [--storage aa] [--launch B,C,X,N] [--placements 1]
    python -m imexlb_amd.runs phased    [--n 256] [--steps 20]
    python -m imexlb_amd.runs placement [--n 512] [--storage aa] [--dtypes f64,f32,f16] [--windows 2]
                                        [--allocations 4]

precision: TGV 512^3, U0 0.05, Re 1600, AA storage in FP64 / FP32 / FP16
    (shifted), each placed with the window search, then autotuned; kinetic
    energy against the analytic early-time decay
    E0 exp(-6 nu k^2 t), and the velocity L2 error of FP32/FP16 against FP64.
cavity: 128^3, Re 400, lid 0.1, halfway bounce-back with the lid momentum
    term; fused pull, fused AA and the phased engine compared every 1000 steps.
sphere: 512x256x256 channel, sphere D = 40 at (128, 128, 128), inlet 0.05,
    Re 200; FP64 / FP32 / FP16 pull; drag proxy every 1000 steps and the final
    velocity field compared with FP64.
ensemble: 256 domains from seed 20261202 (TGV / cavity, edges 32..96,
    Re 50..1000, lid 0.05..0.1), FP32 shifted storage, pull (default) or AA;
    window placement, then the launch autotuner; persistent launch with graph
    replay vs persistent launches vs one launch per domain; per-domain
    velocity error against the same ensemble in FP64.
Every run but timing places its lattices with the window search (the
fastest offset of the lattice inside one address range of lattice + 96 GiB);
--placements N times N separate allocations instead, --placements 1 keeps the
first.
timing: TGV n^3 or NX x NY x NZ (U0 0.05, Re 1600), one dtype and storage, the first
    allocation (or --window / --placements N), a fixed launch shape (or the
    autotuner's, on the kept placement), `steps` timed steps after 4 warm-up
    steps; GLUPS and algorithmic GB/s.  The short run to profile one kernel with.
placement: TGV n^3 in each storage type at a fixed launch shape: the step
    rate with the lattice at every offset of a window search (--windows of
    them, 2 by default, in separate windows: 64 GiB in 2 GiB steps for FP64,
    1 GiB steps for FP32 and FP16), then over --allocations separate
    allocations (4).
phased: the reference's own boundary code (faithful mode) on the phased
    engine's multi-step path: the cavity at n^3 (lid 0.1, tau0 0.096) and the
    sphere input of sphere/input.in (512 x 128 x 128, Re 100, u0 0.05); and the
    same two on the fused engine (ghost ring + one fused pass); GLUPS over
    `steps` timed steps after 3 warm-up steps.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import LBM, Ensemble, copy_bandwidth

PI_R = 3.1415926


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def l2(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y ** 2)) for y in b)
    return math.sqrt(num / den) if den > 0 else float("nan")


def place(g, a, cells, label):
    """Lattice placement search before a timed run: with --placements N > 1,
    the fastest of N separate allocations; by default the fastest offset of
    a window search (the lattice slid through one address range of lattice
    + 96 GiB, whose unused chunks are given back afterwards).
    --placements 1: the first allocation.  Returns the record, None if no
    search ran."""
    def rate(m):
        return round(cells / (m * 1e-3) / 1e9, 2) if m > 0 else None
    if a.placements is not None:
        if a.placements <= 1:
            return None
        kept, ms = g.tune_placement(a.placements, steps=10)
        rec = {"method": "separate allocations", "kept": kept, "glups_each": [rate(m) for m in ms]}
    else:
        lat_gib = g.info()["device_bytes"] / 2 ** 30
        step = 2 if lat_gib >= 16 else (1 if lat_gib >= 4 else 0.5)
        # the fast offsets recur irregularly, 30-60 GiB apart
        slack = 96
        kept, ms = g.tune_placement_window(slack_gib=slack, step_gib=step, steps=10)
        rec = {"method": "window", "slack_gib": slack, "step_gib": step, "kept": kept,
               "glups_each": [rate(m) for m in ms]}
    log(f"[{label}] placement: {rec['method']}, kept {rec['kept']} ({rec['glups_each'][rec['kept']]} GLUPS), "
        f"range {min(r for r in rec['glups_each'] if r)}-{max(rec['glups_each'])}")
    return rec


def precision(a):
    n, steps = a.n, a.steps
    u0, re = 0.05, 1600.0
    tau = u0 * n / re / (1.0 / 3.0) / 2 / PI_R
    nu = tau / 3.0
    k = 2 * PI_R / n
    samples = sorted({s for s in (1, 100, 500, 1000, 2000, 3000, 4000, 5000, steps) if s <= steps})
    runs = {}
    out = {"workload": f"TGV {n}^3 Re {re:g} U0 {u0} AA", "nu": nu, "samples": [], "placement": {}}
    for dt in ("f64", "f32", "f16"):
        g = LBM(n, n, n, tau, u0=u0, case="tgv", dtype=dt, storage="aa")
        g.Initialize()
        out["placement"][dt] = place(g, a, n ** 3, f"precision {dt}")
        g.autotune(None, steps=4)  # on the kept placement; re-initialises
        runs[dt] = g
    ke0 = None
    done = 0
    t_dev = {dt: 0.0 for dt in runs}
    for s in samples:
        rec = {"step": s}
        vel = {}
        for dt, g in runs.items():
            t_dev[dt] += g.step_timed(s - done, moments_every=s - done)
            ke, mass = g.diagnostics()
            rec[f"ke_{dt}"] = ke
            rec[f"mass_{dt}"] = mass
            p, u, v, w = g.moments()
            vel[dt] = (u, v, w)
        done = s
        if ke0 is None:
            ke0 = rec["ke_f64"] / math.exp(-6 * nu * k * k * s)
        rec["ke_analytic"] = ke0 * math.exp(-6 * nu * k * k * s)
        rec["ke_f64_over_analytic"] = rec["ke_f64"] / rec["ke_analytic"]
        for dt in ("f32", "f16"):
            rec[f"vel_l2_{dt}_vs_f64"] = l2(vel[dt], vel["f64"])
        out["samples"].append(rec)
        log(f"[precision] step {s}: KE/analytic {rec['ke_f64_over_analytic']:.6f}, "
            f"L2 f32 {rec['vel_l2_f32_vs_f64']:.3e}, f16 {rec['vel_l2_f16_vs_f64']:.3e}")
        del vel
    cells = n ** 3
    out["glups"] = {dt: cells * steps / (t * 1e-3) / 1e9 for dt, t in t_dev.items()}
    out["note"] = "GLUPS here include the moment-writing steps at the samples"
    return out


def cavity(a):
    n, steps = a.n, a.steps
    lid, re = 0.1, 400.0
    tau = 3.0 * lid * n / re
    engines = {
        "fused_pull": LBM(n, n, n, tau, u0=0.0, case="cavity", mode="config", storage="pull", lid_u=lid),
        "fused_aa": LBM(n, n, n, tau, u0=0.0, case="cavity", mode="config", storage="aa", lid_u=lid),
        "phased": LBM(n, n, n, tau, u0=0.0, case="cavity", mode="config", engine="phased", lid_u=lid),
    }
    out = {"workload": f"cavity {n}^3 Re {re:g} lid {lid}", "tau0": tau, "checks": [], "placement": {}}
    for name, g in engines.items():
        g.Initialize()
        if name != "phased":  # the phased engine keeps its own layout
            out["placement"][name] = place(g, a, n ** 3, f"cavity {name}")
    t = {k: 0.0 for k in engines}
    every = 1000
    for s in range(every, steps + 1, every):
        ref = None
        rec = {"step": s}
        for name, g in engines.items():
            t0 = time.perf_counter()
            g.step(every, moments_every=every)
            g.sync()
            t[name] += time.perf_counter() - t0
            m = g.moments()
            if ref is None:
                ref = m
                rec["umax"] = float(np.abs(m[1]).max())
                rec["ke"], rec["mass"] = g.diagnostics()
            else:
                rec[f"{name}_bitexact_vs_fused_pull"] = all(np.array_equal(x, y) for x, y in zip(m, ref))
        out["checks"].append(rec)
        log(f"[cavity] step {s}: " + json.dumps(rec))
    out["wall_s"] = t
    out["all_bitexact"] = all(v for r in out["checks"] for k, v in r.items() if k.endswith("bitexact_vs_fused_pull"))
    return out


def sphere(a):
    nx, ny, nz = 512, 256, 256
    u_in, d, re = 0.05, 40.0, 200.0
    tau = 3.0 * u_in * d / re
    sph = (128.0, 128.0, 128.0, d / 2)
    runs = {dt: LBM(nx, ny, nz, tau, u0=u_in, case="sphere", mode="config", dtype=dt, storage="pull",
                    sphere=sph) for dt in ("f64", "f32", "f16")}
    out = {"workload": f"sphere channel {nx}x{ny}x{nz} D {d:g} Re {re:g} u_in {u_in}", "tau0": tau, "drag": [],
           "placement": {}}
    for dt, g in runs.items():
        g.autotune(None, steps=4)
        out["placement"][dt] = place(g, a, nx * ny * nz, f"sphere {dt}")
    every = 1000
    t_dev = {dt: 0.0 for dt in runs}
    for s in range(every, a.steps + 1, every):
        rec = {"step": s}
        for dt, g in runs.items():
            t_dev[dt] += g.step_timed(every, moments_every=0)
            rec[f"fx_{dt}"] = float(g.sphere_force()[0])
        for dt in ("f32", "f16"):
            rec[f"fx_relerr_{dt}"] = abs(rec[f"fx_{dt}"] - rec["fx_f64"]) / abs(rec["fx_f64"])
        out["drag"].append(rec)
        if s % 10000 == 0 or s == every:
            log(f"[sphere] step {s}: " + json.dumps(rec))
    vel = {}
    for dt, g in runs.items():
        g.step(1, moments_every=1)
        p, u, v, w = g.moments()
        vel[dt] = (u, v, w)
    out["vel_l2_vs_f64"] = {dt: l2(vel[dt], vel["f64"]) for dt in ("f32", "f16")}
    cells = nx * ny * nz
    out["glups"] = {dt: cells * (a.steps // every * every) / (t * 1e-3) / 1e9 for dt, t in t_dev.items()}
    out["cd_proxy_f64"] = out["drag"][-1]["fx_f64"] / (0.5 * 1.0 * u_in ** 2 * math.pi * (d / 2) ** 2)
    return out


def ensemble(a):
    doms = Ensemble.generate(20261202, a.domains)
    cells = sum(d.nx * d.ny * d.nz for d in doms)
    out = {"workload": f"{a.domains} domains seed 20261202, {cells} cells, {a.storage} storage", "cells": cells,
           "storage": a.storage}
    e = Ensemble(doms, dtype="f32", storage=a.storage)
    e.Initialize()
    def rate(m):
        return round(cells / (m * 1e-3) / 1e9, 2) if m > 0 else None
    if a.placements is not None:
        if a.placements > 1:
            kept, ms_each = e.tune_placement(a.placements, steps=6)
            out["placement"] = {"method": "separate allocations", "kept": kept,
                                "glups_each": [rate(m) for m in ms_each]}
    else:
        kept, ms_each = e.tune_placement_window(slack_gib=96, step_gib=2, steps=6)
        out["placement"] = {"method": "window", "slack_gib": 96, "step_gib": 2, "kept": kept,
                            "glups_each": [rate(m) for m in ms_each]}
    if out.get("placement"):
        log(f"[ensemble] placement: {out['placement']}")
    # launch shape on the kept placement
    out["autotune"] = e.autotune(None, steps=4)
    log(f"[ensemble] autotuned (block, cells/thread, grid): {out['autotune']}")
    res = {}
    for mode in ("graph", "persistent", "per_domain"):
        e.Initialize()
        e.step(2, mode=mode, moments_every=0)
        ms = e.step_timed(a.steps, mode=mode, moments_every=0)
        res[mode] = {"glups": cells * a.steps / (ms * 1e-3) / 1e9, "ms_per_step": ms / a.steps}
        log(f"[ensemble] {mode}: {res[mode]['glups']:.2f} GLUPS")
    out["modes"] = res
    out["graph_vs_per_domain"] = res["graph"]["glups"] / res["per_domain"]["glups"]
    # per-domain error: same ensemble in FP64, same number of steps
    e.Initialize()
    e.step(a.steps, mode="graph")
    ref = Ensemble(doms, dtype="f64", storage=a.storage)
    ref.Initialize()
    ref.step(a.steps, mode="graph")
    errs, fin64, fin32 = [], [], []
    for k in range(len(doms)):
        m32 = e.moments(k)
        m64 = ref.moments(k)
        fin64.append(all(np.isfinite(x).all() for x in m64))
        fin32.append(all(np.isfinite(x).all() for x in m32))
        errs.append(l2(m32[1:], m64[1:]) if fin64[-1] and fin32[-1] else float("nan"))
    errs, fin64, fin32 = np.array(errs), np.array(fin64), np.array(fin32)
    ok = fin64 & fin32
    # BGK is unstable at the smallest edges and highest Reynolds numbers of
    # the draw (tau0 -> 0); such domains blow up in FP64 as in FP32
    out["unstable_domains_f64"] = [int(k) for k in np.where(~fin64)[0]]
    out["unstable_domains_f32"] = [int(k) for k in np.where(~fin32)[0]]
    out["unstable_tau0"] = [doms[k].tau0 for k in np.where(~fin64)[0]]
    if ok.any():
        out["per_domain_vel_l2_vs_f64"] = {
            "domains": int(ok.sum()), "max": float(errs[ok].max()), "median": float(np.median(errs[ok])),
            "tgv_max": float(np.nanmax(np.where(ok[0::2], errs[0::2], np.nan))),
            "cavity_max": float(np.nanmax(np.where(ok[1::2], errs[1::2], np.nan)))}
    out["info"] = e.info()
    return out


def timing(a):
    n = a.n
    shape = tuple(int(v) for v in a.shape.split(",")) if a.shape else (n, n, n)
    tau = 0.05 * shape[0] / 1600.0 / (1.0 / 3.0) / 2 / PI_R
    g = LBM(*shape, tau, u0=0.05, dtype=a.dtype, storage=a.storage)
    if a.launch:
        g.set_launch(*(int(v) for v in a.launch.split(",")))
    g.Initialize()
    out = {"workload": f"TGV {'x'.join(map(str, shape))} Re 1600 {a.dtype} {a.storage}", "placement": None}
    if a.placements is None and not a.window:
        a.placements = 1  # timing keeps the first allocation unless asked
    out["placement"] = place(g, a, shape[0] * shape[1] * shape[2], "timing")
    if not a.launch:
        g.autotune(None, steps=8)  # on the kept placement
    g.step(4, moments_every=0)
    ms = g.step_timed(a.steps, moments_every=0)
    info = g.info()
    glups = shape[0] * shape[1] * shape[2] * a.steps / (ms * 1e-3) / 1e9
    out.update({"steps": a.steps, "ms_per_step": ms / a.steps, "glups": glups,
                "achieved_GBps": glups * info["bytes_per_lup"],
                "launch": [info[k] for k in ("block", "cells_per_thread", "xcd_remap", "nt_loads")]})
    log(f"[timing] {out['workload']}: {glups:.2f} GLUPS, {out['achieved_GBps']:.0f} GB/s")
    return out


def phased(a):
    n = a.n
    r = (512 - 1) / 64.0
    cases = {f"cavity {n}^3": dict(shape=(n, n, n), tau=0.096, u0=0.1, case="cavity", engine="phased"),
             "sphere 512x128x128 Re 100": dict(shape=(512, 128, 128), tau=0.05 * 2 * r / 100.0 / (1.0 / 3.0),
                                               u0=0.05, case="sphere", engine="phased"),
             f"cavity {n}^3 fused engine": dict(shape=(n, n, n), tau=0.096, u0=0.1, case="cavity", engine="fused"),
             "sphere 512x128x128 Re 100 fused engine": dict(shape=(512, 128, 128),
                                                            tau=0.05 * 2 * r / 100.0 / (1.0 / 3.0), u0=0.05,
                                                            case="sphere", engine="fused")}
    out = {"workload": "faithful mode, phased and fused engines", "runs": {}, "placement": {}}
    for name, c in cases.items():
        g = LBM(*c["shape"], c["tau"], u0=c["u0"], case=c["case"], mode="faithful", engine=c["engine"])
        g.Initialize()
        if c["engine"] == "fused":  # placement, then the launch shape on it (the phased engine keeps its layout)
            out["placement"][name] = place(g, a, c["shape"][0] * c["shape"][1] * c["shape"][2], f"phased {name}")
            g.autotune(None, steps=8)
        g.Initialize()
        g.step(3)
        ms = g.step_timed(a.steps)
        cells = c["shape"][0] * c["shape"][1] * c["shape"][2]
        out["runs"][name] = {"glups": cells * a.steps / (ms * 1e-3) / 1e9, "ms_per_step": ms / a.steps}
        log(f"[phased] {name}: {out['runs'][name]['glups']:.2f} GLUPS")
        g.close()
    return out


def placement(a):
    n = a.n
    u0, re = 0.05, 1600.0
    tau = u0 * n / re / (1.0 / 3.0) / 2 / PI_R
    cells = n ** 3

    def rate(m):
        return round(cells / (m * 1e-3) / 1e9, 2) if m > 0 else None
    out = {"workload": f"TGV {n}^3 {a.storage}", "runs": {}}
    for dt, launch, step in (("f64", (128, 1, 1, 1), 2), ("f32", (128, 1, 1, 1), 1), ("f16", (512, 1, 0, 1), 1)):
        if dt not in a.dtypes.split(","):
            continue
        g = LBM(n, n, n, tau, u0=u0, dtype=dt, storage=a.storage)
        g.set_launch(*launch)
        g.Initialize()
        for rep in range(a.windows):
            kept, ms = g.tune_placement_window(slack_gib=64, step_gib=step, steps=10)
            r = [rate(m) for m in ms]
            out["runs"][f"{dt} window {rep}"] = {"launch": launch, "step_gib": step, "kept": kept, "glups": r}
            log(f"[placement] {dt} window {rep}: kept offset {kept} ({r[kept]}), min {min(r)}, max {max(r)}")
        if a.allocations > 1:
            kept, ms = g.tune_placement(a.allocations, steps=10)
            r = [rate(m) for m in ms]
            out["runs"][f"{dt} allocations"] = {"launch": launch, "kept": kept, "glups": r}
            log(f"[placement] {dt} separate allocations: {r}")
        g.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["precision", "cavity", "sphere", "ensemble", "timing", "phased",
                                         "placement"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--domains", type=int, default=256)
    ap.add_argument("--dtype", default="f16", choices=["f64", "f32", "f16"])
    ap.add_argument("--storage", default=None, choices=["aa", "pull"],
                    help="population storage (default: pull for the ensemble, AA otherwise)")
    ap.add_argument("--launch", default=None, help="block,cells_per_thread,xcd_order,nt_loads")
    ap.add_argument("--placements", type=int, default=None,
                    help="separate lattice allocations to try instead of the window search (1: the first "
                         "allocation; timing defaults to 1)")
    ap.add_argument("--shape", default=None, help="nx,ny,nz for timing (default n,n,n)")
    ap.add_argument("--window", action="store_true",
                    help="timing: place the lattice with the window search (default: the first allocation)")
    ap.add_argument("--dtypes", default="f64,f32,f16", help="placement: storage types to search")
    ap.add_argument("--windows", type=int, default=2, help="placement: window searches per storage type")
    ap.add_argument("--allocations", type=int, default=4,
                    help="placement: separate allocations timed after the windows (1: none)")
    a = ap.parse_args(argv)
    defaults = {"precision": (512, 5000), "cavity": (128, 20000), "sphere": (0, 50000), "ensemble": (0, 2000),
                "timing": (512, 20), "phased": (256, 20), "placement": (512, 10)}
    dn, ds = defaults[a.workload]
    a.n = a.n or dn
    a.steps = a.steps or ds
    if a.storage is None:
        # the ensemble runs pull a little faster than AA (400 steps of the
        # 256-domain draw: 26.8-27.1 against 26.1-26.2 GLUPS)
        a.storage = "pull" if a.workload == "ensemble" else "aa"
    fn = {"precision": precision, "cavity": cavity, "sphere": sphere, "ensemble": ensemble, "timing": timing,
          "phased": phased, "placement": placement}[a.workload]
    out = fn(a)
    out["copy_GBps"] = copy_bandwidth(1 << 30, reps=5)[0]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
