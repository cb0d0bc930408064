#!/usr/bin/env python3
"""Throughput of the coupled D3Q19+D3Q7 MRT-LBM Rayleigh-Benard step on one MI355X.

Workload: 3D cubic cavity, 512^3 nodes, Ra=1e8, Pr=0.71, Ma=0.1, hot bottom / cold top
isothermal walls, adiabatic no-slip side walls, seeded initial T (linear profile + uniform
noise +-1e-3 dT), rho=1, u=0.  A "step" is one full time step of both lattices over the whole
domain; inputs are resident in HBM before the timed region.

Prints ONE JSON line:
  value      coupled GLUPS = nodes * steps / seconds / 1e9 (device time of the timed steps,
             HIP events on the solver's stream)
  roofline   algorithmic bytes per node update (464 B in FP64: both lattices read and written
             once plus the M1G history) * nodes / step time, against the 8 TB/s HBM3E peak.
             The unit is one step: the two launches of the two-pass form (k3_thermal, k3_flow)
             or the one launch of the one-pass form (k3_fused).  traffic = measured HBM bytes
             per step from rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a copy kernel's readings
             (profiles/traffic_fp64_512.json), or null when that file has no entry for the form;
             copy_gbs = a device copy measured in the same run (the practical ceiling)
  cpu_baseline  the FP64 OpenMP oracle (a restatement of the reference's CPU path) timed on
             this host over a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "-2d-3d-_amd"))

HBM_PEAK_GBS = 8000.0
SEED = 20261203


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--rayleigh", type=float, default=1e8)
    ap.add_argument("--forms", default="fp64,fp32",
                    help="storage forms to time, comma separated; the first one is `value`")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=48,
                    help="timed oracle steps of the CPU baseline (about 10 s on 16 host cores)")
    ap.add_argument("--cpu-nz", type=int, default=64)
    ap.add_argument("--variant", choices=["auto", "two-pass", "one-pass"], default="auto",
                    help="kernel form of the step (auto: the library's choice, the one-pass form "
                         "for FP64 and the two-pass form for FP32 / FP16 at 512^3)")
    ap.add_argument("--tile", type=int, default=-1,
                    help="one-pass thread tile (0: 32x16, 1: 16x16, 2: 32x8, 3: 32x32, 4: 34x15; "
                         "-1: the library's default, 34x15)")
    ap.add_argument("--prefetch", type=int, default=-1,
                    help="one-pass prefetching (0/1; -1: the library's default, 1)")
    ap.add_argument("--kchunk", type=int, default=0,
                    help="z-planes per one-pass block (0: default, -1: from the resident blocks)")
    ap.add_argument("--layout", choices=["two-array", "aa", "auto", "tuned"], default="two-array",
                    help="population layout (default two-array, with the default two-pass "
                         "form); aa: one array per lattice updated in place (two-pass form only); "
                         "tuned: an LBM_LAYOUT_AUTO domain whose layout and kernel form "
                         "lbm_autotune picks before the first step (--variant, --kchunk, --tile, "
                         "--prefetch are then ignored)")
    ap.add_argument("--tune-steps", type=int, default=8, help="timed steps per autotune candidate")
    ap.add_argument("--full-config", choices=["c1", "c2"], default="",
                    help="instead of the benchmark: the whole 2D 257x129 Ra=1e5 run (20,000 "
                         "steps) or the whole 128^3 Ra=1e6 FP64 run (5,000 steps), timed on the "
                         "GPU and, as the CPU baseline, the FP64 OpenMP oracle over the same run; "
                         "reports both and the largest difference of f, g, u, T and Nu")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run a device-to-device copy of GIB GiB (counter calibration)")
    return ap.parse_args()


def copy_probe(gib):
    """A plain copy kernel of known size: reads and writes gib GiB once per launch."""
    import torch

    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "launches": 3}), flush=True)


def cpu_baseline(size, nz, ra, steps):
    """Time the oracle on a 'size x size x nz' slab of the same cavity (same parameters)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    # an OMP_NUM_THREADS set by the environment (the host's CPU share) takes precedence
    cores = int(os.environ.get("OMP_NUM_THREADS", cores))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ.setdefault("OMP_PROC_BIND", "true")
    os.environ.setdefault("OMP_PLACES", "cores")
    import orc  # the checker / baseline only; never the measured GPU path

    cores = orc.set_threads(cores)
    cfg = orc.Config(dims=3, nx=size, ny=size, nz=nz, rayleigh=ra)
    o = orc.Oracle(cfg)
    o.init_seeded(SEED, 1e-3)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    n = size * size * nz
    o.close()
    return {"value": n * steps / dt / 1e9, "unit": "GLUPS", "cores": cores, "kind": "port",
            "sample": f"{size}x{size}x{nz} slab of the {size}^3 Ra={ra:g} cavity, FP64 OpenMP "
                      f"oracle, {steps} timed steps after 1 warm-up step ({dt:.1f} s)"}


def full_config(name):
    """One whole benchmark configuration on the GPU and, as the CPU baseline, on the oracle."""
    import numpy as np

    import rblbm
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # the checker / baseline only; never the measured GPU path
    if name == "c1":
        kw = dict(dims=2, nx=257, ny=129, nz=1, rayleigh=1e5, prandtl=0.71, thot=1.0, tcold=0.0,
                  vel_x=rblbm.VEL_PERIODIC, temp_x=rblbm.T_PERIODIC)
        steps, label = 20000, "2D RB 257x129 Ra=1e5 Pr=0.71, periodic x, FP64, 20,000 steps"
    else:
        kw = dict(dims=3, nx=128, ny=128, nz=128, rayleigh=1e6, prandtl=0.71)
        steps, label = 5000, "3D RB cavity 128^3 Ra=1e6 Pr=0.71, FP64, 5,000 steps"
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = orc.set_threads(int(os.environ.get("OMP_NUM_THREADS", cores)))
    sol = rblbm.Solver(rblbm.Config(**kw))
    sol.initial_seeded(SEED, 1e-3)
    sol.step_timed(0)                       # the prologue outside the timing
    ms = sol.step_timed(steps)
    m, st, d = sol.macros(), sol.gather(with_bprev=False), sol.diagnostics()
    sol.close()
    o = orc.Oracle(orc.Config(**kw))
    o.init_seeded(SEED, 1e-3)
    t0 = time.perf_counter()
    o.step(steps)
    cpu_s = time.perf_counter() - t0
    s, e = o.state(), o.diagnostics()
    o.close()
    n = kw["nx"] * kw["ny"] * kw["nz"]
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return {"config": label, "nodes": n, "steps": steps,
            "gpu": {"seconds": ms / 1e3, "glups": n * steps / (ms / 1e3) / 1e9},
            "cpu_baseline": {"seconds": cpu_s, "glups": n * steps / cpu_s / 1e9, "cores": cores,
                             "kind": "port"},
            "max_rel_diff": {"f": rel(st["f"], s["f"]), "g": rel(st["g"], s["g"]),
                             "u": rel(m["u"], s["u"]), "T": rel(m["T"], s["T"]),
                             "nu_vol": abs(d["nu_vol"] - e["nu_vol"]) / abs(e["nu_vol"]),
                             "re_vol": abs(d["re_vol"] - e["re_vol"]) / abs(e["re_vol"])},
            "nu_vol": d["nu_vol"], "re_vol": d["re_vol"], "nu_global": d["nu_global"]}


def copy_gbs(torch, gib=2.0, reps=5):
    """Bandwidth of a device-to-device copy (read + write) of `gib` GiB: the practical ceiling."""
    if torch is None:
        return None
    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * n * 8 / (ms / 1e3) / 1e9


def measure(rblbm, torch, n, ra, prec, steps, warmup, variant=0, kchunk=0, tile=-1, prefetch=-1,
            layout=0, tune_steps=0, keep=None):
    """Time `steps` steps of one storage form on the n^3 cavity (inputs resident in HBM).
    tune_steps > 0: lbm_autotune picks the layout and kernel form first (outside the timing)."""
    cfg = rblbm.Config(dims=3, nx=n, ny=n, nz=n, rayleigh=ra, prandtl=0.71, mach=0.1,
                       precision=prec, layout=layout)
    sol = rblbm.Solver(cfg)
    tune_s = None
    if tune_steps > 0:
        t0 = time.perf_counter()
        sol.autotune(None, steps=tune_steps)
        tune_s = time.perf_counter() - t0
    else:
        sol.set_kernel_variant(variant, kchunk)
        if tile >= 0 or prefetch >= 0:
            f0 = sol.kernel_form()
            sol.set_fused_tile(tile if tile >= 0 else f0["tile"],
                               prefetch if prefetch >= 0 else f0["prefetch"])
    form = sol.kernel_form()
    aa = sol.layout == rblbm.LAYOUT_AA
    sol.initial_seeded(SEED, 1e-3)
    mass0 = sol.diagnostics()["mass"]
    sol.step(warmup)
    sol.sync()
    if torch is not None:
        torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ms = sol.step_timed(steps)            # HIP events on the solver's own stream
    if torch is not None:
        torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    nodes = n ** 3
    ms_step = ms / steps
    bpu = sol.bytes_per_update()
    diag = sol.diagnostics()
    # the domain stays allocated until every form has run, so that a later form's domain does
    # not reuse the memory this process allocated first (which streams up to 10% slower)
    if keep is not None:
        keep.append(sol)
    else:
        sol.close()
    one = form["variant"] == rblbm.KERNEL_ONE_PASS
    form["kernels"] = ("k3_fused (one-pass step)" if one else
                       "k3_thermal + k3_flow (two-pass step" + (", in place)" if aa else ")"))
    form["name"] = "one-pass" if one else ("two-pass-aa" if aa else "two-pass")
    form["layout"] = "aa" if aa else "two-array"
    if tune_s is not None:
        form["chosen_by"] = f"lbm_autotune ({tune_s:.1f} s)"
    return {"glups": nodes * steps / (ms / 1e3) / 1e9, "ms_per_step": ms_step, "form": form,
            "bytes_per_update": bpu, "achieved_gbs": bpu * nodes / (ms_step / 1e3) / 1e9,
            "roofline_frac": bpu * nodes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "wall_seconds": t_wall, "nu_vol": diag["nu_vol"], "diag": diag, "mass0": mass0,
            "mass_rel_change": (diag["mass"] - mass0) / mass0}


def main():
    a = parse()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.full_config:
        print(json.dumps(full_config(a.full_config)), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {a.gpus}); the solver runs a single "
              "domain per device and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import rblbm

    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:
        torch, have_torch = None, False

    n = a.size
    nodes = n * n * n
    forms, open_domains = {}, []
    for name in a.forms.split(","):
        prec = {"fp64": rblbm.FP64, "fp32": rblbm.FP32_SHIFT, "fp16": rblbm.FP16_SHIFT}[name]
        variant = {"auto": rblbm.KERNEL_AUTO, "two-pass": rblbm.KERNEL_TWO_PASS,
                   "one-pass": rblbm.KERNEL_ONE_PASS}[a.variant]
        lname = "two-array" if a.variant == "one-pass" and a.layout != "tuned" else a.layout
        layout = {"two-array": rblbm.LAYOUT_TWO_ARRAY, "aa": rblbm.LAYOUT_AA,
                  "auto": rblbm.LAYOUT_AUTO, "tuned": rblbm.LAYOUT_AUTO}[lname]
        forms[name] = measure(rblbm, torch if have_torch else None, n, a.rayleigh, prec,
                              a.steps, a.warmup, variant, a.kchunk, a.tile, a.prefetch, layout,
                              a.tune_steps if lname == "tuned" else 0, open_domains)
    for sol in open_domains:
        sol.close()
    main_form = forms[a.forms.split(",")[0]]
    ms_step = main_form["ms_per_step"]
    glups = main_form["glups"]
    bpu = main_form["bytes_per_update"]
    achieved = main_form["achieved_gbs"]
    t_wall = main_form["wall_seconds"]
    diag = main_form["diag"]
    mass0 = main_form["mass0"]
    if "fp64" in forms and "fp32" in forms:
        forms["fp32"]["glups_vs_fp64"] = forms["fp32"]["glups"] / forms["fp64"]["glups"]
    for f in forms.values():
        f.pop("diag", None)

    main_name = a.forms.split(",")[0]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_fp64_512.json")
    if os.path.exists(tpath) and n == 512 and main_name == "fp64":
        with open(tpath) as fh:
            entry = json.load(fh).get(main_form["form"]["name"])
        if entry:
            traffic = entry["hbm_bytes_per_step"]
    ceiling = copy_gbs(torch if have_torch else None)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(n, min(a.cpu_nz, n), a.rayleigh, a.cpu_steps)

    out = {
        "metric": "coupled D3Q19+D3Q7 GLUPS & %HBM roofline, FP32-shift vs FP64 at 512^3, Nu err",
        "value": glups, "unit": "GLUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16/f32"}[a.forms.split(",")[0]],
        "data": "synthetic",
        "config": {"workload": f"3D RB cavity {n}^3 Ra={a.rayleigh:g} Pr=0.71 Ma=0.1, "
                               f"{main_name} storage form, seeded T noise 1e-3",
                   "nodes": nodes, "precision": main_name, "kernel_form": main_form["form"],
                   "wall_seconds": t_wall},
        "forms": forms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_unit": "bytes per step", "bytes_per_update": bpu,
                     "algorithmic_bytes_per_step": bpu * nodes,
                     "kernel": main_form["form"]["kernels"], "copy_gbs": ceiling,
                     "frac_of_copy": achieved / ceiling if ceiling else None},
        "cpu_baseline": cpu,
        "checks": {"nu_vol": diag["nu_vol"], "re_vol": diag["re_vol"],
                   "mass_rel_change": (diag["mass"] - mass0) / mass0},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
