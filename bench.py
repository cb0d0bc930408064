#!/usr/bin/env python3
"""Throughput of the coupled D3Q19+D3Q7 MRT-LBM Rayleigh-Benard step on one MI355X.

Workload: 3D cubic cavity, 512^3 nodes, Ra=1e8, Pr=0.71, Ma=0.1, hot bottom / cold top
isothermal walls, adiabatic no-slip side walls, seeded initial T (linear profile + uniform
noise +-1e-3 dT), rho=1, u=0.  A "step" is one full time step of both lattices over the whole
domain; inputs are resident in HBM before the timed region.

Prints ONE JSON line:
  value      FP64 coupled GLUPS = nodes * steps / seconds / 1e9 (device time of the timed steps,
             HIP events on the solver's stream)
  forms      the same for every storage form timed (default FP64, FP32-shift, FP16-shift), with
             glups_vs_fp64, and accuracy: the relative differences of Nu_vol, Nu_hot, Nu_cold
             and Re between each reduced form and FP64 at the end of the timed window (same
             seeded start, same steps); the 20,000-step time-averaged study is
             profiles/accuracy_512_ra1e8.md
  roofline   algorithmic bytes per node update (464 B in FP64: both lattices read and written
             once plus the M1G history) * nodes / step time, against the 8 TB/s HBM3E peak.  The
             unit is one step: the one launch of the one-pass form (k3_fused) or the two launches
             of the two-pass form.  traffic = measured HBM bytes per step (rocprofv3 FETCH_SIZE
             and WRITE_SIZE, each scaled by its reading for a copy kernel of known size;
             profiles/traffic_fp64_512.json), null unless that summary was measured on the
             kernel sources of this tree; copy_gbs / read_gbs = the hand-written streaming copy
             and read kernels of hbm_probe.hip in the same run (the practical ceiling);
             tile_copy_gbs = a copy of 29 FP64 arrays of the same grid in the one-pass step's
             access pattern (32 x 13 tiles marching 32-plane chunks), which moves the FP64
             step's algorithmic bytes with no arithmetic
  configs    the other benchmark configurations' rates: C5 (1024^3 Ra=1e10 FP16-shift) and C4
             (the seeded 64-cavity FP32-shift ensemble: HIP-graph table launches, in the
             default forms and in the forms lbm_batch_autotune picks per shape, against one
             launch per cavity)
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
    ap.add_argument("--forms", default="fp64,fp32,fp16",
                    help="storage forms to time, comma separated; the first one is `value`")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the C5 (1024^3 FP16) and C4 (64-cavity ensemble) legs")
    ap.add_argument("--c5-steps", type=int, default=10)
    ap.add_argument("--c4-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=48,
                    help="timed oracle steps of the CPU baseline (about 10 s on 16 host cores)")
    ap.add_argument("--cpu-nz", type=int, default=64)
    ap.add_argument("--variant", choices=["auto", "two-pass", "one-pass"], default="auto",
                    help="kernel form of the step (auto: the library's choice, the one-pass form "
                         "for FP64 and FP32 and the two-pass form for FP16 at 512^3)")
    ap.add_argument("--tile", type=int, default=-1,
                    help="one-pass thread tile (0: 32x16, 1: 16x16, 2: 32x8, 3: 32x32, 4: 34x15, "
                         "5: 66x15; -1: the library's default, 34x15 in FP64 and FP16, 66x15 in "
                         "FP32)")
    ap.add_argument("--prefetch", type=int, default=-1,
                    help="one-pass prefetching (0: none, 1: into registers, 2: staged in LDS, "
                         "FP32 only; -1: the library's default, 1 in FP64 and FP16, 2 in FP32)")
    ap.add_argument("--kchunk", type=int, default=0,
                    help="z-planes per one-pass block (0 or -1: the library's default, 32)")
    ap.add_argument("--layout", choices=["two-array", "aa", "auto", "tuned"], default="two-array",
                    help="population layout (default two-array, with the library's default "
                         "kernel form); aa: one array per lattice updated in place (two-pass form "
                         "only); "
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
                    help="only run the probe's copy and read kernels on GIB GiB, 3 launches each "
                         "(counter calibration)")
    ap.add_argument("--traffic-summary", default="", metavar="DIR",
                    help="only write profiles/traffic_fp64_512.json from the rocprofv3 --pmc "
                         "runs under DIR (fetch/, write/: the bench; fetch_TAG/, write_TAG/: "
                         "optional further bench runs; copy_fetch/, copy_write/: the copy probe)")
    return ap.parse_args()


PROBE_LIB = os.path.join(ROOT, "-2d-3d-_amd", "lib", "libhbmprobe.so")
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_fp64_512.json")


def hbm_probe(gib=4.0, reps=5):
    """Streaming copy and read kernels (hbm_probe.hip) over gib GiB: (copy GB/s, read GB/s)."""
    import ctypes as C

    lib = C.CDLL(PROBE_LIB)
    lib.hbm_probe.argtypes = [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    cp, rd = C.c_double(), C.c_double()
    rc = lib.hbm_probe(0, int(gib * (1 << 30)), reps, C.byref(cp), C.byref(rd))
    if rc != 0:
        raise RuntimeError(f"hbm_probe failed: HIP error {rc}")
    return cp.value, rd.value


def hbm_probe_tile(n, reps=3):
    """The step's access pattern without arithmetic (hbm_probe_tile): GB/s over an n^3 grid."""
    import ctypes as C

    lib = C.CDLL(PROBE_LIB)
    lib.hbm_probe_tile.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    g = C.c_double()
    rc = lib.hbm_probe_tile(0, n, n, n, reps, C.byref(g))
    if rc != 0:   # e.g. 2 x 29 arrays of n^3 doubles do not fit the device
        print(f"bench: tile-pattern copy not measured (HIP error {rc})", file=sys.stderr, flush=True)
        return None
    return g.value


def copy_probe(gib):
    """The probe's copy and read kernels of known size (3 launches each) for the counters."""
    n = int(gib * (1 << 30)) // 16 * 16
    hbm_probe(gib, 3)
    print(json.dumps({"probe_bytes": n, "copy": "k_copy reads and writes probe_bytes a launch",
                      "read": "k_read reads probe_bytes a launch", "launches": 3}), flush=True)


def kernel_sha():
    """Hash of the kernel and launch sources, the key of the measured traffic summary."""
    import hashlib

    d = os.path.join(ROOT, "-2d-3d-_amd", "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".h")) and f != "hbm_probe.hip":
            with open(os.path.join(d, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def _pmc(path, counter):
    """Per-kernel mean of a rocprofv3 --pmc counter (KiB per dispatch) from a run directory."""
    import csv
    import glob

    out = {}
    for f in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] == counter:
                out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in out.items()}


# The kernels of one steady step per (storage form, kernel form), as rocprofv3 names them.
STEP_KERNELS = {
    ("fp64", "one-pass"): ["k3_fused<rblbm::FormF64"],
    ("fp64", "two-pass"): ["k3_thermal<false, 0>", "k3_flow<0>"],
    ("fp32", "one-pass"): ["k3_fused<rblbm::FormLP<float>"],
    ("fp32", "two-pass"): ["lp3_thermal<float, false, 0>", "lp3_flow<float, 0>"],
    # _Float16 instances reach the counter files with their mangled names
    ("fp16", "one-pass"): ["k3_fused<rblbm::FormLP<_Float16>", "k3_fusedINS_6FormLPIDF16_E"],
    ("fp16", "two-pass"): ["lp3_thermal<_Float16, false, 0>", "lp3_flow<_Float16, 0>",
                           "lp3_thermalIDF16_Lb0ELi0E", "lp3_flowIDF16_Li0E"],
}


def traffic_summary(d):
    """profiles/traffic_fp64_512.json from the counter runs under d (see --traffic-summary)."""
    cf, cw = _pmc(os.path.join(d, "copy_fetch"), "FETCH_SIZE"), _pmc(os.path.join(d, "copy_write"), "WRITE_SIZE")
    probe = [json.loads(l) for l in open(os.path.join(d, "copy_fetch.log")) if l.startswith("{")][-1]
    nbytes = probe["probe_bytes"]
    kc = [k for k in cf if "k_copy" in k][0]
    f_fetch = nbytes / (cf[kc] * 1024.0)
    f_write = nbytes / (cw[kc] * 1024.0)
    out = {"kernel_sha": kernel_sha(), "nodes": None,
           "copy_calibration": {"probe_bytes": nbytes, "fetch_kib": cf[kc], "write_kib": cw[kc],
                                "fetch_factor": f_fetch, "write_factor": f_write},
           "forms": {}}
    # fetch/, write/: the default bench run; fetch_<tag>/, write_<tag>/ (optional): further
    # bench runs, e.g. with --variant two-pass
    tags = [""] + sorted(n[len("fetch_"):] for n in os.listdir(d)
                         if n.startswith("fetch_") and os.path.isdir(os.path.join(d, n)))
    for tag in tags:
        sfx = "_" + tag if tag else ""
        fetch = _pmc(os.path.join(d, "fetch" + sfx), "FETCH_SIZE")
        write = _pmc(os.path.join(d, "write" + sfx), "WRITE_SIZE")
        bench = [json.loads(l) for l in open(os.path.join(d, f"fetch{sfx}.log")) if l.startswith("{")][-1]
        nodes = out["nodes"] = bench["config"]["nodes"]
        for name, f in bench["forms"].items():
            key = f"{name}/{f['form']['name']}"
            kern = [k for k in fetch if any(p in k for p in STEP_KERNELS[(name, f["form"]["name"])])]
            if not kern or key in out["forms"]:
                continue
            fb = sum(fetch[k] for k in kern) * 1024.0 * f_fetch
            wb = sum(write[k] for k in kern) * 1024.0 * f_write
            out["forms"][key] = {
                "kernel": " + ".join(k.split("(")[0].replace("void ", "") for k in kern),
                "fetch_bytes_per_step": fb, "write_bytes_per_step": wb, "hbm_bytes_per_step": fb + wb,
                "algorithmic_bytes_per_step": f["bytes_per_update"] * nodes,
                "ratio": (fb + wb) / (f["bytes_per_update"] * nodes)}
    with open(TRAFFIC, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out), flush=True)


def measured_traffic(name, form_name):
    """HBM bytes per step of the summary, or None when it was measured on other sources."""
    if not os.path.exists(TRAFFIC):
        return None, "no traffic summary"
    with open(TRAFFIC) as fh:
        t = json.load(fh)
    if t.get("kernel_sha") != kernel_sha():
        return None, f"summary measured on kernel sources {t.get('kernel_sha')}, tree has {kernel_sha()}"
    e = t.get("forms", {}).get(f"{name}/{form_name}")
    if not e:
        return None, f"no entry for {name}/{form_name}"
    return e["hbm_bytes_per_step"], f"profiles/traffic_fp64_512.json ({e['kernel']})"


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


def measure(rblbm, n, ra, prec, steps, warmup, variant=0, kchunk=0, tile=-1, prefetch=-1,
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
    t_wall = time.perf_counter()
    ms = sol.step_timed(steps)            # HIP events on the solver's own stream
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
            "wall_seconds": t_wall, "diag": diag, "mass0": mass0,
            "mass_rel_change": (diag["mass"] - mass0) / mass0}


def c5_rate(rblbm, steps):
    """C5: 1024^3 Ra=1e10 FP16-shift, the library's layout choice (two arrays when they fit)."""
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=1024, ny=1024, nz=1024, rayleigh=1e10,
                                    prandtl=0.71, mach=0.1, precision=rblbm.FP16_SHIFT,
                                    layout=rblbm.LAYOUT_AUTO))
    form = sol.kernel_form()
    lay = "aa" if sol.layout == rblbm.LAYOUT_AA else "two-array"
    sol.initial_seeded(SEED, 1e-3)
    sol.step(2)
    sol.sync()
    ms = sol.step_timed(steps)
    d = sol.diagnostics()
    sol.close()
    nodes = 1024 ** 3
    return {"workload": "3D RB cavity 1024^3 Ra=1e10 Pr=0.71 Ma=0.1, FP16-shift storage",
            "glups": nodes * steps / (ms / 1e3) / 1e9, "ms_per_step": ms / steps, "steps": steps,
            "layout": lay, "kernel_form": "one-pass" if form["variant"] == rblbm.KERNEL_ONE_PASS else "two-pass",
            "bytes_per_update": 128.0,
            "roofline_frac": 128.0 * nodes / (ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS,
            "nu_vol": d["nu_vol"]}


def c4_rate(rblbm, steps):
    """C4: the seeded 64-cavity FP32-shift ensemble; graph-replayed table launches against one
    launch per cavity (the same steps on the same cavities)."""
    sys.path.insert(0, os.path.join(ROOT, "-2d-3d-_amd"))
    import ensemble

    cfgs = ensemble.ensemble(64, SEED, rblbm.FP32_SHIFT)
    b = rblbm.Batch(cfgs)
    b.initial_seeded(SEED, 1e-3)
    nodes = b.nodes
    out = {"workload": "64 cavities, edges {32,48,64,96,128}^3 seeded, Ra log-uniform [1e4,1e7], "
                       "Pr in {0.71,7}, FP32-shift", "nodes": nodes, "steps": steps}
    for name, mode in (("per_cavity", rblbm.BATCH_PER_DOMAIN), ("graph", rblbm.BATCH_GRAPH)):
        b.set_mode(mode)
        b.step(4)
        ms = b.step_timed(steps)
        out[name] = {"glups": nodes * steps / (ms / 1e3) / 1e9, "ms_per_step": ms / steps}
    # the autotuner's choice per shape (lbm_batch_autotune, no cache), then graph replay again
    t0 = time.perf_counter()
    b.autotune(None, steps=4)
    tune_s = time.perf_counter() - t0
    b.set_mode(rblbm.BATCH_GRAPH)
    b.step(4)
    ms = b.step_timed(steps)
    forms = [d.kernel_form() for d in b.domains]
    out["graph_autotuned"] = {
        "glups": nodes * steps / (ms / 1e3) / 1e9, "ms_per_step": ms / steps, "tune_seconds": tune_s,
        "one_pass_domains": sum(f["variant"] == rblbm.KERNEL_ONE_PASS for f in forms),
        "tile": forms[0]["tile"], "prefetch": forms[0]["prefetch"]}
    out["glups"] = out["graph_autotuned"]["glups"]
    out["graph_vs_per_cavity"] = out["graph_autotuned"]["glups"] / out["per_cavity"]["glups"]
    b.close()
    return out


def main():
    a = parse()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.traffic_summary:
        traffic_summary(a.traffic_summary)
        return
    if a.full_config:
        print(json.dumps(full_config(a.full_config)), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {a.gpus}); the solver runs a single "
              "domain per device and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import rblbm

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
        forms[name] = measure(rblbm, n, a.rayleigh, prec, a.steps, a.warmup, variant, a.kchunk,
                              a.tile, a.prefetch, layout, a.tune_steps if lname == "tuned" else 0,
                              open_domains)
        print(f"bench: {name} {forms[name]['ms_per_step']:.3f} ms per step", file=sys.stderr, flush=True)
    for sol in open_domains:
        sol.close()
    main_name = a.forms.split(",")[0]
    main_form = forms[main_name]
    diag, mass0 = main_form["diag"], main_form["mass0"]
    # reduced forms against FP64 at the end of the timed window (same seeded start and steps)
    if "fp64" in forms:
        ref = forms["fp64"]["diag"]
        for name, f in forms.items():
            if name == "fp64":
                continue
            f["glups_vs_fp64"] = f["glups"] / forms["fp64"]["glups"]
            f["accuracy_vs_fp64"] = {
                k: abs(f["diag"][k] - ref[k]) / abs(ref[k]) for k in ("nu_vol", "nu_hot", "nu_cold", "re_vol")}
            f["accuracy_vs_fp64"]["after_steps"] = a.warmup + a.steps
            f["accuracy_vs_fp64"]["time_averaged_study"] = "profiles/accuracy_512_ra1e8.md"
    for f in forms.values():
        f["nu_vol"] = f.pop("diag")["nu_vol"]

    copy_gbs, read_gbs = hbm_probe()
    tile_gbs = hbm_probe_tile(n)
    traffic, traffic_src = (None, "not measured at this size / form")
    if n == 512:
        traffic, traffic_src = measured_traffic(main_name, main_form["form"]["name"])
    configs = {}
    if not a.no_extras:
        configs["c5_1024_fp16"] = c5_rate(rblbm, a.c5_steps)
        print(f"bench: C5 {configs['c5_1024_fp16']['ms_per_step']:.3f} ms per step", file=sys.stderr, flush=True)
        configs["c4_ensemble_fp32"] = c4_rate(rblbm, a.c4_steps)
        print(f"bench: C4 {configs['c4_ensemble_fp32']['glups']:.2f} GLUPS", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(n, min(a.cpu_nz, n), a.rayleigh, a.cpu_steps)

    achieved = main_form["achieved_gbs"]
    bpu = main_form["bytes_per_update"]
    out = {
        "metric": "coupled D3Q19+D3Q7 GLUPS & %HBM roofline, FP32-shift vs FP64 at 512^3, Nu err",
        "value": main_form["glups"], "unit": "GLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": main_form["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16/f32"}[main_name],
        "data": "synthetic",
        "config": {"workload": f"3D RB cavity {n}^3 Ra={a.rayleigh:g} Pr=0.71 Ma=0.1, "
                               f"{main_name} storage form, seeded T noise 1e-3",
                   "nodes": nodes, "precision": main_name, "kernel_form": main_form["form"],
                   "wall_seconds": main_form["wall_seconds"]},
        "forms": forms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_unit": "bytes per step", "traffic_source": traffic_src,
                     "bytes_per_update": bpu, "algorithmic_bytes_per_step": bpu * nodes,
                     "kernel": main_form["form"]["kernels"], "copy_gbs": copy_gbs,
                     "read_gbs": read_gbs, "frac_of_copy": achieved / copy_gbs,
                     "tile_copy_gbs": tile_gbs,
                     "frac_of_tile_copy": achieved / tile_gbs if tile_gbs else None,
                     "kernel_sha": kernel_sha()},
        "configs": configs,
        "cpu_baseline": cpu,
        "checks": {"nu_vol": diag["nu_vol"], "re_vol": diag["re_vol"],
                   "mass_rel_change": (diag["mass"] - mass0) / mass0},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
