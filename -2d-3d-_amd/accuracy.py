#!/usr/bin/env python3
"""Accuracy of the reduced-precision storage forms against the FP64 form on one GPU.

Runs the same seeded cavity in the FP64 form and in each requested reduced-precision form,
samples the diagnostics every `--every` steps, and reports for each form:
  * time-averaged Nu_vol, Nu_global, Nu_hot, Re_vol over the samples (all samples, and the
    samples from --avg-from on) and their relative error against the FP64 run,
  * relative L2 error of u and T against the FP64 run at the end,
  * throughput of each run (GLUPS over the whole run, diagnostics included).

The FP64 form is bitwise identical to the reference's CPU path (tests/test_gpu_parity.py), so
it stands in for the CPU oracle at sizes the CPU cannot run.

Example (512^3, Ra=1e8, 20,000 steps):
    python ./-2d-3d-_amd/accuracy.py --n 512 --ra 1e8 --steps 20000 --every 500 --forms fp32
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import rblbm  # noqa: E402

FORMS = {"fp64": rblbm.FP64, "fp32": rblbm.FP32_SHIFT, "fp16": rblbm.FP16_SHIFT}
KEYS = ("nu_vol", "nu_global", "nu_hot", "nu_cold", "re_vol")


LAYOUTS = {"two-array": rblbm.LAYOUT_TWO_ARRAY, "aa": rblbm.LAYOUT_AA, "auto": rblbm.LAYOUT_AUTO}


def run(form, n, ra, steps, every, seed, keep_fields, mach, layout="auto"):
    cfg = rblbm.Config(dims=3, nx=n, ny=n, nz=n, rayleigh=ra, mach=mach, precision=FORMS[form],
                       layout=LAYOUTS[layout])
    sol = rblbm.Solver(cfg)
    sol.initial_seeded(seed, 1e-3)
    series = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(every, steps - done)
        sol.step(k)
        done += k
        d = sol.diagnostics()
        series.append({"itc": done, **{key: d[key] for key in KEYS}})
        print(f"{form} {done}/{steps} Nu_vol={d['nu_vol']:.6f} Re={d['re_vol']:.3f}", flush=True)
    sol.sync()
    wall = time.perf_counter() - t0
    fields = None
    if keep_fields:
        m = sol.macros()
        fields = {k: m[k].astype(np.float64) for k in ("u", "v", "w", "T")}
    sol.close()
    return series, fields, n ** 3 * steps / wall / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ra", type=float, default=1e8)
    ap.add_argument("--mach", type=float, default=0.1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--avg-from", type=int, default=-1,
                    help="first step of the second set of time averages (default: half the run)")
    ap.add_argument("--forms", default="fp32")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--out", default="")
    ap.add_argument("--layout", choices=sorted(LAYOUTS), default="auto",
                    help="population layout of every run (auto: two arrays when they fit)")
    ap.add_argument("--ref-form", choices=sorted(FORMS), default="fp64",
                    help="form of the reference run (fp32 at the same edge isolates the "
                         "storage precision where FP64 does not fit)")
    ap.add_argument("--ref-n", type=int, default=0,
                    help="run the FP64 reference at this (coarser) edge; fields are compared "
                         "after averaging the fine run over blocks of (n/ref_n)^3 nodes")
    a = ap.parse_args()

    if a.avg_from < 0:
        a.avg_from = a.steps // 2
    rk = a.ref_form   # key of the reference run in the result
    assert rk not in a.forms.split(","), "the reference form is not compared with itself"
    ref_n = a.ref_n or a.n
    ratio = a.n // ref_n
    assert ref_n * ratio == a.n, "--n must be a multiple of --ref-n"
    # the same physical time: the coarse run takes n/ref_n fewer steps per free-fall time
    ref_steps, ref_every = a.steps // ratio, max(1, a.every // ratio)
    ref, ref_fields, ref_glups = run(a.ref_form, ref_n, a.ra, ref_steps, ref_every, a.seed, True,
                                     a.mach, a.layout)
    for s in ref:
        s["itc"] *= ratio
    result = {"workload": f"3D RB cavity {a.n}^3 Ra={a.ra:g} Pr=0.71 Ma={a.mach}",
              "reference": f"{a.ref_form.upper()} at {ref_n}^3, {ref_steps} steps",
              "steps": a.steps, "sample_every": a.every, "avg_from": a.avg_from,
              rk: {"glups_incl_diag": ref_glups, "series": ref}}

    def coarsen(x):
        if ratio == 1:
            return x
        nz, ny, nx = x.shape
        return x.reshape(nz // ratio, ratio, ny // ratio, ratio, nx // ratio, ratio).mean(
            axis=(1, 3, 5))

    def tavg(series, key, start=0):
        v = [s[key] for s in series if s["itc"] >= start]
        return float(np.mean(v))

    for k in KEYS:
        result[rk][f"tavg_{k}"] = tavg(ref, k)
        result[rk][f"tavg2_{k}"] = tavg(ref, k, a.avg_from)
    for form in a.forms.split(","):
        series, fields, glups = run(form, a.n, a.ra, a.steps, a.every, a.seed, True, a.mach,
                                    a.layout)
        fields = {k: coarsen(v) for k, v in fields.items()}
        if ratio > 1:   # samples at the coarse run's times only
            times = {s["itc"] for s in ref}
            series = [s for s in series if s["itc"] in times]
        r = {"glups_incl_diag": glups, "series": series}
        for k in KEYS:
            for tag, start in (("tavg", 0), ("tavg2", a.avg_from)):
                r[f"{tag}_{k}"] = tavg(series, k, start)
                r[f"{tag}_{k}_relerr"] = abs(r[f"{tag}_{k}"] - result[rk][f"{tag}_{k}"]) / \
                    abs(result[rk][f"{tag}_{k}"])
        num = sum(np.sum((fields[k] - ref_fields[k]) ** 2) for k in ("u", "v", "w"))
        den = sum(np.sum(ref_fields[k] ** 2) for k in ("u", "v", "w"))
        r["l2_u_relerr"] = float(np.sqrt(num / den))
        r["l2_T_relerr"] = float(np.sqrt(np.sum((fields["T"] - ref_fields["T"]) ** 2) /
                                         np.sum(ref_fields["T"] ** 2)))
        result[form] = r
        print(json.dumps({form: {k: v for k, v in r.items() if k != "series"}}), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(result, fh, indent=1)
    summary = {f: {k: v for k, v in result[f].items() if k != "series"}
               for f in [rk] + a.forms.split(",")}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
