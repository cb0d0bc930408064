#!/usr/bin/env python3
"""Batched-ensemble throughput: many independent cavities on one GPU.

Ensemble: 64 cubic cavities, edge drawn uniformly (seeded) from {32, 48, 64, 96, 128},
Ra log-uniform in [1e4, 1e7], Pr in {0.71, 7.0}, FP32-shifted storage, seeded initial noise
(instance index in the noise key).  Aggregate GLUPS = sum of nodes * steps / device seconds,
timed for each kernel form (auto: the library's default, which runs the one-launch modes in
the one-pass form once the batch's total node count reaches 2^22 FP64 or 2^24 FP32-shift, and
each cavity in its own form in per_domain mode; two_pass; one_pass) and stepping mode of the
batch:
  per_domain   one launch per cavity per pass
  one_launch   one launch per pass for all cavities (per-cavity parameter/offset table)
  graph        one_launch, two steps captured in a HIP graph and replayed
and, with the diagnostics of every cavity every --every steps (the default form, wall-clock
time including the diagnostics' host work):
  diag_eager   lbm_batch_step + lbm_diagnostics per cavity (one synchronisation per cavity)
  diag_graph   lbm_batch_run_diagnostics (steps and reductions of a period replayed from a
               HIP graph; one synchronisation per run)
Prints one JSON line.
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


def ensemble(n, seed, precision):
    rng = np.random.default_rng(seed)
    edges = rng.choice([32, 48, 64, 96, 128], size=n)
    ras = 10.0 ** rng.uniform(4.0, 7.0, size=n)
    prs = rng.choice([0.71, 7.0], size=n)
    return [rblbm.Config(dims=3, nx=int(e), ny=int(e), nz=int(e), rayleigh=float(r),
                         prandtl=float(p), precision=precision)
            for e, r, p in zip(edges, ras, prs)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "fp32", "fp16"], default="fp32")
    ap.add_argument("--every", type=int, default=10, help="steps between diagnostics samples")
    a = ap.parse_args()
    prec = {"fp64": rblbm.FP64, "fp32": rblbm.FP32_SHIFT, "fp16": rblbm.FP16_SHIFT}[a.precision]
    cfgs = ensemble(a.n, a.seed, prec)
    b = rblbm.Batch(cfgs)
    b.initial_seeded(a.seed, 1e-3)
    nodes = b.nodes
    out = {"workload": f"{a.n} cavities, edges {sorted(set(c.nx for c in cfgs))}, "
                       f"Ra log-uniform [1e4,1e7], Pr in {{0.71,7}}, {a.precision}",
           "nodes": nodes, "steps": a.steps, "modes": {}}
    for form, variant in (("auto", rblbm.KERNEL_AUTO), ("two_pass", rblbm.KERNEL_TWO_PASS),
                          ("one_pass", rblbm.KERNEL_ONE_PASS)):
        b.set_kernel_variant(variant)
        for name, mode in (("per_domain", rblbm.BATCH_PER_DOMAIN),
                           ("one_launch", rblbm.BATCH_ONE_LAUNCH),
                           ("graph", rblbm.BATCH_GRAPH)):
            b.set_mode(mode)
            b.step(a.warmup)
            ms = b.step_timed(a.steps)
            key = f"{form}/{name}"
            out["modes"][key] = {"ms_per_step": ms / a.steps,
                                 "glups": nodes * a.steps / (ms / 1e3) / 1e9}
            print(f"{key}: {out['modes'][key]}", file=sys.stderr, flush=True)
    b.set_kernel_variant(rblbm.KERNEL_AUTO)
    b.set_mode(rblbm.BATCH_GRAPH)
    b.step(a.warmup)
    nsamp = max(1, a.steps // a.every)
    b.domains[0].sync()
    t0 = time.perf_counter()
    for _ in range(nsamp):
        b.step(a.every)
        eager = [d.diagnostics() for d in b.domains]
    t_eager = time.perf_counter() - t0
    t0 = time.perf_counter()
    series = b.run_diagnostics(nsamp * a.every, a.every)
    t_graph = time.perf_counter() - t0
    for name, t in (("diag_eager", t_eager), ("diag_graph", t_graph)):
        out["modes"][name] = {"ms_per_step": t * 1e3 / (nsamp * a.every),
                              "glups": nodes * nsamp * a.every / t / 1e9, "every": a.every}
        print(f"{name}: {out['modes'][name]}", file=sys.stderr, flush=True)
    assert len(series) == nsamp and len(eager) == len(b.domains)
    d = b.domains[0].diagnostics()
    out["domain0_nu_vol"] = d["nu_vol"]
    out["graph_vs_per_domain"] = {
        f: out["modes"][f + "/graph"]["glups"] / out["modes"][f + "/per_domain"]["glups"]
        for f in ("auto", "two_pass", "one_pass")}
    b.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
