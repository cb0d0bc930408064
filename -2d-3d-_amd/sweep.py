#!/usr/bin/env python3
"""Time every kernel form of the 3D step on one cavity and print one JSON line per form.

Forms: the two-pass step in both population layouts (two arrays, in place), and the one-pass
step for each thread tile, prefetch setting and z-chunk.  Each form runs `--warmup` untimed
steps, then `--steps` steps timed with HIP events on the solver's stream; the results do not
depend on the form (the FP64 form is bitwise identical in all of them).  lbm_autotune runs the
same comparison on a scratch domain; this script lists every candidate.

Example:
    python ./-2d-3d-_amd/sweep.py --n 512 --forms fp64,fp32,fp16 --kchunks 8,16 --out sweep.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import rblbm  # noqa: E402

FORMS = {"fp64": rblbm.FP64, "fp32": rblbm.FP32_SHIFT, "fp16": rblbm.FP16_SHIFT}
TILES = {rblbm.TILE_32X16: "32x16", rblbm.TILE_16X16: "16x16", rblbm.TILE_32X8: "32x8",
         rblbm.TILE_32X32: "32x32"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ra", type=float, default=1e8)
    ap.add_argument("--forms", default="fp64,fp32,fp16")
    ap.add_argument("--kchunks", default="8,16")
    ap.add_argument("--tiles", default="0,1,2,3", help="one-pass tiles (3 = 32x32: FP32/FP16 only)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    kcs = [int(x) for x in a.kchunks.split(",") if x]
    tiles = [int(x) for x in a.tiles.split(",") if x]
    n = a.n
    res = []

    def timed(sol, **tag):
        sol.step(a.warmup)
        ms = sol.step_timed(a.steps) / a.steps
        r = dict(tag, ms=round(ms, 3), glups=round(n ** 3 / (ms / 1e3) / 1e9, 3))
        res.append(r)
        print(json.dumps(r), flush=True)

    for f in a.forms.split(","):
        for layout, name in ((rblbm.LAYOUT_AA, "aa"), (rblbm.LAYOUT_TWO_ARRAY, "two-array")):
            sol = rblbm.Solver(rblbm.Config(dims=3, nx=n, ny=n, nz=n, rayleigh=a.ra,
                                            precision=FORMS[f], layout=layout))
            sol.initial_seeded(1, 1e-3)
            sol.step(1)
            sol.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
            timed(sol, form=f, variant="two-pass", layout=name)
            if layout == rblbm.LAYOUT_TWO_ARRAY:
                for tile in tiles:
                    if tile == rblbm.TILE_32X32 and f == "fp64":
                        continue   # 1024 threads leave 128 registers a lane: too few for FP64
                    for pf in (0, 1):
                        for kc in kcs:
                            sol.set_kernel_variant(rblbm.KERNEL_ONE_PASS, kc)
                            sol.set_fused_tile(tile, pf)
                            timed(sol, form=f, variant="one-pass", layout=name,
                                  tile=TILES[tile], prefetch=pf, kchunk=kc)
            sol.close()
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1)


if __name__ == "__main__":
    main()
