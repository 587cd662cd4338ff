#!/usr/bin/env python
"""Benchmark of the IB timestep hot path on one MI355X.

Workload (C4): 200 RBCs (8832 sample sites each) + 400 platelets (1552 sample
sites each) = 2,387,200 markers in a 32 x 64 x 32 um box, walls in y,
fd::mac{128} -> 128 x 256 x 128 cells; cosine delta; FP64.

A step is one spread of all marker forces onto the MAC grid plus one
interpolation of the grid velocity back to all markers, each a single
launch over every cell (uneven RBC/platelet marker counts in one pass).
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _load(name, path):
    import importlib.util
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def touched_points(ibm, grids, xs, delta_support=4):
    """Exact count of distinct grid points touched by any marker stencil, per
    component (T_c in the interpolation's algorithmic bytes), on the GPU."""
    import torch
    out = []
    for c in range(3):
        g = grids[c]
        mask = torch.zeros(g.size, dtype=torch.bool, device="cuda")
        pts = g.points
        for x in xs:
            n = x.numel() // 3
            idx = []
            for j in range(3):
                a = g.axis[j]
                xj = x[j * n:(j + 1) * n]
                r = a.resolution
                s = a.shift
                if a.solid:
                    u = xj * r
                else:
                    u = torch.remainder(xj - s / r, a.length) * r + s
                base = torch.floor(u - s).long()
                if not a.solid:
                    base = torch.remainder(base, a.cells)
                else:
                    base = base.clamp(-1, a.points - 1)
                offs = base[:, None] + torch.arange(-1, 3, device="cuda")[None, :]
                if a.solid:
                    offs = offs.clamp(0, a.points - 1)
                else:
                    offs = torch.remainder(offs, a.cells)
                idx.append(offs)
            lin = (idx[0][:, None, None, :] + pts[0] * (idx[1][:, None, :, None] + pts[1] * idx[2][:, :, None, None]))
            mask[lin.reshape(-1)] = True
        out.append(int(mask.sum()))
    return out


def cpu_baseline(orc, inputs, cfg, L, refinement):
    """The oracle (a scalar restatement of the reference algorithms) on a
    bounded sample: 20 RBCs + 40 platelets on the full C4 grid."""
    spec = [(L[0], 0, (0.0, 0.0), (0.0, 0.0)), (L[1], 1, (1.0, 0.0), (1.0, 0.0)), (L[2], 0, (0.0, 0.0), (0.0, 0.0))]
    axes = orc.mac_axes(spec, refinement)
    nr = cfg["n_rbc"] // 10
    npl = cfg["n_plt"] // 10
    x = np.concatenate([cfg["x_rbc"].reshape(3, -1)[:, :nr], cfg["x_plt"].reshape(3, -1)[:, :npl]], axis=1)
    f = np.concatenate([cfg["f_rbc"].reshape(3, -1)[:, :nr], cfg["f_plt"].reshape(3, -1)[:, :npl]], axis=1)
    x = np.ascontiguousarray(x.reshape(-1))
    f = np.ascontiguousarray(f.reshape(-1))
    n = x.size // 3
    t0 = time.perf_counter()
    fe = orc.spread(axes, "cosine", x, f)
    t1 = time.perf_counter()
    orc.interpolate(axes, "cosine", x, fe)
    t2 = time.perf_counter()
    pts = 2 * 3 * 64 * n
    return {"value": pts / (t2 - t0) / 1e9, "unit": "Gpts/s", "cores": 1, "kind": "port",
            "sample": f"20 RBCs + 40 platelets ({n} markers, 10% of C4) spread + interpolate on the full "
                      f"128x256x128 MAC grid; spread {t1 - t0:.2f} s, interpolate {t2 - t1:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU "
                                   "decomposition is a design only)"}))
        sys.exit(2)

    import torch

    ibm = _load("ibm", os.path.join(ROOT, "ib.cu_amd", "ibm.py"))
    inputs = _load("ibm_inputs", os.path.join(ROOT, "ib.cu_amd", "inputs.py"))

    cfg = inputs.config_c4()
    L = cfg["L"]
    refinement = cfg["refinement"]
    dom = [ibm.periodic(L[0]), ibm.dirichlet(L[1]), ibm.periodic(L[2])]
    grids = ibm.mac_grids(dom, refinement)
    dev = torch.device("cuda", 0)
    xs = [torch.from_numpy(cfg["x_rbc"]).to(dev), torch.from_numpy(cfg["x_plt"]).to(dev)]
    fs = [torch.from_numpy(cfg["f_rbc"]).to(dev), torch.from_numpy(cfg["f_plt"]).to(dev)]
    n = cfg["n_rbc"] + cfg["n_plt"]
    fe = [torch.empty(grids[c].size, dtype=torch.float64, device=dev) for c in range(3)]
    ul = [torch.empty_like(x) for x in xs]
    ctx = ibm.Context(0)

    def step():
        ibm.spread_sets(ctx, grids, "cosine", xs, fs, out=fe)
        ibm.interpolate_sets(ctx, grids, "cosine", xs, fe, outs=ul)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    K = args.steps
    ms_per_step = (t1 - t0) * 1e3 / K
    kernels = {}
    for name in ("bin", "spread_brick", "interp"):
        tot, cnt = ctx.profile_read(name)
        kernels[name] = {"ms": tot / max(cnt, 1), "launches": cnt}
    ctx.profile(False)

    Nc = [grids[c].size for c in range(3)]
    T = touched_points(ibm, grids, xs)
    bytes_spread = 48 * n + 8 * sum(Nc)
    bytes_interp = 48 * n + 8 * sum(T)
    dom_name = max(("spread_brick", "interp"), key=lambda k: kernels[k]["ms"])
    alg = bytes_spread if dom_name == "spread_brick" else bytes_interp
    achieved = alg / (kernels[dom_name]["ms"] * 1e-3) / 1e9
    gpts = 2 * 3 * 64 * n / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        orc = _load("orc", os.path.join(ROOT, "oracle", "orc.py"))
        cpu = cpu_baseline(orc, inputs, cfg, L, refinement)

    line = {
        "metric": "IB timesteps/s at 128³ + 600 cells; spread/interp Gpts/s; % HBM peak",
        "value": gpts,
        "unit": "Gpts/s (spread+interp marker-stencil interactions)",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded RBC/platelet surfaces, N(0,1) forces)",
        "config": {"workload": "C4 spread+interpolate: 200 RBCs x 8832 + 400 platelets x 1552 markers, "
                               "128x256x128 MAC grid, y walls, cosine delta",
                   "markers": n, "grid": [128, 256, 128]},
        "kernels_ms": {k: round(v["ms"], 4) for k, v in kernels.items()},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
