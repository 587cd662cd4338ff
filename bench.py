#!/usr/bin/env python
"""Benchmark of the immersed-boundary timestep on one MI355X.

Workload (C4): 200 RBCs (2210 data sites, 8832 sample sites each) and 400
platelets (605 / 1552) in a 32 x 64 x 32 um box, periodic in x and z, no-slip
walls in y with the top wall moving at 1e-3 m/s, fd::mac{128} ->
128 x 256 x 128 cells; cosine delta; water (nu = 1e-6 m^2/s, rho = 1 g/mL);
dt = 0.1 us.  Cells are tethered to a 2% larger copy of their shape, so
every step spreads nonzero forces.

A step is the library's whole IB timestep (ibm_solver_step_fixed): RBF
geometry of all cells (one matrix-core launch), tether forces, spread of
2,387,200 sample-site forces, advection, three Helmholtz solves and the
pressure projection (multigrid), the FP64 divergence check, interpolation at
684,000 data sites and their update.  The headline form ("mixed") runs the
multigrid V-cycles in FP32 inside FP64 defect correction; the whole step is
recorded once as a HIP graph and replayed.  The plain FP64 form (FP64
V-cycles) is timed the same way for comparison.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6

UM = 0.1                 # library length unit = 10 um
NU = 10.0                # 1e-6 m^2/s
RHO = 1e-3               # 1 g/mL
DT = 1e-4                # 0.1 us
WALL = 0.1               # 1e-3 m/s
STIFFNESS = 0.02
TETHER_SCALE = 1.02
TOL = 1e-10
FRAC = 0.3               # share of the cells the CPU baseline times
ADAPTIVE_STEPS = 8       # adaptive steps before the capture


def _load(name, path):
    import importlib.util
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def c4_cells(inputs, surfaces):
    L = (32 * UM, 64 * UM, 32 * UM)
    rp, pp = inputs.cell_poses(*L, 200, 400, 4)
    types = []
    for shape, nd, ns, deg, poses in ((inputs.rbc_shape, 2210, 8832, 5, rp),
                                      (inputs.platelet_shape, 605, 1552, 0, pp)):
        ct = surfaces.CellType(shape, nd, ns, deg, inputs.bauer_params)
        dref, sref = ct.data_reference(), ct.sample_reference()
        X = np.stack([inputs.posed(dref, q) for q in poses], axis=1)                      # (3, nc, nd)
        tether = np.stack([inputs.posed(sref, q, TETHER_SCALE) for q in poses], axis=1)   # (3, nc, ns)
        types.append({"ct": ct, "X": X, "tether": tether, "nc": len(poses)})
    return L, types


def make_solver(ibm, ctx, L, types, precision):
    import torch
    dev = torch.device("cuda", 0)
    cell_types = []
    for T in types:
        ct = T["ct"]
        mats = [torch.from_numpy(m).to(dev) for m in (ct.E, ct.Et, ct.Ep)]
        op = ibm.RbfOperator(ctx, "f64", mats, ct.nd)
        cell_types.append({"ops": op, "cells": T["nc"], "weights": torch.from_numpy(ct.w).to(dev),
                           "tether": torch.from_numpy(np.ascontiguousarray(T["tether"].reshape(-1))).to(dev),
                           "stiffness": STIFFNESS})
    dom = [ibm.periodic(L[0]), ibm.dirichlet(L[1]), ibm.periodic(L[2])]
    s = ibm.Solver(ctx, dom, 128, DT, NU, density=RHO, tolerance=TOL, precision=precision, max_iterations=60,
                   cell_types=cell_types)
    g, _ = s.grids()
    ub = [torch.zeros(g[i].size, dtype=torch.float64, device=dev) for i in range(3)]
    ibm.boundary_apply(ctx, g[0], 1, 1, ub[0], value=WALL, order=2)
    st = s.zero_state([torch.from_numpy(np.ascontiguousarray(T["X"].reshape(-1))).to(dev) for T in types])
    return s, st, ub, cell_types


KERNELS = ("spread_brick", "bin", "interp", "rbf_gemm", "geometry", "stencil_f64", "stencil_f32", "restrict",
           "prolong", "twopoint", "coarse_vcycle", "advection", "chebyshev_f32", "chebyshev_f64", "residual_restrict",
           "prolong_residual")


def run_form(ibm, ctx, stream, L, types, precision, steps, warmup):
    """Calibrates, captures and replays one form; then profiles eager
    fixed-count steps.  Returns timings and the per-kernel breakdown."""
    import torch
    s, st, ub, keep = make_solver(ibm, ctx, L, types, precision)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        # Adaptive steps set the V-cycle counts of the captured step (the
        # largest of the last four, + 1): the first two warm-start from the
        # previous solution, the later ones from the extrapolation of the
        # last two (the form every replay uses).
        for _ in range(ADAPTIVE_STEPS):
            s.step(st, boundary=ub)
        s.capture(st, boundary=ub)
        for _ in range(warmup):
            s.replay(st)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            s.replay(st)
        e1.record(stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    info = s.info()
    ms_graph = e0.elapsed_time(e1) / steps
    wall_ms = (t1 - t0) * 1e3 / steps
    # Per-kernel breakdown from eager fixed-count steps (same kernels).
    prof_steps = 3
    with torch.cuda.stream(stream):
        ctx.profile(True)
        t2 = time.perf_counter()
        for _ in range(prof_steps):
            s.step_fixed(st, boundary=ub)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        kern = {}
        for name in KERNELS:
            ms, n = ctx.profile_read(name)
            if n:
                kern[name] = {"ms_per_step": ms / prof_steps, "launches_per_step": n / prof_steps,
                              "work_per_step": ctx.profile_work(name) / prof_steps}
        ctx.profile(False)
    res = {"ms_graph": ms_graph, "wall_ms": wall_ms, "eager_ms": (t3 - t2) * 1e3 / prof_steps,
           "info": info, "kernels": kern, "state": st}
    s.close()
    return res


def cpu_baseline(inputs, orc, fd, rbf, L, types, gpu_iters):
    """The oracle (single-threaded restatement of the reference algorithms) on
    bounded samples of one C4 step, scaled to a whole step."""
    from threadpoolctl import threadpool_limits
    spec = [(L[0], 0, (0.0, 0.0), (0.0, 0.0)), (L[1], 1, (1.0, 0.0), (1.0, 0.0)), (L[2], 0, (0.0, 0.0), (0.0, 0.0))]
    axes = orc.mac_axes(spec, 128)
    parts = {}
    with threadpool_limits(limits=1):
        # (a) spread of 30% of the sample sites, interpolation at 30% of the data sites
        xs, fs, xd = [], [], []
        for T in types:
            ct, nc = T["ct"], T["nc"]
            k = max(int(round(nc * FRAC)), 1)
            y = rbf.evaluate(ct.E, ct.nd, T["X"][:, :k].reshape(3 * k, ct.nd), dtype=np.float64)
            xs.append(y.reshape(3, -1))
            fs.append((-STIFFNESS * (y.reshape(3, k, -1) - T["tether"][:, :k])).reshape(3, -1))
            xd.append(T["X"][:, :k].reshape(3, -1))
        x = np.ascontiguousarray(np.concatenate(xs, axis=1).reshape(-1))
        f = np.ascontiguousarray(np.concatenate(fs, axis=1).reshape(-1))
        t0 = time.perf_counter()
        fe = orc.spread(axes, "cosine", x, f)
        t1 = time.perf_counter()
        orc.interpolate(axes, "cosine", np.ascontiguousarray(np.concatenate(xd, axis=1).reshape(-1)), fe)
        t2 = time.perf_counter()
        parts["spread"] = (t1 - t0) / FRAC
        parts["interpolate"] = (t2 - t1) / FRAC
        # (b) RBF geometry (3 operators) for 30% of the cells, FP64 BLAS on one thread
        t3 = time.perf_counter()
        for T in types:
            ct, k = T["ct"], max(int(round(T["nc"] * FRAC)), 1)
            Xk = T["X"][:, :k].reshape(3 * k, ct.nd)
            for M in (ct.E, ct.Et, ct.Ep):
                rbf.evaluate(M, ct.nd, Xk, dtype=np.float64)
        parts["rbf"] = (time.perf_counter() - t3) / FRAC
        # (c) one V-cycle and one advection on a 1/8 sub-box (64 x 128 x 64), scaled x8
        sub = [(L[0] / 2, 0, (0.0, 0.0), (0.0, 0.0)), (L[1] / 2, 1, (1.0, 0.0), (1.0, 0.0)),
               (L[2] / 2, 0, (0.0, 0.0), (0.0, 0.0))]
        af = lambda r, al=(0.0, 0.5, 0.5): [orc.grid_axes(sub, r, list(al))[j] for j in range(3)]
        mg = fd.Multigrid(af, 64, 1.0, -DT * NU / 2, order="x-fastest")
        b = np.random.default_rng(0).standard_normal(mg.levels[0]["A"].shape[0])
        t4 = time.perf_counter()
        mg.vcycle(b)
        parts["v_cycle"] = (time.perf_counter() - t4) * 8
        MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
        axf = lambda al: [orc.grid_axes(sub, 64, list(al))[j] for j in range(3)]
        u = [np.random.default_rng(1).standard_normal(np.prod([a.points for a in axf(al)])) for al in MAC]
        t5 = time.perf_counter()
        fd.advection(axf, MAC, u)
        parts["advection"] = (time.perf_counter() - t5) * 8
    cycles = sum(gpu_iters)
    total = parts["spread"] + parts["interpolate"] + parts["rbf"] + parts["advection"] + cycles * parts["v_cycle"]
    return {"value": 1.0 / total, "unit": "timesteps/s", "cores": 1, "kind": "port",
            "sample": (f"oracle timed on samples of one C4 step and scaled to the whole step: spread of 30% of the "
                       f"sample sites ({parts['spread']:.1f} s scaled), interpolation at 30% of the data sites "
                       f"({parts['interpolate']:.1f} s), FP64 RBF geometry of 30% of the cells "
                       f"({parts['rbf']:.1f} s), advection and one CSR multigrid V-cycle on a 64x128x64 sub-box "
                       f"x8 ({parts['advection']:.1f} s, {parts['v_cycle']:.1f} s per cycle x {cycles} cycles, "
                       f"the GPU's counts); total {total:.0f} s per step")}


# Kernel-name prefixes of the counter files for each timed kernel group.
PMC_KERNELS = {"chebyshev_f32": "cheb_march_kernel<float", "chebyshev_f64": "cheb_march_kernel<double",
               "stencil_f32": "stencil_kernel<float", "stencil_f64": "stencil_kernel<double",
               "restrict": "::restrict_kernel<", "residual_restrict": "residual_restrict_kernel",
               "prolong": "::prolong_add_kernel", "prolong_residual": "prolong_residual_kernel",
               "rbf_gemm": "rbf_gemm_kernel",
               "spread_brick": "spread_brick_kernel", "interp": "interp_kernel", "advection": "advection_kernel"}
COPY_BYTES = 1 << 28     # the calibration copy: 256 MiB read, 256 MiB written


def pmc_probe(torch, ibm, L, types):
    """A known copy (the counters' calibration), then one adaptive and one
    fixed-count C4 step in the headline form."""
    x = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    print("copy done", file=sys.stderr, flush=True)
    stream = torch.cuda.Stream()
    ctx = ibm.Context(0, stream=stream)
    s, st, ub, _keep = make_solver(ibm, ctx, L, types, 32)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        s.step(st, boundary=ub)
        s.step_fixed(st, boundary=ub)
    torch.cuda.synchronize()
    print("steps done", file=sys.stderr, flush=True)


def pmc_summary(d):
    """Per-kernel HBM bytes per launch from two rocprofv3 --pmc runs (FETCH_SIZE
    and WRITE_SIZE, KiB per dispatch), each scaled so that the 256 MiB copy
    reads 256 MiB: bytes = 1024 * (fetch_KiB * f_fetch + write_KiB * f_write)."""
    import csv
    import glob

    def load(kind, counter):
        (path,) = glob.glob(os.path.join(d, f"{kind}_step_counter_collection.csv"))
        return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(open(path))
                if r["Counter_Name"] == counter]
    fetch, write = load("fetch", "FETCH_SIZE"), load("write", "WRITE_SIZE")
    copy_f = max(v for n, v in fetch if "copyBuffer" in n)
    copy_w = max(v for n, v in write if "copyBuffer" in n)
    f_fetch = COPY_BYTES / 1024 / copy_f     # one dispatch copies the whole buffer
    f_write = COPY_BYTES / 1024 / copy_w
    print(f"calibration: copy dispatch fetch {copy_f:.1f} KiB, write {copy_w:.1f} KiB of "
          f"{COPY_BYTES // 1024} KiB moved each way -> f_fetch {f_fetch:.4f}, f_write {f_write:.4f}")
    for name, prefix in PMC_KERNELS.items():
        fs = [v for n, v in fetch if prefix in n]
        ws = [v for n, v in write if prefix in n]
        if not fs or not ws:
            continue
        per = 1024 * (sum(fs) / len(fs) * f_fetch + sum(ws) / len(ws) * f_write)
        print(f"{name}: launches {len(fs)} fetch {sum(fs) / len(fs):.1f} KiB write {sum(ws) / len(ws):.1f} KiB "
              f"-> traffic_bytes_per_launch: {per:.0f}")


def recorded_traffic(kernel):
    """HBM bytes per launch of `kernel` as recorded from the counter runs
    (the --pmc-summary output kept in profiles/c4_step_kernels_hbm_traffic.md)."""
    path = os.path.join(ROOT, "profiles", "c4_step_kernels_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"{kernel}:") and "traffic_bytes_per_launch:" in line:
            return {"bytes_per_launch": float(line.rsplit(":", 1)[1]), "source": os.path.relpath(path, ROOT)}
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain FP64 form")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run a 256 MiB copy and two C4 steps only (the program rocprofv3 --pmc profiles)")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="print per-kernel HBM traffic from DIR/fetch_step_* and DIR/write_step_* counter files")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU "
                                   "decomposition is a design only)"}))
        sys.exit(2)

    import torch

    ibm = _load("ibm", os.path.join(ROOT, "ib.cu_amd", "ibm.py"))
    inputs = _load("ibm_inputs", os.path.join(ROOT, "ib.cu_amd", "inputs.py"))
    surfaces = _load("ibm_surfaces", os.path.join(ROOT, "ib.cu_amd", "surfaces.py"))
    t_setup = time.perf_counter()
    L, types = c4_cells(inputs, surfaces)
    print(f"setup: RBF operators built in {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    if args.pmc_probe:
        pmc_probe(torch, ibm, L, types)
        return

    stream = torch.cuda.Stream()
    ctx = ibm.Context(0, stream=stream)
    mixed = run_form(ibm, ctx, stream, L, types, 32, args.steps, args.warmup)
    print(f"mixed: {mixed['ms_graph']:.3f} ms/step (graph)", file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        plain = run_form(ibm, ctx, stream, L, types, 64, args.steps, args.warmup)
        print(f"plain: {plain['ms_graph']:.3f} ms/step (graph)", file=sys.stderr, flush=True)

    # Roofline of the kernel taking the most time per step.
    kern = mixed["kernels"]
    dom_name = max(kern, key=lambda k: kern[k]["ms_per_step"])
    d = kern[dom_name]
    if dom_name == "rbf_gemm":
        achieved = d["work_per_step"] / (d["ms_per_step"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s"}
    else:
        achieved = d["work_per_step"] / (d["ms_per_step"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    traffic = recorded_traffic(dom_name)
    roofline = {"bound": roof["bound"], "kernel": dom_name, "achieved": achieved, "peak": roof["peak"],
                "unit": roof["unit"], "frac": achieved / roof["peak"],
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "algorithmic_per_launch": d["work_per_step"] / d["launches_per_step"],
                "avg_launch_ms": d["ms_per_step"] / d["launches_per_step"]}
    per_kernel = {}
    for k, v in kern.items():
        if k == "rbf_gemm":
            rate = {"TFLOP/s": v["work_per_step"] / (v["ms_per_step"] * 1e-3) / 1e12}
        elif v["work_per_step"] > 0:
            rate = {"GB/s": v["work_per_step"] / (v["ms_per_step"] * 1e-3) / 1e9}
        else:
            rate = {}
        per_kernel[k] = {"ms_per_step": round(v["ms_per_step"], 4), "launches": v["launches_per_step"], **rate}

    n_sample = sum(T["ct"].ns * T["nc"] for T in types)
    n_data = sum(T["ct"].nd * T["nc"] for T in types)
    t_si = sum(kern[k]["ms_per_step"] for k in ("spread_brick", "interp") if k in kern)
    t_si += kern.get("bin", {"ms_per_step": 0})["ms_per_step"]
    gpts = 3 * 64 * (n_sample + n_data) / (t_si * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        orc = _load("orc", os.path.join(ROOT, "oracle", "orc.py"))
        fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
        rbf = _load("orc_rbf", os.path.join(ROOT, "oracle", "rbf.py"))
        cpu = cpu_baseline(inputs, orc, fd, rbf, L, types, mixed["info"]["iterations"])

    line = {
        "metric": "IB timesteps/s at 128³ + 600 cells; spread/interp Gpts/s; % HBM peak",
        "value": 1e3 / mixed["ms_graph"],
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_graph"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 multigrid cycles in f64 defect correction)",
        "data": "synthetic (seeded RBC/platelet poses, Bauer-spiral sites, tether forces, sheared channel)",
        "config": {"workload": "C4 IB timestep: 200 RBCs (2210/8832 sites) + 400 platelets (605/1552), "
                               "128x256x128 MAC grid, y walls with moving top wall, cosine delta, hipGraph replay",
                   "sample_sites": n_sample, "data_sites": n_data, "grid": [128, 256, 128], "dt_us": 0.1,
                   "tolerance": TOL},
        "plain_fp64_ms_per_step": plain["ms_graph"] if plain else None,
        "speedup_vs_plain_fp64": plain["ms_graph"] / mixed["ms_graph"] if plain else None,
        "eager_ms_per_step": mixed["eager_ms"],
        "solves": mixed["info"],
        "spread_interp_gpts": gpts,
        "kernels": per_kernel,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
