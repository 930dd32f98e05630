"""Throughput of the implicit residual-and-update iteration on one MI355X.

Workload: the curvilinear wall block of 256x128x64 cells (2,097,152 real
cells), laminar Navier-Stokes, Roe-MUSCL with minmod, viscous fluxes, wall and
far-field ghosts, local time step, matrix-free point-Jacobi implicit phase with
lmax=10 sub-iterations, all FP64 (codens_amd.configs.wall_ns).  One step is one
implicit iteration (one atsch_num stage); inputs are resident in HBM.

With --workload multiblock: 48 Cartesian blocks of seeded uneven sizes (each
axis 16..192 cells, 39,991,040 cells in all) joined 1-to-1 along a seeded
adjacency tree, Spalart-Allmaras RANS at Re=5e6, M=0.7, the same implicit
phase (codens_amd.configs.multiblock_sa); stages captured in a HIP graph
(--graph, on by default for this workload).

With --workload adjoint: the gradient of the wall force of one 4-stage RK
step on the 128x64x32 viscous block with respect to the state and 4,096
wall-node displacements (codens_amd.adjoint): `value` is real cells x RK
stages / the time of one whole gradient (forward step with its
checkpoints, reverse sweep, mesh adjoint), next to the forward step alone.

With --workload converge: iterations and wall time to a 1e-8 density-
residual drop on the inclined-free-stream wall block 32x16x8 (the case of
tests/golden/converged_wall_32x16x8.npz), FP64 implicit phase and FP32
defect correction; `value` is the FP32 run's cell-updates/s to the drop.

Prints one JSON line: Mcell-updates/s, the roofline of the dominant kernel
(implimf_sweep: 264 algorithmic bytes per real cell per launch) measured with
HIP events on the launch stream, and the reference Fortran build timed on the
host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SWEEP_BYTES_PER_CELL = 8 * (5 + 5 + 5 + 9 + 1 + 3 + 5)  # u, v, d_in, sn, diag, coefe, d_out
# Memory traffic of one sweep launch on the 256x128x64 workload: 2.0 x FETCH_SIZE
# + 1.0 x WRITE_SIZE (factors from a 512 MiB device copy in the same run),
# 658,503 KiB = 674.3 MB against 553.6 MB algorithmic (64x4 tiles)
# (profiles/traffic_wall_ns_256x128x64.md).
SWEEP_TRAFFIC_BYTES = 658503 * 1024
HBM_PEAK_GBS = 8000.0


def cpu_baseline(case, stages, sample=None):
    """The reference routines (flang build, 1 core) on `stages` implicit
    iterations of `case`; falls back to the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    import cpu_stage
    use_ref = oracle.ref_lib() is not None
    st = cpu_stage.CpuStage(case, use_ref=use_ref)
    t0 = time.perf_counter()
    for _ in range(stages):
        st.stage()
    dt = time.perf_counter() - t0
    return {"value": case.layout.ncells * stages / dt / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": "reference" if use_ref else "port",
            "sample": f"{stages} implicit iterations of {sample or 'the same 256x128x64 case'}, lmax=10, "
                      f"{'flang -O2 build of the reference routines' if use_ref else 'oracle C restatement'}"}


def cpu_allcores(case, stages, sample=None):
    """The oracle's C restatement of the same stage (OpenMP over the block's
    planes, bit for bit the serial result) on every core OpenMP is given:
    the all-cores CPU comparator beside the one-core reference build."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_stage
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    st = cpu_stage.CpuStage(case, use_ref=False)
    st.stage()
    t0 = time.perf_counter()
    for _ in range(stages):
        st.stage()
    dt = time.perf_counter() - t0
    return {"value": case.layout.ncells * stages / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "cpu": model, "nproc": os.cpu_count(),
            "sample": f"{stages} implicit iterations of {sample or 'the same 256x128x64 case'} (after one untimed), "
                      f"lmax=10, oracle C restatement with OpenMP"}


# kernels (or launch groups) timed with HIP events on the launch stream
TIMED_KERNELS = ("implimf_sweep", "implimf_sweep1", "implimf_coef", "implimf_update", "implimf32_sweep",
                 "implimf32_sweep1", "implimf32_coef", "implimf32_update", "roe_prim", "roe_limiter_chain",
                 "roe_face_flux", "roe_gather", "roe_fused", "roe_zero_ghost", "zvisqc", "chronos", "zpres",
                 "zvismo")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-stages", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dims", type=str, default="256,128,64")
    ap.add_argument("--workload", choices=("wall", "multiblock", "adjoint", "converge"), default="wall",
                    help="wall: the 256x128x64 laminar wall block; multiblock: 48 uneven SA blocks, ~40M cells")
    ap.add_argument("--streams", type=int, default=0,
                    help="side streams for the per-block work (0: 4 for multiblock, 1 otherwise)")
    ap.add_argument("--batched", default="auto",
                    help="phases of all blocks in one launch per pass: auto (the implicit phase on the multiblock "
                         "workload, the fastest combination there; the residual and turbulence phases stay per block "
                         "on the concurrent streams), on (residual,implicit,turbulence), off, or a comma-separated list "
                         "of those names")
    ap.add_argument("--graph", choices=("auto", "on", "off"), default="auto",
                    help="replay stages from a captured HIP graph (auto: on for multiblock)")
    ap.add_argument("--implicit", choices=("fp64", "fp32"), default="fp64",
                    help="implicit phase: implimf_3d in FP64 (the reference's), or its FP32 defect-correction form")
    ap.add_argument("--residual", choices=("fused", "split", "persistent"), default="fused",
                    help="sch_roe as one fused kernel (default), as face passes + gather, or fused with the "
                         "all-blocks launch persistent over a tile queue")
    ap.add_argument("--sweep-tile", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="tile of the implicit sub-iteration: -1 timed choice per block shape (default), "
                         "0 64x4, 1 32x8, 2 16x16")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="finish with 3 device-to-device copies of MIB MiB: the known-size kernel that "
                         "calibrates FETCH_SIZE/WRITE_SIZE counter passes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path (block sharding with face halos over RCCL) "
              "is a design, not a build, in this version", file=sys.stderr)
        sys.exit(2)
    import torch
    from codens_amd import abi, configs
    from codens_amd.solver import Solver

    abi.residual_path({"split": 0, "fused": 1, "persistent": 2}[args.residual])
    abi.sweep_tile(args.sweep_tile)
    if args.workload == "adjoint":
        return adjoint_bench(args)
    if args.workload == "converge":
        return converge_bench(args)
    multi = args.workload == "multiblock"
    ci, cj, ck = (int(x) for x in args.dims.split(","))
    case = configs.multiblock_sa() if multi else configs.wall_ns(ci, cj, ck)
    nstreams = args.streams or (4 if multi else 1)
    batched = {"auto": "implicit" if multi else "", "on": True, "off": ""}.get(args.batched, args.batched)
    s = Solver(case, implicit=args.implicit, streams=nstreams, batched=batched)
    ncell = case.layout.ncells
    nblk = len(case.layout.blocks)
    graph = args.graph == "on" or (args.graph == "auto" and multi)
    if graph:
        s.capture()  # one eager stage, then the captured one
    for _ in range(args.warmup):
        s.stage()
    torch.cuda.synchronize()
    abi.timing_enable(True)
    for name in TIMED_KERNELS:
        abi.timing_read(name, reset=True)
    abi.timing_enable(False)
    # timed region (no events inside)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.stage()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel timing pass with HIP events on the launch stream (eager
    # stages: the same launches as the graph's)
    g_saved, s.graph = s.graph, None
    side_saved, s.side = s.side, []  # one stream: events bracket one kernel at a time
    abi.timing_enable(True)
    for _ in range(max(3, args.steps // 4)):  # n_timing stages
        s.stage()
    torch.cuda.synchronize()
    kern = {}
    for name in TIMED_KERNELS:
        ms, cnt = abi.timing_read(name, reset=True)
        if cnt:
            kern[name] = {"avg_us": 1e3 * ms / cnt, "launches": cnt}
    abi.timing_enable(False)
    s.graph, s.side = g_saved, side_saved
    res = s.density_residual()
    ms_step = 1e3 * wall / args.steps
    value = ncell * args.steps / wall / 1e6
    fp32 = args.implicit == "fp32"
    sweep_name = "implimf32_sweep" if fp32 else "implimf_sweep"
    sweep_bytes = SWEEP_BYTES_PER_CELL // 2 if fp32 else SWEEP_BYTES_PER_CELL  # the same 33 values in FP32
    n_timing = max(3, args.steps // 4)
    sw = kern.get(sweep_name, {"avg_us": float("nan"), "launches": nblk * (case.lmax - 1) * n_timing})
    # bytes of all sweep launches over their summed time: the timing stages
    # sweep every cell lmax-1 times, in one launch per block and
    # sub-iteration, or one per sub-iteration when the blocks are batched
    cells_per_launch = ncell * (case.lmax - 1) * n_timing / sw["launches"]
    achieved = sweep_bytes * cells_per_launch / (sw["avg_us"] * 1e-6) / 1e9
    out = {
        "metric": "Mcell-updates/s per implicit iter to 1e-8 residual drop; % of HBM roofline",
        "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 residual/update, f32 implicit" if fp32 else "f64", "data": "synthetic",
        "config": {"workload": (f"multiblock_sa {nblk} blocks (axes 16..192, seed 20261203), coincident patches, "
                                "Spalart-Allmaras RANS Re=5e6 M=0.7, Roe-MUSCL minmod + viscous, implimf_3d lmax=10, "
                                if multi else
                                f"wall_ns {ci}x{cj}x{ck} laminar NS, Roe-MUSCL minmod + viscous, implimf_3d lmax=10, ")
                               + ("FP32 implicit phase (defect correction)" if fp32 else "FP64")
                               + (", stages replayed from a HIP graph" if graph else "")
                               + (f", blocks on {nstreams} concurrent streams" if len(s.side) > 1 else "")
                               + (", " + ", ".join(p for p in ("residual", "implicit", "turbulence")
                                                    if p in s.batch)
                                  + " of all blocks in one launch per pass" if s.batch else ""),
                   "cells": ncell, "blocks": nblk, "lmax": case.lmax,
                   "sweep_tile": ({0: "64x4", 1: "32x8", 2: "16x16"}.get(
                       abi.sweep_tile_of(abi.block(case.layout.blocks[0]), fp32=fp32), "none")
                       if nblk == 1 else "per block")},
        "roofline": {"bound": "hbm", "kernel": sweep_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": SWEEP_TRAFFIC_BYTES if (ci, cj, ck) == (256, 128, 64) and not fp32 and not multi
                     else None,
                     "bytes_per_cell": sweep_bytes,
                     "avg_us": round(sw["avg_us"], 2), "cells_per_launch": round(cells_per_launch)},
        "kernels": kern,
        "density_residual": res,
    }
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu:
        if multi:
            del s
            torch.cuda.empty_cache()
            small = configs.multiblock_sa(48, 16, 40, None)
            out["cpu_baseline"] = cpu_baseline(
                small, 2, f"a 48-block draw with axes 16..40 ({small.layout.ncells} cells, same seed, "
                          "same physics)")
        else:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_stages)
            out["cpu_baseline_all_cores"] = cpu_allcores(case, 3)
    print(json.dumps(out))


def converge_bench(args):
    import numpy as np
    import torch
    from codens_amd import configs
    from codens_amd.solver import Solver
    case = configs.adjoint_wall(32, 16, 8, eta=10.0, amp=0.0, bump=0.0)
    ncell = case.layout.ncells
    runs, fields = {}, {}
    inner = case.layout.interior()
    for mode in ("fp64", "fp32"):
        s = Solver(case, implicit=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, hist = s.run_to(drop=1e-8, max_iter=60000, check_every=250)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        runs[mode] = {"iterations": it, "seconds": round(dt, 3), "drop": hist[-1][1],
                      "Mcell_updates_per_s": round(ncell * it / dt / 1e6, 3)}
        w = s.host_state()[:5, inner]
        g1 = case.phys["gam"] - 1.0
        fields[mode] = {"rho": w[0], "u": w[1] / w[0],
                        "p": g1 * (w[4] - 0.5 * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2) / w[0])}
    # the converged FP32 defect-correction fields against the FP64 solve
    runs["fp32_vs_fp64"] = {
        q: {"L2": float(np.linalg.norm(fields["fp32"][q] - fields["fp64"][q]) / np.linalg.norm(fields["fp64"][q])),
            "Linf": float(np.abs(fields["fp32"][q] - fields["fp64"][q]).max() / np.abs(fields["fp64"][q]).max())}
        for q in ("rho", "u", "p")}
    out = {
        "metric": "Mcell-updates/s per implicit iter to 1e-8 residual drop; % of HBM roofline",
        "value": runs["fp32"]["Mcell_updates_per_s"], "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": runs["fp32"]["iterations"], "warmup": 0,
        "ms_per_step": round(1e3 * runs["fp32"]["seconds"] / runs["fp32"]["iterations"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 residual/update, f32 implicit", "data": "synthetic",
        "config": {"workload": "wall block 32x16x8, free stream inclined 10 and 5 degrees, laminar NS, unlimited MUSCL, "
                               "implimf_3d lmax=10, CFL 10, run to a 1e-8 density-residual drop (the residual check "
                               "every 250 iterations is inside the timing)", "cells": ncell},
        "to_1e-8": runs, "roofline": None,
    }
    print(json.dumps(out))


def adjoint_bench(args):
    import torch
    from codens_amd import abi, configs
    from codens_amd.adjoint import AdjointProblem, gradient, RK_ALPHA
    case = configs.adjoint_wall(128, 64, 32)
    P = AdjointProblem(case)
    v0 = torch.as_tensor(case.v0, device="cuda")
    delta = torch.zeros(len(case.wall_nodes), dtype=torch.float64, device="cuda")
    ncell = case.layout.ncells
    for _ in range(args.warmup):
        gradient(P, v0, delta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gradient(P, v0, delta)
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / args.steps
    b = case.layout.blocks[0]
    sn = torch.as_tensor(case.sn[9 * b.npc: 9 * (b.npc + b.nnn)], device="cuda")
    vol = torch.as_tensor(case.vol, device="cuda")
    t0 = time.perf_counter()
    with torch.no_grad():
        for _ in range(args.steps):
            P.objective(P.rk_step(v0, sn, vol), sn)
    torch.cuda.synchronize()
    tf = (time.perf_counter() - t0) / args.steps
    # per-entry-point times with HIP events around single calls
    ub = torch.ones((5, case.layout.ip11), dtype=torch.float64, device="cuda") * P.real
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    ev[0].record(); P.residual(v0, sn, vol); ev[1].record()
    P.jvp(v0, sn, vol, v0, sn, vol); ev[2].record()
    P.vjp(v0, sn, vol, ub); ev[3].record()
    torch.cuda.synchronize()
    nst = len(RK_ALPHA)
    out = {
        "metric": "Mcell-updates/s per implicit iter to 1e-8 residual drop; % of HBM roofline",
        "value": round(ncell * nst / tg / 1e6, 3), "unit": "Mcell-updates/s (RK stages of the gradient)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * tg, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint_wall 128x64x32 laminar NS, Roe + unlimited MUSCL + viscous, 4-stage RK step, "
                               "gradient of the wall x force w.r.t. the state and 4096 wall-node displacements",
                   "cells": ncell, "rk_stages": nst},
        "forward_step_ms": round(1e3 * tf, 3),
        "note": "the gradient's time includes the mesh metrics of the displaced nodes, their JVP and adjoint, all on the device",
        "entry_points_ms": {"residual (production)": round(ev[0].elapsed_time(ev[1]), 3),
                            "codens_adj_residual_jvp": round(ev[1].elapsed_time(ev[2]), 3),
                            "codens_adj_residual_vjp": round(ev[2].elapsed_time(ev[3]), 3)},
        "roofline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
