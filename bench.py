"""Benchmark: FLIP 128^3 dam break on one MI355X.

A step is one substep of the hybrid particle-grid path (PhysicsAnimation::
AdvanceSingleFrame with frame dt = 1/480, which takes one substep at this CFL):
P2G, fluid SDF + markers, extrapolation, walls, gravity, pressure projection
(multigrid-preconditioned CG to ||b-Ax||/||b|| <= 1e-6, checked in FP64,
warm-started from the previous pressure), G2P and midpoint advection, on 4,021,248 particles seeded
2x2x2 per fluid cell in x < 0.4, y < 0.6.

Prints one JSON line: Mparticle-substeps/s (whole job), the median pressure-solve
ms, the roofline of the dominant kernel (average launch duration measured with
HIP events on the scene's stream over the timed region) and the CPU baseline
(the FP64 oracle on this host's cores, on a bounded sample of the same state).
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mparticle-substeps/s + pressure-solve ms @1e-6 residual, FLIP 128^3"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per cell for one launch of each timed kernel (FP32
# vectors, FP64 solution, 1-byte markers whose 6 neighbours come from cache).
KERNEL_BYTES_PER_CELL = {
    # k_cg_update: read x(8) d(4) q(4) r(4), write x'(8) r(4); the diagonal
    # preconditioner adds the row code(2) and z(4) (see update_bytes)
    "cg_update": 32,
    # k_cg_spmv: read z(4) d(4) row code(2), write d'(4) q(4)
    "cg_spmv": 18,
    # k_jacobi: read x(4) b(4) row code(2), write x'(4)
    "jacobi": 14,
    # k_mg_smooth on the finest level: read x(4) b(4) row code(2), write x'(4)
    "mg_smooth": 14,
    # k_mg_down on the finest level: read b(4) row code(2), write x2(4) and the coarse b (4/8)
    "mg_down": 10.5,
    # k_mg_up on the finest level: read x2(4) b(4) row code(2) and the coarse correction (4/8), write z(4)
    "mg_up": 14.5,
}
KERNEL_IDS = {"none": 0, "cg_update": 2, "cg_spmv": 1, "jacobi": 6, "p2g": 3, "g2p_move": 4, "sdf": 5,
              "mg_smooth": 7, "mg_down": 8, "mg_up": 9}
SOLVERS = {"jacobi": 0, "cg": 1, "pcg": 2, "mgpcg": 3}
SOLVER_TEXT = {"jacobi": "Jacobi sweeps", "cg": "CG", "pcg": "Jacobi-PCG", "mgpcg": "multigrid-PCG"}


def kernel_bytes_per_cell(kernel, solver, mg_split=False):
    b = KERNEL_BYTES_PER_CELL[kernel]
    if kernel == "cg_update" and solver == "pcg":
        b += 6
    if kernel == "cg_update" and solver == "mgpcg" and mg_split:  # row code(2) + the V-cycle's first sweep(4)
        b += 6
    return b


def kernel_bytes_per_launch(kernel, solver, n, npart, mg_split=False):
    """Algorithmic HBM bytes of one launch of a timed kernel on the n^3 FLIP scene."""
    cells = n ** 3
    if kernel in KERNEL_BYTES_PER_CELL:
        return int(kernel_bytes_per_cell(kernel, solver, mg_split) * cells)
    faces = (n + 1) * n * n  # per component
    if kernel == "p2g":  # one of three component launches: a third of the x, v read (8 of 24 per particle),
        return 8 * npart + 8 * faces  # the face velocity and the FLIP snapshot written (8 per face)
    if kernel == "g2p_move":  # read x, v (24) and write x, v (24) per particle; read u, v, w + snapshot (24 per face)
        return 48 * npart + 8 * 3 * faces
    if kernel == "sdf":  # read every particle position (12), cell starts (4) and write the SDF (8) per cell
        return 12 * npart + 12 * cells
    return None


def kernel_impl_bytes_per_launch(kernel, n, npart):
    """What this implementation's kernel reads and writes by design, where that
    differs from the algorithmic bytes (P2G: each particle read by about 1.76 face blocks)."""
    if kernel == "p2g":  # the scatter reads a float4 position (16) + v_c (4) per candidate: a particle is a
        # candidate of about 900 / 512 face blocks (10 x 10 x 9 candidate cells per 8 x 8 x 8 faces); two bin
        # starts (8) per row of candidate cells (100 rows per 512 faces); u + marker + FLIP snapshot (9) per face
        faces = (n + 1) * n * n
        return int(20 * npart * 900 / 512 + 8 * 100 * faces / 512 + 9 * faces)
    return None


def roofline(kernel, solver, n, npart, kms, launches, traffic, mg_split=False):
    bpl = kernel_bytes_per_launch(kernel, solver, n, npart, mg_split)
    if bpl is None or launches <= 0:
        return None
    avg_ms = kms / launches
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    out = {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "avg_launch_us": round(avg_ms * 1e3, 3), "launches": launches, "bytes_per_launch": bpl}
    impl = kernel_impl_bytes_per_launch(kernel, n, npart)
    if impl is not None:
        out["impl_bytes_per_launch"] = impl
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--kernel", default="p2g", choices=sorted(KERNEL_IDS),
                    help="kernel of the roofline object (default: P2G, the largest share of the substep)")
    ap.add_argument("--solver", default="mgpcg", choices=sorted(SOLVERS))
    ap.add_argument("--mg", default="fused", choices=["fused", "split"],
                    help="multigrid V-cycle kernels: one descent and one ascent per level, or one per step")
    # one substep (about 20 s on 16 host threads) keeps the CPU sample within
    # half a minute; the rate is per substep, so more substeps only average it
    ap.add_argument("--cpu-substeps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solver-comparison", action="store_true",
                    help="skip the plain-CG / multigrid-PCG comparison after the timed run")
    ap.add_argument("--batch-frames", type=int, default=10,
                    help="frames of the 96-scene batch workload after the headline run (0 skips it)")
    ap.add_argument("--no-graph-comparison", action="store_true",
                    help="skip the comparison with substeps run as captured HIP graphs")
    ap.add_argument("--batch-workers", type=int, nargs="+", default=[4],
                    help="host worker threads of the batch (several values: one measurement each)")
    ap.add_argument("--learned-steps", type=int, default=1500,
                    help="training steps of the learned warm start (0 skips that workload)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--summarize-trace", metavar="STATS_CSV",
                    help="write profiles/flip128_kernel_trace.md from a rocprofv3 --stats kernel summary and exit")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/flip128_hbm_traffic.{json,md} from FETCH_SIZE/WRITE_SIZE passes and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (k_copy_probe), the reference kernel "
                         "that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    return ap.parse_args()


KERNEL_SYMBOLS = {"cg_update": "k_cg_update", "cg_spmv": "k_cg_spmv", "jacobi": "k_jacobi", "p2g": "k_p2g_scatter",
                  "g2p_move": "k_g2p_move", "sdf": "k_sdf", "mg_smooth": "k_mg_smooth", "mg_down": "k_mg_down<finest>",
                  "mg_up": "k_mg_up<finest>"}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "flip128_hbm_traffic.json")
TRACE_CMD = ("rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/kt -o kt -- "
             "python3 bench.py --steps 20 --warmup 5 --kernel none --no-cpu-baseline --no-solver-comparison "
             "--no-graph-comparison --batch-frames 0 --learned-steps 0")
PMC_CMD = ("rocprofv3 --pmc {counter} --output-format csv -d gpurun_out/pmc_{counter} -o pmc -- "
           "python3 bench.py --steps 3 --warmup 2 --kernel none --no-cpu-baseline --no-solver-comparison "
           "--no-graph-comparison --batch-frames 0 --learned-steps 0 --copy-calibration 256")


def short_name(kernel):
    """'fe::k_cg_update(fe::Geom, ...)' -> 'k_cg_update'; template arguments
    dropped, except that the fused V-cycle kernels keep their level class
    ('k_mg_down<finest>' for the finest level, 'k_mg_down<coarse>' below it)."""
    name = kernel.split("(")[0].replace("void ", "").strip()
    base, _, targs = name.partition("<")
    base = base.split("::")[-1]
    if base in ("k_mg_down", "k_mg_up") and targs:
        base += "<finest>" if targs.startswith("true") else "<coarse>"
    return base


def summarize_trace(stats_csv):
    """Markdown table of a rocprofv3 --kernel-trace --stats summary."""
    import csv

    rows = list(csv.DictReader(open(stats_csv)))
    lines = ["# FLIP 128^3 dam break: kernel trace (rocprofv3)", "",
             "Command (repo root on the MI355X box; 5 warm-up + 20 timed substeps, 4,021,248 particles, "
             "multigrid-PCG to 1e-6 relative residual, warm start):", "", "```", TRACE_CMD, "```", "",
             f"Source: `{os.path.basename(stats_csv)}`, summarised by `python3 bench.py --summarize-trace <csv>`. "
             "25 substeps in total (warm-up included).", "",
             "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| `{r['Name'].split('(')[0]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")
    return "\n".join(lines) + "\n"


def summarize_pmc(fetch_csv, write_csv):
    """HBM bytes per launch per kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE passes,
    each scaled by what the counter reads for k_copy_probe (a plain copy of known size)."""
    import csv

    def load(path, counter):
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            acc.setdefault(short_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]) * 1024.0)
        return acc

    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_bytes = 256 * 1024 * 1024
    f_scale = copy_bytes / fetch["k_copy_probe"][0]
    w_scale = copy_bytes / write["k_copy_probe"][0]
    out = {"fetch_scale": f_scale, "write_scale": w_scale, "calibration": "k_copy_probe, 256 MiB read + 256 MiB "
           "written", "commands": [PMC_CMD.format(counter=c) for c in ("FETCH_SIZE", "WRITE_SIZE")], "kernels": {}}
    for k in sorted(fetch):
        if k == "k_copy_probe" or k not in write:
            continue
        fb = sum(fetch[k]) / len(fetch[k]) * f_scale
        wb = sum(write[k]) / len(write[k]) * w_scale
        out["kernels"][k] = {"launches": len(fetch[k]), "read_bytes": round(fb), "write_bytes": round(wb),
                             "bytes_per_launch": round(fb + wb)}
    return out


def traffic_for(kernel):
    try:
        with open(TRAFFIC_JSON) as f:
            return json.load(f)["kernels"][KERNEL_SYMBOLS[kernel]]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(ctx, mib):
    ms = ctx.copy_probe(mib * 1024 * 1024)
    print(f"bench: copy calibration k_copy_probe {mib} MiB read + {mib} MiB written, {ms:.3f} ms",
          file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.summarize_trace:
        md = summarize_trace(args.summarize_trace)
        with open(os.path.join(ROOT, "profiles", "flip128_kernel_trace.md"), "w") as f:
            f.write(md)
        print(md)
        return
    if args.summarize_pmc:
        t = summarize_pmc(*args.summarize_pmc)
        with open(TRAFFIC_JSON, "w") as f:
            json.dump(t, f, indent=1)
        lines = ["# FLIP 128^3 dam break: HBM traffic per kernel launch (rocprofv3 PMC)", "",
                 "Commands (repo root on the MI355X box), one counter per run:", "", "```"] + t["commands"] + [
                 "```", "", f"Scale factors from `k_copy_probe` (256 MiB read + 256 MiB written): FETCH_SIZE x "
                 f"{t['fetch_scale']:.4f}, WRITE_SIZE x {t['write_scale']:.4f} (KiB counters). Summarised by "
                 "`python3 bench.py --summarize-pmc <fetch_csv> <write_csv>`; machine-readable copy in "
                 "`flip128_hbm_traffic.json` (read by bench.py for `roofline.traffic`).", "",
                 "| kernel | launches | read MB/launch | write MB/launch | total MB/launch |",
                 "|---|---:|---:|---:|---:|"]
        for k, v in sorted(t["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
            lines.append(f"| `{k}` | {v['launches']} | {v['read_bytes'] / 1e6:.2f} | {v['write_bytes'] / 1e6:.2f} | "
                         f"{v['bytes_per_launch'] / 1e6:.2f} |")
        md = "\n".join(lines) + "\n"
        with open(os.path.join(ROOT, "profiles", "flip128_hbm_traffic.md"), "w") as f:
            f.write(md)
        print(md)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding is a design only "
              "(z-slabs with face/pressure halos) -- run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from fluide_engine_amd import capi

    n = args.n
    dt = 1.0 / 480.0
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_cfl=3.0, gravity=(0.0, -9.81, 0.0),
                           max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1.0, -1.0, -1.0), (0.4, 0.6, 2.0))], seed=0)
    v = np.zeros_like(x)
    npart = x.shape[0]
    if args.solver == "jacobi":  # the reference's CUDA default: 1000 sweeps, ||b-Ax|| < 1e-14 checked every 5
        solver = capi.solver(capi.SOLVER_JACOBI, 1000, 5, 1e-14, warm_start=True)
    else:
        solver = capi.solver(SOLVERS[args.solver], 2000, 0, 1e-6, relative=True, warm_start=True)

    ctx = capi.Context(0)
    scene = capi.Scene(ctx, desc)
    scene.set_mg_variant(capi.MG_FUSED if args.mg == "fused" else capi.MG_SPLIT)
    scene.upload_particles(x, v)
    print(f"bench: {ctx.device_name} FLIP {n}^3 {npart} particles, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    fluid_cells = []  # FLUID cells (the fluid volume in cells) after the first frame and around the timed region
    for i in range(args.warmup):
        scene.advance_frame(dt, solver)
        if i == 0:
            fluid_cells.append(int((scene.get(capi.GRID_FLAGS) == capi.FLUID).sum()))
    scene.sync()
    fluid_cells.append(int((scene.get(capi.GRID_FLAGS) == capi.FLUID).sum()))

    scene.kernel_timer(KERNEL_IDS[args.kernel])
    pressure_ms, iters, resid, sub_total = [], [], [], 0
    stage_acc = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        # one frame of frame dt = substep dt; the stats of the last substep are kept per frame
        nsub = scene.advance_frame(dt, solver)
        sub_total += nsub
        st = scene.last_stats()
        pressure_ms.append(st.ms)
        iters.append(st.iterations)
        resid.append(st.rel_residual)
        for k, val in scene.stage_times().items():
            stage_acc[k] = stage_acc.get(k, 0.0) + val
        if args.verbose and i % 20 == 0:
            print(f"bench: step {i} pressure {st.ms:.3f} ms it {st.iterations} res {st.rel_residual:.2e}",
                  file=sys.stderr, flush=True)
    scene.sync()
    t1 = time.perf_counter()
    kms, klaunch = scene.kernel_timer_read(reset=True)
    fluid_cells.append(int((scene.get(capi.GRID_FLAGS) == capi.FLUID).sum()))
    wall = t1 - t0
    value = npart * sub_total / wall / 1e6
    on_profile = args.n == 128 and args.solver == "mgpcg"
    roof = roofline(args.kernel, args.solver, n, npart, kms, klaunch,
                    traffic_for(args.kernel) if on_profile else None, scene.mg_variant() == capi.MG_SPLIT)
    # the other timed kernels' rooflines, each over a few more substeps after the timed region
    rooflines = {}
    if args.kernel != "none":
        others = ["g2p_move", "sdf"] + ([] if args.solver == "jacobi" else ["cg_update", "cg_spmv"])
        if args.solver == "mgpcg":
            others += ["mg_down", "mg_up"] if scene.mg_variant() == capi.MG_FUSED else ["mg_smooth"]
        for kern in others:
            if kern == args.kernel:
                continue
            scene.kernel_timer(KERNEL_IDS[kern])
            for i in range(min(20, max(args.steps, 1))):
                scene.advance_frame(dt, solver)
            scene.sync()
            cms, clau = scene.kernel_timer_read(reset=True)
            rooflines[kern] = roofline(kern, args.solver, n, npart, cms, clau,
                                       traffic_for(kern) if on_profile else None,
                                       scene.mg_variant() == capi.MG_SPLIT)
    scene.kernel_timer(KERNEL_IDS["none"])
    graphs = None
    if not args.no_graph_comparison:
        graphs = graph_substeps(capi, ctx, desc, x, scene.mg_variant(), dt, solver, min(args.steps, 50), npart)

    solve_cmp = None if (args.solver == "jacobi" or args.no_solver_comparison) else solver_comparison(scene, capi, dt)
    learned = None
    if args.learned_steps > 0 and args.solver != "jacobi":
        learned = learned_warm_start(capi, ctx, scene, n, dt, args.learned_steps)
    batch = None
    if args.batch_frames > 0:
        batch = [batch_workload(capi, ctx, args.batch_frames, w) for w in args.batch_workers]

    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, capi, desc, dt, args)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mparticle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(sub_total, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"FLIP (PIC weight 0.03) dam break {n}^3, {npart} particles, dt=1/480, "
                               f"{SOLVER_TEXT[args.solver]} to 1e-6 relative residual (FP64-checked), warm start",
                   "solver": args.solver, "mg_kernels": args.mg if args.solver == "mgpcg" else None,
                   "grid": [n, n, n], "particles": npart, "substeps": sub_total},
        "pressure_ms_median": round(statistics.median(pressure_ms), 4),
        "pressure_iterations_median": statistics.median(iters),
        "pressure_rel_residual_max": max(resid),
        "fluid_volume_cells": {"seeded": seeded_cells(n), "after_first_frame": fluid_cells[0],
                               "before_timed": fluid_cells[-2], "after_timed": fluid_cells[-1],
                               "drift_over_run": round((fluid_cells[-1] - fluid_cells[0]) / max(fluid_cells[0], 1), 5)},
        "stage_ms_per_step": {k: round(val / max(args.steps, 1), 4) for k, val in stage_acc.items()},
        "roofline": roof,
        "rooflines_other": rooflines,
        "pressure_solver_comparison": solve_cmp,
        "learned_warm_start": learned,
        "graph_substeps": graphs,
        "batch_96_scenes": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    scene.close()
    ctx.close()


def seeded_cells(n):
    """Cells of the n^3 dam-break box x < 0.4, y < 0.6 (cell centres inside)."""
    c = (np.arange(n) + 0.5) / n
    return int((c < 0.4).sum() * (c < 0.6).sum() * n)


def solver_comparison(scene, capi, dt):
    """One pressure projection (build + solve + apply, device ms) on the input a
    substep gives it -- the current particles through P2G, SDF, extrapolation,
    walls and gravity -- by plain CG (FP32 vectors, FP64 dots) from a zero
    guess, multigrid-PCG from a zero guess and multigrid-PCG warm-started from
    the previous pressure, all to 1e-6 relative residual.  The grid and the
    pressure are restored before each run."""
    for stage in (capi.STAGE_P2G, capi.STAGE_SDF, capi.STAGE_EXTRAPOLATE, capi.STAGE_BC, capi.STAGE_GRAVITY,
                  capi.STAGE_BC):
        scene.run_stage(stage, dt)
    fields = (capi.GRID_U, capi.GRID_V, capi.GRID_W, capi.GRID_PRESSURE)
    saved = {f: scene.get(f) for f in fields}
    out = {}
    for name, kind, warm in (("cg_zero_guess", capi.SOLVER_CG, False),
                             ("mgpcg_zero_guess", capi.SOLVER_MGPCG, False),
                             ("mgpcg_warm_start", capi.SOLVER_MGPCG, True)):
        best = None
        for _ in range(3):
            for f, a in saved.items():
                scene.set(f, a)
            st = scene.run_stage(capi.STAGE_PRESSURE, dt,
                                 capi.solver(kind, 5000, 0, 1e-6, relative=True, warm_start=warm))
            if best is None or st.ms < best.ms:
                best = st
        out[name] = {"ms": round(best.ms, 4), "iterations": best.iterations,
                     "rel_residual": float(f"{best.rel_residual:.3e}")}
    for f, a in saved.items():
        scene.set(f, a)
    out["mgpcg_warm_over_cg_time"] = round(out["mgpcg_warm_start"]["ms"] / out["cg_zero_guess"]["ms"], 4)
    return out


def learned_warm_start(capi, ctx, scene, n, dt, steps):
    """The learned warm start at n^3: samples (pressure systems and solutions)
    from 20 randomised scenes (seeds 100-119; even: a dam break x < U[0.25,
    0.6], y < U[0.4, 0.8]; odd: a pool y < U[0.1, 0.3] and a sphere of radius
    U[0.1, 0.2] above it), one after every 30 of 120 substeps; `steps` Adam
    steps on the device; then
    plain CG and multigrid-PCG to 1e-6 from zero and from the network's guess
    on the dam-break input of the timed scene (build + guess + solve + apply,
    device ms, best of 3)."""
    net = capi.Net(ctx, seed=7)
    t0 = time.perf_counter()
    mg = capi.solver(capi.SOLVER_MGPCG, 2000, 0, 1e-6, relative=True, warm_start=True)
    for seed in range(100, 120):
        rng = np.random.default_rng(seed)
        if seed % 2 == 0:
            shapes = [capi.box((-1.0, -1.0, -1.0), (rng.uniform(0.25, 0.6), rng.uniform(0.4, 0.8), 2.0))]
        else:
            r = rng.uniform(0.1, 0.2)
            shapes = [capi.box((-1.0, -1.0, -1.0), (2.0, rng.uniform(0.1, 0.3), 2.0)),
                      capi.sphere((rng.uniform(0.3, 0.7), rng.uniform(0.55, 0.75), rng.uniform(0.3, 0.7)), r)]
        desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
        x = capi.seed_particles(desc, shapes, seed=seed)
        s = capi.Scene(ctx, desc)
        s.upload_particles(x, np.zeros_like(x))
        for k in range(1, 121):
            s.substep(dt, mg)
            if k % 30 == 0:
                net.add_sample(s)
        s.close()
    t_gen = time.perf_counter() - t0
    t1 = time.perf_counter()
    losses = net.train(steps, learning_rate=2e-3, seed=1)
    t_train = time.perf_counter() - t1
    print(f"bench: learned warm start: {net.sample_count()} samples in {t_gen:.1f} s, {steps} steps in "
          f"{t_train:.1f} s, loss {losses[:20].mean():.3f} -> {losses[-50:].mean():.4f}", file=sys.stderr, flush=True)

    for stage in (capi.STAGE_P2G, capi.STAGE_SDF, capi.STAGE_EXTRAPOLATE, capi.STAGE_BC, capi.STAGE_GRAVITY,
                  capi.STAGE_BC):
        scene.run_stage(stage, dt)
    fields = (capi.GRID_U, capi.GRID_V, capi.GRID_W, capi.GRID_PRESSURE)
    saved = {f: scene.get(f) for f in fields}
    scene.set_warm_start_net(net)
    out = {"samples": net.sample_count(), "train_steps": steps, "train_s": round(t_train, 2),
           "train_ms_per_step": round(t_train * 1e3 / steps, 3),
           "train_loss_first20": round(float(losses[:20].mean()), 5),
           "train_loss_last50": round(float(losses[-50:].mean()), 5),
           "loss": "relative A-norm error^2 after the best scaling"}
    for name, kind, warm in (("cg_zero_guess", capi.SOLVER_CG, capi.WARM_START_ZERO),
                             ("cg_learned_guess", capi.SOLVER_CG, capi.WARM_START_LEARNED),
                             ("mgpcg_zero_guess", capi.SOLVER_MGPCG, capi.WARM_START_ZERO),
                             ("mgpcg_learned_guess", capi.SOLVER_MGPCG, capi.WARM_START_LEARNED)):
        best = None
        for _ in range(3):
            for f, a in saved.items():
                scene.set(f, a)
            st = scene.run_stage(capi.STAGE_PRESSURE, dt,
                                 capi.solver(kind, 5000, 0, 1e-6, relative=True, warm_start=warm))
            if best is None or st.ms < best.ms:
                best = st
        out[name] = {"ms": round(best.ms, 4), "iterations": best.iterations,
                     "rel_residual": float(f"{best.rel_residual:.3e}")}
    for f, a in saved.items():
        scene.set(f, a)
    scene.set_warm_start_net(None)
    net.close()
    return out


def graph_substeps(capi, ctx, desc, x, mg, dt, solver, steps, npart):
    """Eager launches against captured HIP graphs (fe_scene_set_graph) on the
    same work: two fresh scenes from the seeded particles, one in graph mode,
    advance 3 untimed frames (the graph scene's first substep runs eagerly,
    the next two capture one graph pair per parity of the particle buffers),
    then `steps` timed frames each, one scene after the other.  The runs are
    bit-identical (particles and pressure compared at the end), so both solve
    the same systems with the same iteration counts."""
    pair = []
    for graph in (False, True):
        sc = capi.Scene(ctx, desc)
        sc.set_mg_variant(mg)
        sc.upload_particles(x, np.zeros_like(x))
        sc.set_graph(graph)
        pair.append(sc)

    def run(sc, k):
        its = []
        t0 = time.perf_counter()
        sub = 0
        for _ in range(k):
            sub += sc.advance_frame(dt, solver)
            its.append(sc.last_stats().iterations)
        sc.sync()
        return time.perf_counter() - t0, sub, its

    eager, graph = pair
    run(eager, 3)
    run(graph, 3)
    we, se, ie = run(eager, steps)
    wg, sg, ig = run(graph, steps)
    n_graph, captures, resumed = graph.graph_stats()
    xe, ve, _ = eager.download_particles()
    xg, vg, _ = graph.download_particles()
    same = bool(np.array_equal(xe, xg) and np.array_equal(ve, vg) and ie == ig
                and np.array_equal(eager.get(capi.GRID_PRESSURE), graph.get(capi.GRID_PRESSURE)))
    for sc in pair:
        sc.close()
    return {"frames": steps, "after_frames": 3, "eager_ms_per_step": round(we / se * 1e3, 4),
            "graph_ms_per_step": round(wg / sg * 1e3, 4), "iterations_median": float(np.median(ie)),
            "graph_mparticle_substeps_per_s": round(npart * sg / wg / 1e6, 2), "bit_identical": same,
            "graph_substeps": n_graph, "captures": captures, "resumed": resumed}


def batch_workload(capi, ctx, frames, workers, warmup=2):
    """96 independent FLIP dam breaks on cubic grids n ~ U{24,32,48,64,96}
    (seed 1), each one fluid box with extents ~ U[0.2, 0.7], 8 particles per
    fluid cell, multigrid-PCG to 1e-6, frame dt = 1/480, advanced together by
    fe_batch_advance_frame.  Reports scenes/s (scene-frames per second), the
    whole batch's particle-substeps/s and the per-scene substep wall time."""
    specs = capi.batch_scene_specs()
    scenes, npart = [], 0
    for n, lo, hi in specs:
        desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_cfl=3.0, gravity=(0.0, -9.81, 0.0),
                               max_particles=8 * n ** 3)
        x = capi.seed_particles(desc, [capi.box(lo, hi)], seed=n)
        desc.max_particles = max(1, x.shape[0])
        sc = capi.Scene(ctx, desc)
        sc.upload_particles(x, np.zeros_like(x))
        scenes.append((sc, x.shape[0]))
        npart += x.shape[0]
    solver = capi.solver(capi.SOLVER_MGPCG, 2000, 0, 1e-6, relative=True, warm_start=True)
    b = capi.Batch(ctx, [sc for sc, _ in scenes], workers)
    dt = 1.0 / 480.0
    for _ in range(warmup):
        b.advance_frame(dt, solver)

    def timed():
        per_substep, psub = [], 0
        t0 = time.perf_counter()
        for _ in range(frames):
            sub, ms = b.advance_frame(dt, solver)
            psub += sum(cnt * s for (_, cnt), s in zip(scenes, sub))
            per_substep.extend(m / max(s, 1) for m, s in zip(ms, sub))
        wall = time.perf_counter() - t0
        per_substep.sort()
        q = lambda f: per_substep[min(len(per_substep) - 1, int(f * len(per_substep)))]
        return {"scenes_per_s": round(len(specs) * frames / wall, 2),
                "mparticle_substeps_per_s": round(psub / wall / 1e6, 2),
                "scene_substep_ms_p50": round(q(0.5), 4), "scene_substep_ms_p99": round(q(0.99), 4),
                "wall_s": round(wall, 3)}

    out = {"scenes": len(specs), "particles": npart, "frames": frames, "workers": workers}
    out.update(timed())  # default kernel variants (P2G scatter, G2P quads)
    # the autotuner picks per scene between bit-identical kernel variants
    t0 = time.perf_counter()
    picks = {}
    for (sc, _), (n, _, _) in zip(scenes, specs):
        (p2g, g2p), _ = sc.autotune(reps=3)
        key = f"{n}^3"
        picks.setdefault(key, {"p2g_tiles": 0, "p2g_faces": 0, "g2p_pairs": 0, "scenes": 0})
        picks[key]["scenes"] += 1
        picks[key]["p2g_tiles"] += int(p2g == capi.P2G_TILES)
        picks[key]["p2g_faces"] += int(p2g == capi.P2G_FACES)
        picks[key]["g2p_pairs"] += int(g2p == capi.G2P_PAIRS)
    out["autotune_s"] = round(time.perf_counter() - t0, 3)
    out["autotune_picks"] = picks
    out["autotuned"] = timed()
    # the same batch with every scene's substeps as captured HIP graphs
    for sc, _ in scenes:
        sc.set_graph(True)
    for _ in range(2):  # captures (one per parity of each scene's particle buffers)
        b.advance_frame(dt, solver)
    out["graphs"] = timed()
    st = [sc.graph_stats() for sc, _ in scenes]
    out["graphs"].update({"graph_substeps": sum(a for a, _, _ in st), "captures": sum(c for _, c, _ in st),
                          "resumed": sum(r for _, _, r in st)})
    b.close()
    for sc, _ in scenes:
        sc.close()
    return out


def cpu_baseline(scene, capi, desc, dt, args):
    """The FP64 oracle (a port of the reference substep) on this host's cores,
    continuing from the GPU's current state for a few substeps."""
    from oracle import oracle as orc

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    n = tuple(desc.n)
    gx, gv, _ = scene.download_particles()
    o = orc.Scene(n, method=orc.FLIP, blend=desc.blend, max_cfl=desc.max_cfl, gravity=tuple(desc.gravity),
                  threads=threads)
    o.set_particles(gx.astype(np.float64), gv.astype(np.float64))
    o.set(orc.PRESSURE, scene.get(capi.GRID_PRESSURE))
    # the reference's CPU ConjugateGradientSolver (src/linear_system/conjugate_gradient_solver.cpp:14-55),
    # stopped at the metric's 1e-6 relative residual, warm-started from the previous pressure
    solv = orc.solver(orc.CG, 4000, 0, 1e-6, relative=True, warm_start=True)
    print(f"bench: cpu baseline, {threads} threads, {args.cpu_substeps} substeps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.cpu_substeps):
        o.substep(dt, solv)
    t = time.perf_counter() - t0
    return {"value": round(gx.shape[0] * args.cpu_substeps / t / 1e6, 4), "unit": "Mparticle-substeps/s",
            "cores": threads, "kind": "port",
            "sample": f"{args.cpu_substeps} substeps of the same FLIP {n[0]}^3 scene from the GPU state after "
                      f"warm-up ({gx.shape[0]} particles), FP64 oracle with CG to 1e-6, {t:.1f} s"}


if __name__ == "__main__":
    main()
