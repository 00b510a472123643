"""Benchmark: MPS rotating-drum timestep at 1M particles on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--l0 2.012e-3] [--solver learned|mixed|fp64]
                    [--precond two|jacobi] [--no-cpu-baseline] [--no-plain]
    python bench.py --workload batched [--buckets 64] [--steps K] [--warmup W] [--no-tune]
    python bench.py --workload precision --l0 1.6e-3 --steps 200
    python bench.py --workload radix [--steps K] [--warmup W]

Workload: the rotating drum (cylinder R = 0.15 m, L = 0.3 m, 9984-triangle STL, 30 rpm, 40 % fill)
at l0 = 2.012e-3 m: 1,008,472 fluid particles and 2,000 rigid spheres (d = 5 l0, 2500 kg/m3,
Schiller-Naumann drag), FP32 particle operators, pressure from the pressure
Poisson equation solved every step to an FP64 relative residual of 1e-8.  A step is one body of
the reference's CalculateNew loop: two binnings, wall binning, two STL distance passes,
predictor, PND, PPE assembly + PCG, corrector, two relocations, wall rotation.  Inputs are
resident in HBM when the timed region starts; inside it only the PCG's convergence flag (once per
32 iterations) and the per-step timings cross PCIe.

Production solver: mixed precision (FP32 matrix and SpMV, FP64 dots and residuals, FP64
refinement) with the cell V-cycle preconditioner: Jacobi plus one symmetric V-cycle on the
cell-level Galerkin operator whose coarse level is the dense inverse on 2^3-cell aggregates
(rebuilt every 50 solves; the JSON line gives the set-up time of one rebuild and its share per
step), warm-started by the previous pressure.  --solver learned adds the trained network's
warm start (simulation_project_amd/weights/mpnn_drum.npy).  The plain comparison solves the same
steps with FP64 Jacobi-PCG from the previous pressure.

One JSON line: value = particle-updates/s over the K timed steps; PPE ms/step of both solvers;
roofline of the dominant kernel (the PCG SpMV) from HIP events around every SpMV launch of the
timed region, on the solver's stream; a whole-step roofline from the algorithmic bytes of a step;
the CPU oracle timed on the host cores on one step of the same drum.
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
# HBM bytes per SpMV launch of the mixed-precision solver on the default drum (l0 = 2.012e-3 m),
# from two rocprofv3 counter passes (FETCH_SIZE scaled by the factor a plain copy kernel of known
# size reads, WRITE_SIZE unscaled): profiles/k_pcg_spmv_drum1M_traffic.md
SPMV_TRAFFIC_BYTES = {(2.012e-3, "mixed"): 958.6e6}
WEIGHTS = os.path.join(ROOT, "simulation_project_amd", "weights", "mpnn_drum.npy")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ppe_options(abi, kind, precond, warm):
    solver = abi.MPSR_PCG_FP64 if kind == "fp64" else abi.MPSR_PCG_MIXED
    pre = {"vcycle": abi.MPSR_PRECOND_CELL_VCYCLE, "two": abi.MPSR_PRECOND_TWO_LEVEL}.get(precond, abi.MPSR_PRECOND_JACOBI)
    return abi.PpeOptions(solver, 1e-8, 20000, warm, 0.0, 0.0, pre, 2, 50)


def run_steps(s, first, count):
    """Per-step timings of `count` steps; returns (elapsed s, list of timing dicts)."""
    out = []
    s.synchronize()
    t0 = time.perf_counter()
    for k in range(count):
        s.step(first + k)
        out.append(s.timings())
    s.synchronize()
    return time.perf_counter() - t0, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--l0", type=float, default=2.012e-3)
    ap.add_argument("--solver", choices=["learned", "mixed", "fp64"], default=None)
    ap.add_argument("--precond", choices=["vcycle", "two", "jacobi"], default="vcycle")
    ap.add_argument("--spheres", type=int, default=2000, help="rigid spheres of diameter 5 l0, density 2500")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 PCG comparison run")
    ap.add_argument("--workload", choices=["drum", "batched", "precision", "radix"], default="drum",
                    help="drum: the 1M-particle drum (default); batched: 64 independent rotating buckets; "
                         "precision: the mixed-precision drum against the FP64 build after --steps steps")
    ap.add_argument("--buckets", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="batched: keep the default launch shape (256/8)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one domain per GPU and no multi-GPU path is "
              "built; run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return batched(args)
    if args.workload == "precision":
        return precision(args)
    if args.workload == "radix":
        return radix(args)

    from simulation_project_amd import Solver, abi, cases, load_library
    lib = load_library()
    solver = args.solver or "mixed"
    t0 = time.time()
    case = cases.drum(lib, l0=args.l0, steps=args.warmup + args.steps)
    n = case.n
    sph = cases.drum_spheres(case, args.spheres) if args.spheres > 0 else None
    log(f"[bench] drum l0={args.l0} n={n} triangles={len(case.vertices)} built in {time.time() - t0:.1f}s; "
        f"solver {solver}, preconditioner {args.precond}")

    warm = abi.MPSR_WARM_LEARNED if solver == "learned" else abi.MPSR_WARM_PREVIOUS
    kind = "fp64" if solver == "fp64" else "mixed"
    s = Solver(case, precision=abi.MPSR_FP32, pressure_mode=abi.MPSR_PRESSURE_PPE,
               ppe=ppe_options(abi, kind, args.precond, warm))
    if solver == "learned":
        s.mpnn_set_weights(np.load(WEIGHTS).astype(np.float32))
    if sph is not None:
        s.set_spheres(sph)
    s.set_profiling(True)
    log("[bench] warm-up")
    if args.warmup:
        run_steps(s, 0, args.warmup)
    log(f"[bench] timed region: {args.steps} steps")
    elapsed, per_step = run_steps(s, args.warmup, args.steps)
    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    ppe_ms = [t["ppe_ms"] for t in per_step]
    iters = [t["ppe_iterations"] for t in per_step]
    relres = [t["ppe_rel_residual"] for t in per_step]
    launches = sum(t["spmv_launches"] for t in per_step)
    spmv_total_ms = sum(t["spmv_ms"] * t["spmv_launches"] for t in per_step)
    spmv_avg_ms = spmv_total_ms / max(1, launches)
    spmv_b = float(np.mean([t["spmv_bytes"] for t in per_step]))
    achieved = spmv_b / (spmv_avg_ms * 1e-3) / 1e9 if spmv_avg_ms > 0 else 0.0
    phases = {k: round(float(np.mean([t[k] for t in per_step])), 4) for k in
              ("bin_ms", "tri_ms", "stl_ms", "explicit_ms", "relocate_ms", "pnd_ms", "ppe_ms", "gradient_ms",
               "wall_ms", "sphere_ms", "ppe_setup_ms")}
    # whole-step algorithmic bytes: the particle phases (FP32: 442 B per particle) plus the PPE
    # assembly and every PCG iteration (SpMV values and columns, 19 vector words per row)
    nnz = float(np.mean([t["ppe_nnz"] for t in per_step]))
    rows = float(np.mean([t["ppe_rows"] for t in per_step]))
    sv, sx = (4.0, 4.0) if kind == "mixed" else (8.0, 8.0)
    step_bytes = n * 442.0 + rows * 5 * 4 + nnz * (sv + 4) + float(np.mean(iters)) * (nnz * (sv + 4) + rows * 19 * sx)
    step_achieved = step_bytes / (ms_per_step * 1e-3) / 1e9
    log(f"[bench] {ms_per_step:.2f} ms/step, PPE {np.mean(ppe_ms):.2f} ms, iterations {iters}, "
        f"SpMV {spmv_avg_ms * 1e3:.1f} us x {launches}")
    learned_fb = sum(t.get("ppe_learned_fallback", 0) for t in per_step)
    # the coarse operator is rebuilt every 50 solves (and when the iterations grow by a quarter):
    # one more step, after the timed region, with a forced rebuild gives its set-up time
    opt = ppe_options(abi, kind, args.precond, warm)
    every = opt.coarse_rebuild
    opt.coarse_rebuild = 1
    s.set_ppe_options(opt)
    run_steps(s, args.warmup + args.steps, 1)
    tr = s.timings()
    coarse = {"rows": tr["ppe_coarse_rows"], "rebuild_every": every, "setup_ms": round(tr["ppe_setup_ms"], 3),
              "amortized_ms_per_step": round(tr["ppe_setup_ms"] / every, 3),
              "rebuilds_in_timed_steps": int(sum(t["ppe_coarse_rebuilt"] for t in per_step))}
    log(f"[bench] coarse operator: {coarse}")
    s.close()

    plain = None
    if not args.no_plain:
        log("[bench] plain FP64 Jacobi-PCG, same steps")
        p = Solver(case, precision=abi.MPSR_FP32, pressure_mode=abi.MPSR_PRESSURE_PPE,
                   ppe=ppe_options(abi, "fp64", "jacobi", abi.MPSR_WARM_PREVIOUS))
        if sph is not None:
            p.set_spheres(sph)
        p.set_profiling(True)
        if args.warmup:
            run_steps(p, 0, args.warmup)
        pe, pt = run_steps(p, args.warmup, args.steps)
        p.close()
        pm = float(np.mean([t["ppe_ms"] for t in pt]))
        plain = {"ppe_ms_per_step": round(pm, 3), "iterations": [t["ppe_iterations"] for t in pt],
                 "ms_per_step": round(1e3 * pe / args.steps, 3),
                 "ppe_speedup": round(pm / float(np.mean(ppe_ms)), 3)}
        log(f"[bench] plain: {plain}")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            from oracle.oracle import Oracle
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            o = Oracle(case, precision=abi.MPSR_FP32, pressure_mode=abi.MPSR_PRESSURE_PPE,
                       ppe=abi.PpeOptions(abi.MPSR_PCG_FP64, 1e-8, 20000, abi.MPSR_WARM_PREVIOUS, 0.0, 0.0))
            if sph is not None:
                o.set_spheres(sph)
            log(f"[bench] CPU oracle: 1 step of the same drum on {threads} threads")
            tc = time.perf_counter()
            o.step_many(0, 1)
            tcpu = time.perf_counter() - tc
            cpu = {"value": round(n / tcpu, 1), "unit": "particle-updates/s", "cores": threads, "kind": "port",
                   "sample": f"step 0 of the same {n}-particle drum (FP32 operators, FP64 Jacobi-PCG to 1e-8), "
                             f"{tcpu:.1f} s, OpenMP over particles"}
            o.close()
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] CPU baseline skipped: {e}")

    out = {
        "metric": "particle-updates/s at 1M particles; PPE ms/step learned vs plain PCG @1e-8",
        "value": round(value, 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rotating drum, {n} particles (l0={args.l0} m), 9984-triangle STL, 30 rpm, "
                               f"40% fill, {0 if sph is None else len(sph.position)} rigid spheres (d=5 l0, 2500 kg/m3, "
                               f"Schiller-Naumann drag), PPE to 1e-8 ({solver}, {args.precond} preconditioner)",
                   "particles": n, "triangles": len(case.vertices), "spheres": 0 if sph is None else len(sph.position),
                   "dt": case.params.Dt},
        "ppe_ms_per_step": round(float(np.mean(ppe_ms)), 3),
        "ppe_iterations": iters,
        "ppe_rel_residual_max": max(relres),
        "ppe_learned_fallbacks": learned_fb,
        "ppe_entries": {"real": int(nnz), "stored": int(np.mean([t["ppe_stored"] for t in per_step]))},
        "ppe_coarse": coarse,
        "plain_fp64_pcg": plain,
        "phase_ms": phases,
        "step_roofline": {"achieved": round(step_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_achieved / HBM_PEAK_GBS, 4), "bytes_per_step": step_bytes},
        "roofline": {"bound": "hbm", "kernel": "k_pcg_spmv", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": SPMV_TRAFFIC_BYTES.get((args.l0, kind)),
                     "bytes_per_launch": spmv_b, "avg_launch_ms": spmv_avg_ms, "launches": launches},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def batched(args):
    """64 independent rotating buckets (boxes and drums, 5k-200k particles drawn log-uniformly,
    10-60 rpm, l0 = 5e-3 m, EOS pressure, FP32): the launch shape is autotuned first (each shape
    timed on the same saved state, which is then restored), then K steps run as one graph launch
    per step and K more with the same kernels launched directly, bucket by bucket."""
    from simulation_project_amd import Batch, cases, load_library
    lib = load_library()
    t0 = time.time()
    bl = cases.buckets(lib, count=args.buckets)
    n = sum(c.n for c in bl)
    tris = sum(len(c.vertices) for c in bl)
    b = Batch(bl)
    log(f"[bench] {len(bl)} buckets, {n} particles, {tris} triangles, built in {time.time() - t0:.1f}s")
    if args.no_tune:
        blk, items, table = 256, 8, {}
    else:
        blk, items, table = b.autotune(0, 2)
        log(f"[bench] autotune: block {blk}, sort items {items}; " +
            ", ".join(f"{k[0]}/{k[1]}: {v:.2f}" for k, v in sorted(table.items())))
    if args.warmup:
        b.step_many(0, args.warmup)
    b.step_many(args.warmup, args.steps)      # timed on the batch stream with HIP events
    info = b.info()
    ms_graph = info["ms_per_step"]
    b.step_many_eager(args.warmup + args.steps, args.steps)
    ms_eager = b.info()["ms_per_step"]
    log(f"[bench] graph {ms_graph:.3f} ms/step ({info['graph_nodes']} nodes), direct launches {ms_eager:.3f} ms/step")
    b.close()
    out = {
        "metric": "particle-updates/s, 64 batched rotating buckets",
        "value": round(n / (ms_graph * 1e-3), 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(bl)} rotating buckets (boxes of 12 and drums of up to 9984 triangles), "
                               f"{n} particles (log-uniform 5k-200k per bucket), 10-60 rpm, l0=5e-3 m, "
                               f"EOS pressure, one graph launch per step",
                   "buckets": len(bl), "particles": n, "triangles": tris},
        "direct_launch_ms_per_step": round(ms_eager, 4),
        "graph_nodes": info["graph_nodes"],
        "launch": {"block": blk, "sort_items": items,
                   "autotune_ms": {f"{k[0]}/{k[1]}": round(v, 4) for k, v in sorted(table.items())}},
    }
    print(json.dumps(out), flush=True)


# CalcTurbulenceViscosity per particle (FP32): perm, id, position and velocity records (2 x 16 B),
# wall distance and the two results; neighbour records and cell ranges come from L2
TURB_BYTES_PER_PARTICLE = 4 + 4 + 16 + 16 + 4 + 4 + 4


def radix(args):
    """The default drum with the radix solver variant's two terms switched on: the Smagorinsky
    viscosity (Cs = 0.17, filter width l0, clamps 1e-3 m2/s and 1 m2/s2) and the negative-pressure
    term of the rotating drum (influence 2 l0, gas pressure 0, direction 1e-4 along +z).  Same steps with both
    terms off beside it; per-phase HIP-event times of both; the turbulence pass of the CPU oracle
    on the same binning."""
    from simulation_project_amd import Solver, abi, cases, load_library
    lib = load_library()
    case = cases.drum(lib, l0=args.l0, steps=args.warmup + args.steps)
    n = case.n
    sph = cases.drum_spheres(case, args.spheres) if args.spheres > 0 else None
    l0 = case.params.InitialParticleDistance
    terms = dict(bTurbulance=1, SmagorinskyConstant=0.17, FilterWidth=l0, TurbulanViscosityRatio=1e-3, Cnue=0.094,
                 Cepsilon=1.048, TurbulanceIntensity=1.0, bNegativePressure=1)
    # the term scales with DirectionVector and 1/l0^2 (0.5 (Pg + p) Z(l0) / l0^2): at l0 = 2 mm a unit
    # vector and kPa pressures kick wall particles by metres per second per step, so a short vector
    # keeps the drum physical; the kernel's work (every particle's nearest-wall search) is the same
    neg = [dict(RotationInRPM=30.0, NegativePressureCheck=0, InfluenceRegion=2.0, DirectionVector=(0.0, 0.0, 1e-4),
                PressureGas=0.0)]
    log(f"[bench] radix terms: drum l0={args.l0} n={n}")
    res = {}
    for name, on in (("on", True), ("off", False)):
        s = Solver(case, precision=abi.MPSR_FP32, pressure_mode=abi.MPSR_PRESSURE_PPE,
                   ppe=ppe_options(abi, "mixed", "vcycle", abi.MPSR_WARM_PREVIOUS))
        if sph is not None:
            s.set_spheres(sph)
        if on:
            s.set_radix_terms(terms, neg)
        s.set_profiling(True)
        if args.warmup:
            run_steps(s, 0, args.warmup)
        el, per = run_steps(s, args.warmup, args.steps)
        ph = {k: round(float(np.mean([t[k] for t in per])), 4) for k in
              ("turbulence_ms", "explicit_ms", "gradient_ms", "ppe_ms", "pnd_ms")}
        res[name] = {"ms_per_step": 1e3 * el / args.steps, "phase_ms": ph}
        if on:
            nut = s.get_turbulence()[0]
            res[name]["nu_t"] = {"max": float(nut.max()), "mean": float(nut.mean())}
        s.close()
        log(f"[bench] terms {name}: {res[name]['ms_per_step']:.2f} ms/step, {ph}")
    tms = res["on"]["phase_ms"]["turbulence_ms"]
    achieved = n * TURB_BYTES_PER_PARTICLE / (tms * 1e-3) / 1e9 if tms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        try:
            from oracle.oracle import Oracle
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            o = Oracle(case, precision=abi.MPSR_FP32)
            o.set_radix_terms(terms, neg)
            o.RegisterParticleTopology()
            o.RegisterTriangleTopology()
            o.CalcSTLDistance()
            tc = time.perf_counter()
            o.CalcTurbulenceViscosity()
            tcpu = time.perf_counter() - tc
            cpu = {"value": round(n / tcpu, 1), "unit": "particles/s (turbulence pass)", "cores": threads,
                   "kind": "port", "sample": f"one CalcTurbulenceViscosity pass over the {n}-particle drum, "
                                             f"{tcpu:.2f} s, OpenMP over particles"}
            o.close()
        except Exception as e:
            log(f"[bench] CPU baseline skipped: {e}")
    on, off = res["on"], res["off"]
    out = {
        "metric": "particle-updates/s at 1M particles with the radix turbulence and negative-pressure terms",
        "value": round(n / (on["ms_per_step"] * 1e-3), 1),
        "unit": "particle-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(on["ms_per_step"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"rotating drum, {n} particles (l0={args.l0} m), "
                               f"{0 if sph is None else len(sph.position)} spheres, PPE to 1e-8 (mixed, cell V-cycle), "
                               f"Smagorinsky viscosity and negative pressure on", "particles": n},
        "terms_off_ms_per_step": round(off["ms_per_step"], 3),
        "phase_ms": {"on": on["phase_ms"], "off": off["phase_ms"]},
        "nu_t": on.get("nu_t"),
        "roofline": {"bound": "hbm", "kernel": "k_turbulence", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": n * TURB_BYTES_PER_PARTICLE, "avg_launch_ms": tms,
                     "note": "a neighbour operator: about 800 candidates per particle from L2, VALU-bound"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def precision(args):
    """The mixed-precision drum (FP32 particle operators and SpMV, FP64 dots, residuals and
    refinement, cell V-cycle preconditioner) against the FP64 build (FP64 operators, FP64 PCG with
    the same preconditioner, 1e-8) from the same start, both on the GPU, after --steps steps:
    relative L2 and max errors of pressure, velocity and position over the particles alive in
    both, and the bulk kinetic energy and mean pressure.  The FP64 build is the stand-in for the
    FP64 oracle at this size (it equals the oracle bit for bit in every particle operator)."""
    from simulation_project_amd import Solver, abi, cases, load_library
    lib = load_library()
    t0 = time.time()
    case = cases.drum(lib, l0=args.l0, steps=args.steps)
    n = case.n
    sph = cases.drum_spheres(case, args.spheres) if args.spheres > 0 else None
    log(f"[bench] precision: drum l0={args.l0} n={n}, {args.steps} steps, built in {time.time() - t0:.1f}s")
    res = {}
    for name, prec, kind in (("mixed", abi.MPSR_FP32, "mixed"), ("fp64", abi.MPSR_FP64, "fp64")):
        s = Solver(case, precision=prec, pressure_mode=abi.MPSR_PRESSURE_PPE,
                   ppe=ppe_options(abi, kind, "vcycle", abi.MPSR_WARM_PREVIOUS))
        if sph is not None:
            s.set_spheres(sph)
        s.synchronize()
        tt = time.perf_counter()
        for k0 in range(0, args.steps, 50):
            s.step_many(k0, min(50, args.steps - k0))
            log(f"[bench] {name}: {min(k0 + 50, args.steps)} steps")
        s.synchronize()
        el = time.perf_counter() - tt
        pa = s.get_particles(0)
        tl = s.timings()
        res[name] = {"ms_per_step": 1e3 * el / args.steps, "p": pa.pressure.copy(), "v": pa.velocity.copy(),
                     "x": pa.position.copy(), "id": pa.id.copy(),
                     "last_solve": {"iterations": tl["ppe_iterations"], "coarse_rows": tl["ppe_coarse_rows"],
                                    "rel_residual": tl["ppe_rel_residual"]}}
        s.close()
    a, b = res["mixed"], res["fp64"]
    alive = (a["id"] >= 0) & (b["id"] >= 0)
    def rel(x, y):
        return float(np.linalg.norm((x - y)[alive]) / max(np.linalg.norm(y[alive]), 1e-300))
    dv = np.linalg.norm((a["v"] - b["v"])[alive], axis=1)
    dx = np.linalg.norm((a["x"] - b["x"])[alive], axis=1)
    ke_a = 0.5 * float(np.sum(a["v"][alive] ** 2))
    ke_b = 0.5 * float(np.sum(b["v"][alive] ** 2))
    out = {
        "metric": f"mixed-precision drum vs FP64 build after {args.steps} steps",
        "value": rel(a["p"], b["p"]),
        "unit": "relative L2 error of pressure",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(a["ms_per_step"], 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"rotating drum, {n} particles (l0={args.l0} m), "
                               f"{0 if sph is None else len(sph.position)} spheres, PPE to 1e-8, "
                               f"mixed (FP32 operators and SpMV, FP64 dots and refinement) vs FP64",
                   "particles": n},
        "errors": {"pressure_rel_l2": rel(a["p"], b["p"]),
                   "pressure_max_abs_over_max": float(np.abs(a["p"] - b["p"])[alive].max() / np.abs(b["p"][alive]).max()),
                   "velocity_rel_l2": rel(a["v"], b["v"]),
                   "velocity_max_abs": float(dv.max()),
                   "position_rms_over_l0": float(np.sqrt(np.mean(dx ** 2)) / args.l0),
                   "position_max_over_l0": float(dx.max() / args.l0),
                   "kinetic_energy_rel": abs(ke_a - ke_b) / ke_b,
                   "mean_pressure_rel": abs(float(a["p"][alive].mean()) - float(b["p"][alive].mean())) /
                                        abs(float(b["p"][alive].mean())),
                   "alive_mixed": int((a["id"] >= 0).sum()), "alive_fp64": int((b["id"] >= 0).sum())},
        "fp64_ms_per_step": round(b["ms_per_step"], 3),
        "last_solve": {"mixed": a["last_solve"], "fp64": b["last_solve"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
