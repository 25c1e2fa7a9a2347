"""Throughput of the D3Q19 MRT channel step on one MI355X.

A step is one iteration of the reference's main loop body
(Channel-Flow/main.f90:157-161: collision_MRT with its periodic exchange,
then macrovar) over the whole lattice, run by liblbm's fused kernel.

Workload (default): turbulent channel Re_tau = 180, nu = 0.0036, MRTtype 1,
768 streamwise x 384 spanwise x 192 wall-normal nodes (reference axes
lx=192, ly=768, lz=384; 56.6 M nodes), log-law mean profile plus a seeded
divergence-free random-phase perturbation of 10% amplitude, populations
resident in HBM before timing.  `value` is FP64 MLUPS; FP32, FP16-storage and
FP16C-storage (1 sign, 4 exponent, 11 significand bits) variants of the same
workload are reported under "variants".  The timed
steps stay far below the ~4k steps after which MRTtype 1's linear
instability at this centreline speed shows (the rates do not change the
kernel's cost).

--workload c4: the large channel Re_tau = 395 on 2048 x 1024 x 512 nodes
(1.07 G); `value` is FP32 MLUPS there (FP64 lattices would need 326 GB) and
FP16 storage is the variant.

`value` runs lbm_run's two-step blocked sweep (pairs of steps through a
ring of slabs held in the Infinity Cache -- y slabs in FP64, z slabs in
FP32 -- bitwise equal to single steps); "single_step" is the same workload
with the sweep off.

"forcingp": the same step under the perturbation forcing FORCINGP
(Channel-Flow/collision.f90:529-602), its force evaluated per node inside
the step, timed by wall clock around the host loop.

"closure_coarse": the 2x-coarse channel (384 x 192 x 96, nu = 0.0018) with
the fused learned closure, FP32 storage; its ratio of wall time per unit of
physical time against the fine grid (one coarse step = two fine steps, 1/8
of the nodes) bounds the time-to-statistics speed-up of the closure run.

Roofline: the step kernel is HBM-bound; algorithmic traffic is one read and
one write of the 19 populations per node (304 B FP64, 152 B FP32, 76 B FP16).
`achieved` = algorithmic bytes per step / mean device time per step from HIP
events recorded on the library's stream around every single step or blocked
pair (halved); for the blocked sweep this is an effective bandwidth, since
the Infinity Cache serves the intermediate lattice.  `traffic` is the
rocprofv3 FETCH_SIZE + WRITE_SIZE figure between L2 and the memory fabric.

"per_call": the same FP64 steps driven as the reference's main loop calls
them, collision_MRT then macrovar per step through the C ABI (the library
defers and pairs them into blocked sweeps), wall clock around the loop.

"particles": the coupled step with 2000 spheres of radius 5 at the same
turbulent constants (--particles 0 leaves it out).

"diagnostics": diag, statistc2 and dissipation_calc reduced on the device,
ms per call, beside the download of the four macroscopic fields into host
arrays that a host-side evaluation of them starts from.

cpu_baseline: the CPU oracle's out-of-place schedule (same arithmetic as the
reference, OpenMP over z planes) on the host cores, on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "d3q19-single-phase_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import d3q19  # noqa: E402
from d3q19 import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = {"fp64": 304, "fp32": 152, "fp16": 76, "fp16c": 76}
WORKLOADS = {
    # name: (lx, ly, lz, description)
    "c2": (192, 768, 384, "turbulent channel Re_tau=180, 768x384x192 (streamwise x spanwise x wall-normal)"),
    "c1": (32, 64, 32, "channel 64x32x32 (streamwise x spanwise x wall-normal)"),
    "c5": (96, 384, 192, "coarse channel 384x192x96 (streamwise x spanwise x wall-normal)"),
    # 1.07 G nodes: FP32 lattices 2 x 81.6 GB, FP16 2 x 40.8 GB; FP64 (2 x 163 GB) does not fit in 288 GB
    "c4": (512, 2048, 1024, "large channel Re_tau=395, 2048x1024x512 (streamwise x spanwise x wall-normal)"),
}
RSTAR = {"c4": 395.0}  # Re_tau per workload (para.f90:62 otherwise: 180)
MAIN_PREC = {"c4": "fp32"}  # precision of `value` (FP64 unless it does not fit)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(prec, workload):
    """L2-to-fabric bytes per step measured by rocprofv3 --pmc (committed summary), or None."""
    path = os.path.join(ROOT, "profiles", f"step_{prec}_{workload}_traffic.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("hbm_bytes_per_step")
    return None


def init_channel(ch, lx, ly, lz):
    """Log-law mean profile plus the seeded divergence-free perturbation (10%),
    summed on the device from plane x profile terms; returns the terms."""
    terms = workloads.loglaw_divfree_terms(lx, ly, lz, ch.p["ustar"], ch.p["ystar"], seed=2)
    ch.set_velocity_separable(terms)
    ch.FORCING()
    ch.initpop()
    ch.macrovar()
    return terms


def measure(prec, workload, steps, warmup, single_steps=False):
    """Main-loop steps on the workload; single_steps=True turns lbm_run's
    two-step blocked sweep off (LBM_BLOCK_PLANES=0 while the handle is made)."""
    lx, ly, lz, _ = WORKLOADS[workload]
    n = lx * ly * lz
    saved = os.environ.get("LBM_BLOCK_PLANES")
    if single_steps:
        os.environ["LBM_BLOCK_PLANES"] = "0"
    try:
        ch = d3q19.Channel(lx, ly, lz, prec, rstar=RSTAR.get(workload))
    finally:
        if single_steps:
            if saved is None:
                os.environ.pop("LBM_BLOCK_PLANES", None)
            else:
                os.environ["LBM_BLOCK_PLANES"] = saved
    with ch:
        axis, slabs = ch.block_info()
        rows = ch.block_rows()
        init_channel(ch, lx, ly, lz)
        m0 = ch.mass()
        ch.run(warmup)
        ch.sync()
        t0 = time.perf_counter()
        avg_ms, total_ms = ch.time_steps(steps)
        ch.sync()
        wall = time.perf_counter() - t0
        m1 = ch.mass()
    mlups = n * steps / wall / 1e6
    algo = n * BYTES_PER_NODE[prec]
    achieved = algo / (avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(prec, workload)
    return {
        "mlups": mlups,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": avg_ms,
        "event_total_ms": total_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": algo,
                     # achieved: one lattice read and write per step (the single
                     # step's HBM traffic) over the device time per step; a
                     # blocked pair keeps its intermediate lattice in the
                     # Infinity Cache, so this is an effective bandwidth
                     "basis": "algorithmic bytes per step (read + write of 19 populations) / device time per step",
                     # traffic: rocprofv3 FETCH_SIZE + WRITE_SIZE, measured
                     # between L2 and the memory fabric: an upper bound on HBM
                     # bytes when the Infinity Cache serves the ring
                     "traffic_point": "L2-to-fabric (upper bound on HBM bytes)"},
        # change of the mean density fluctuation (sum of shifted populations / N)
        "mass_drift_per_node": (m1 - m0) / n,
        # lbm_run's pairing of steps: sweep axis, slabs per chunk (0: single
        # steps) and rows per window of the slabs (0: whole slabs)
        "blocked_sweep": {"axis": axis, "slabs_per_chunk": slabs, "rows_per_window": rows},
    }


def measure_per_call(prec, workload, steps, warmup):
    """The main loop as a driver calling the reference's subroutines one at a
    time (collision_MRT then macrovar per step, Channel-Flow/main.f90:157-161)
    through the C ABI; the library pairs those calls into blocked sweeps.
    Wall clock around the host loop."""
    lx, ly, lz, _ = WORKLOADS[workload]
    n = lx * ly * lz
    with d3q19.Channel(lx, ly, lz, prec, rstar=RSTAR.get(workload)) as ch:
        init_channel(ch, lx, ly, lz)
        for _ in range(warmup):
            ch.collision_MRT()
            ch.macrovar()
        ch.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ch.collision_MRT()
            ch.macrovar()
        ch.sync()
        wall = time.perf_counter() - t0
    achieved = n * BYTES_PER_NODE[prec] / (wall / steps) / 1e9
    return {"mlups": n * steps / wall / 1e6, "ms_per_step": wall / steps * 1e3,
            "wall_frac": achieved / HBM_PEAK_GBS}


def measure_forcingp(prec, workload, steps, warmup):
    """The same step under the perturbation forcing FORCINGP (collision.f90:
    529-602), the force evaluated per node inside the step: one FORCINGP,
    collision_MRT and macrovar per step from the host, wall-clock timed."""
    lx, ly, lz, _ = WORKLOADS[workload]
    n = lx * ly * lz
    with d3q19.Channel(lx, ly, lz, prec, rstar=RSTAR.get(workload)) as ch:
        init_channel(ch, lx, ly, lz)
        ch.run_forcingp(0, warmup, 10**9)
        ch.sync()
        t0 = time.perf_counter()
        ch.run_forcingp(warmup, steps, 10**9)
        ch.sync()
        wall = time.perf_counter() - t0
    achieved = n * BYTES_PER_NODE[prec] / (wall / steps) / 1e9
    return {"mlups": n * steps / wall / 1e6, "ms_per_step": wall / steps * 1e3,
            "wall_frac": achieved / HBM_PEAK_GBS}


def measure_closure(steps, warmup):
    """The 2x-coarse channel (96 x 384 x 192, nu = 0.0018, rate set 3) with the
    fused closure, FP32 storage.  The kernel cost does not depend on the
    weights, so a seeded network stands in for a trained one."""
    lx, ly, lz, _ = WORKLOADS["c5"]
    n = lx * ly * lz
    with d3q19.Channel(lx, ly, lz, "fp32", mrttype=3, visc=0.0018) as ch:
        init_channel(ch, lx, ly, lz)
        ch.closure_set(d3q19.closure_params(seed=1, nu=0.0018), ystar=ch.p["ystar"])
        ch.run(warmup)
        ch.sync()
        t0 = time.perf_counter()
        avg_ms, _ = ch.time_steps(steps)
        ch.sync()
        wall = time.perf_counter() - t0
    return {"mlups": n * steps / wall / 1e6, "kernel_ms": avg_ms, "nodes": n}


def measure_diagnostics(prec, workload, reps=5):
    """Output-cadence work on the device, ms per call: diag
    (Channel-Flow/saveload.f90:1507-1676), statistc2 (:1348-1502) and
    dissipation_calc (:4644-4882), each reduced to a few hundred values on the
    device; beside them the download of macrovar's four fields (4 x N doubles
    into preallocated host arrays) that a host-side evaluation starts with."""
    lx, ly, lz, _ = WORKLOADS[workload]
    with d3q19.Channel(lx, ly, lz, prec, rstar=RSTAR.get(workload)) as ch:
        init_channel(ch, lx, ly, lz)
        ch.run(2)
        host = [np.empty((lz, ly, lx)) for _ in range(4)]

        def timed(fn, n):
            fn()  # first call: allocations, page faults
            ch.sync()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            ch.sync()
            return round((time.perf_counter() - t0) / n * 1e3, 3)

        out = {"diag_ms": timed(ch.diag, reps), "statistc2_ms": timed(ch.statistc2, reps),
               "dissipation_calc_ms": timed(ch.dissipation, reps),
               "fields_download_ms": timed(lambda: ch.macrovar(download=True, out=host), 2),
               "fields_download_bytes": 4 * 8 * lx * ly * lz}
    return out


def measure_particles(prec, workload, npart, rad, steps, warmup):
    """Coupled particle step (beads_collision, collision_MRT with solid nodes
    skipped, macrovar, beads_move, beads_links, beads_filling) with npart
    seeded neutrally buoyant spheres of radius rad on the workload's grid, at
    the turbulent constants of the main workload (tau = 0.5108, rate set 1,
    log-law start with its divergence-free perturbation), lubrication on.
    The spheres start at the fluid velocity of the node nearest their
    centre."""
    lx, ly, lz, _ = WORKLOADS[workload]
    n = lx * ly * lz
    pos = workloads.sphere_positions(lx, ly, lz, npart, rad, seed=3)
    with d3q19.Channel(lx, ly, lz, prec, rstar=RSTAR.get(workload)) as ch:
        terms = init_channel(ch, lx, ly, lz)
        ix, iy, iz = (np.floor(pos[:, k]).astype(int) for k in range(3))
        vel = np.stack(workloads.separable_at(terms, ix, iy, iz), axis=1)
        ch.particles_create(pos, vel=vel, rad=rad, visc=ch.p["visc"])
        solid = int(np.count_nonzero(ch.particles_solid() >= 0))
        ch.run_particles(warmup)
        ch.sync()
        t0 = time.perf_counter()
        ch.run_particles(steps)
        ch.sync()
        wall = time.perf_counter() - t0
        st = ch.particles_get()
    return {"npart": npart, "radius": rad, "solid_fraction": round(solid / n, 5),
            "value": round(n * steps / wall / 1e6, 1), "unit": "MLUPS", "ms_per_step": round(wall / steps * 1e3, 4),
            "regime": "turbulent constants (tau 0.5108), spheres released at the local fluid velocity",
            "max_particle_speed": round(float(np.abs(st["vel"]).max()), 5)}


COPY_KERNEL_TAG = "MulFunctor"  # torch.mul(src, 2.0, out=dst): reads and writes 8 B per element
PROBE_STEPS = 8  # steps in the traffic probe (even: whole blocked pairs)
STEP_KERNEL_TAGS = ("k_step<", "k_step_planes<")  # single-step kernel, blocked-sweep launches


def traffic_probe(prec, workload):
    """A short run for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: PROBE_STEPS
    main-loop steps (single steps or blocked pairs, as lbm_run runs them),
    then one elementwise kernel of known traffic (read 2 GiB, write 2 GiB)
    that calibrates the counters."""
    import torch
    lx, ly, lz, _ = WORKLOADS[workload]
    steps = PROBE_STEPS
    with d3q19.Channel(lx, ly, lz, prec) as ch:
        init_channel(ch, lx, ly, lz)
        ch.run(steps)
        ch.sync()
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    info = {"prec": prec, "workload": workload, "step_launches": steps, "copy_read_bytes": n * 8,
            "copy_write_bytes": n * 8, "copy_kernel_tag": COPY_KERNEL_TAG}
    print(json.dumps(info), flush=True)


def _counter_means(csv_path, counter):
    """`counter` summed over the step kernels' dispatches per step, and its
    mean per dispatch of the copy kernel."""
    import csv
    step, copy = [], []
    with open(csv_path, newline="") as fh:
        for row in csv.DictReader(fh):
            if row.get("Counter_Name") != counter:
                continue
            name, v = row.get("Kernel_Name", ""), float(row["Counter_Value"])
            if any(t in name for t in STEP_KERNEL_TAGS):
                step.append(v)
            elif COPY_KERNEL_TAG in name:
                copy.append(v)
    if not step or not copy:
        raise SystemExit(f"{csv_path}: no {counter} rows for the step ({len(step)}) or copy ({len(copy)}) kernel")
    return sum(step) / PROBE_STEPS, sum(copy) / len(copy), len(step)


def traffic_summary(prec, workload, fetch_csv, write_csv):
    """Bytes per step between L2 and the memory fabric from the two counter
    runs, each counter scaled by what it reads for the copy kernel of known
    traffic.  With the blocked sweep the 256 MB Infinity Cache behind this
    point serves part of it, so the figure bounds HBM traffic from above."""
    lx, ly, lz, _ = WORKLOADS[workload]
    algo = lx * ly * lz * BYTES_PER_NODE[prec]
    f_step, f_copy, nf = _counter_means(fetch_csv, "FETCH_SIZE")
    w_step, w_copy, nw = _counter_means(write_csv, "WRITE_SIZE")
    copy_bytes = (1 << 28) * 8
    f_scale = copy_bytes / (f_copy * 1024.0)
    w_scale = copy_bytes / (w_copy * 1024.0)
    fetch = f_step * 1024.0 * f_scale
    write = w_step * 1024.0 * w_scale
    return {"prec": prec, "workload": workload, "step_dispatches": [nf, nw], "steps": PROBE_STEPS,
            "fetch_kib_per_step": f_step, "write_kib_per_step": w_step,
            "copy_fetch_kib": f_copy, "copy_write_kib": w_copy,
            "fetch_scale": f_scale, "write_scale": w_scale,
            "fetch_bytes": fetch, "write_bytes": write,
            "hbm_bytes_per_step": round(fetch + write), "algorithmic_bytes": algo,
            "traffic_over_algorithmic": (fetch + write) / algo}


# The oracle's step against the compiled reference on one core of the build
# container (Intel Xeon 2.7 GHz): 5.35 against 5.45 MLUPS on 192x384x96.
PORT_VS_REFERENCE_1CORE = round(5.35 / 5.45, 3)


def host_cores():
    """Cores this process may use: its affinity set, capped by OMP_NUM_THREADS
    when that is set (a job scheduler that gives one GPU's job 16 host
    threads sets it to 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(sample_steps, threads):
    """Oracle out-of-place step (collision + macrovar) on the full default grid."""
    import oracle
    lx, ly, lz, _ = WORKLOADS["c2"]
    o = oracle.OracleChannel(lx, ly, lz)
    o.FORCING()
    ux, uy, uz = workloads.loglaw_divfree(lx, ly, lz, o.p.ustar, o.p.ystar, seed=2)
    o.ux[:], o.uy[:], o.uz[:] = ux, uy, uz
    del ux, uy, uz
    o.initpop()
    o.macrovar()
    o.run(1, ab=True, nthreads=threads)  # warm the pages
    t0 = time.perf_counter()
    o.run(sample_steps, ab=True, nthreads=threads)
    dt = time.perf_counter() - t0
    n = lx * ly * lz
    return {"value": round(n * sample_steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{sample_steps} steps (collision_MRT + macrovar) of the full {lx}x{ly}x{lz} FP64 grid, "
                      f"oracle out-of-place schedule, OpenMP {threads} threads (every core this job may use), "
                      f"{dt:.1f} s",
            "port_vs_reference_1core": PORT_VS_REFERENCE_1CORE}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--grid", default=None, metavar="LX,LY,LZ",
                    help="measure this grid instead of the workload's (shape studies; Re_tau = 180)")
    ap.add_argument("--variants", default="fp32,fp16,fp16c",
                    help="comma list of extra storage precisions ('' = none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-closure", action="store_true", help="skip the coarse closure-run throughput")
    ap.add_argument("--no-forcingp", action="store_true", help="skip the perturbation-forcing step throughput")
    ap.add_argument("--no-diagnostics", action="store_true", help="skip the device diagnostics timing")
    ap.add_argument("--particles", type=int, default=2000, metavar="NPART",
                    help="also time the coupled particle step with NPART spheres of radius 5 (FP64 and FP32)")
    ap.add_argument("--cpu-steps", type=int, default=26)
    ap.add_argument("--traffic-probe", metavar="PREC", choices=sorted(BYTES_PER_NODE),
                    help="run the short program profiled by rocprofv3 --pmc (5 step launches + calibration copy)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 counter_collection.csv files (FETCH_SIZE, WRITE_SIZE runs of "
                         "--traffic-probe) into profiles/step_<prec>_<workload>_traffic.json")
    ap.add_argument("--prec", default="fp64", choices=sorted(BYTES_PER_NODE), help="precision for --traffic-from")
    args = ap.parse_args()
    if args.grid:
        lx, ly, lz = (int(v) for v in args.grid.split(","))
        WORKLOADS["grid"] = (lx, ly, lz, f"channel {ly}x{lz}x{lx} (streamwise x spanwise x wall-normal)")
        args.workload = "grid"
    if args.traffic_probe:
        traffic_probe(args.traffic_probe, args.workload)
        return
    if args.traffic_from:
        s = traffic_summary(args.prec, args.workload, *args.traffic_from)
        path = os.path.join(ROOT, "profiles", f"step_{args.prec}_{args.workload}_traffic.json")
        with open(path, "w") as fh:
            json.dump(s, fh, indent=1)
        print(json.dumps(s, indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "the slab decomposition over several GPUs is designed, not implemented", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    main_prec = MAIN_PREC.get(args.workload, "fp64")
    log(f"bench: workload {args.workload} {main_prec} steps={args.steps} warmup={args.warmup}")
    main_m = measure(main_prec, args.workload, args.steps, args.warmup)
    log(f"bench: {main_prec} {main_m['mlups']:.1f} MLUPS, kernel {main_m['kernel_ms']:.3f} ms, "
        f"{main_m['roofline']['achieved']:.0f} GB/s")
    single = measure(main_prec, args.workload, args.steps, args.warmup, single_steps=True)
    log(f"bench: {main_prec} single steps {single['mlups']:.1f} MLUPS, {single['roofline']['achieved']:.0f} GB/s")
    pc = measure_per_call(main_prec, args.workload, args.steps, args.warmup)
    per_call = {"value": round(pc["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(pc["ms_per_step"], 4),
                "hbm_frac_wall": round(pc["wall_frac"], 4),
                "calls": "collision_MRT + macrovar per step through the C ABI (paired by the library)"}
    log(f"bench: {main_prec} per-call loop {pc['mlups']:.1f} MLUPS ({pc['wall_frac']:.3f} of 8 TB/s)")
    variants = {}
    for prec in [v for v in args.variants.split(",") if v and v != main_prec]:
        log(f"bench: variant {prec}")
        v = measure(prec, args.workload, args.steps, args.warmup)
        variants[prec] = {"value": round(v["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(v["ms_per_step"], 4),
                          "kernel_ms": round(v["kernel_ms"], 4), "roofline": v["roofline"],
                          f"speedup_vs_{main_prec}": round(v["mlups"] / main_m["mlups"], 3),
                          "mass_drift_per_node": v["mass_drift_per_node"], "blocked_sweep": v["blocked_sweep"]}
        log(f"bench: {prec} {v['mlups']:.1f} MLUPS, {v['roofline']['achieved']:.0f} GB/s")

    forcingp = None
    if not args.no_forcingp:
        log(f"bench: perturbation forcing (FORCINGP), {main_prec}")
        fp = measure_forcingp(main_prec, args.workload, args.steps, args.warmup)
        forcingp = {"value": round(fp["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(fp["ms_per_step"], 4),
                    "hbm_frac_wall": round(fp["wall_frac"], 4),
                    "vs_uniform_force": round(fp["mlups"] / main_m["mlups"], 3)}
        log(f"bench: forcingp {fp['mlups']:.1f} MLUPS")

    diagnostics = None
    if not args.no_diagnostics:
        log(f"bench: device diagnostics, {main_prec}")
        diagnostics = measure_diagnostics(main_prec, args.workload)
        log(f"bench: diag {diagnostics['diag_ms']} ms, statistc2 {diagnostics['statistc2_ms']} ms, "
            f"dissipation_calc {diagnostics['dissipation_calc_ms']} ms, fields download "
            f"{diagnostics['fields_download_ms']} ms")

    closure = None
    if args.workload == "c2" and not args.no_closure:
        log("bench: 2x-coarse channel with the fused closure")
        c = measure_closure(args.steps, args.warmup)
        # one coarse step covers two fine steps of physical time on 1/8 of the nodes
        per_phys = lambda mlups, nodes, steps_per_unit: nodes * steps_per_unit / mlups  # noqa: E731
        fine_n = WORKLOADS["c2"][0] * WORKLOADS["c2"][1] * WORKLOADS["c2"][2]
        closure = {"workload": WORKLOADS["c5"][3], "value": round(c["mlups"], 1), "unit": "MLUPS",
                   "kernel_ms": round(c["kernel_ms"], 4),
                   "time_per_physical_time_vs_fp64_fine": round(
                       per_phys(main_m["mlups"], fine_n, 2) / per_phys(c["mlups"], c["nodes"], 1), 2)}
        if "fp32" in variants:
            closure["time_per_physical_time_vs_fp32_fine"] = round(
                per_phys(variants["fp32"]["value"], fine_n, 2) / per_phys(c["mlups"], c["nodes"], 1), 2)
        log(f"bench: closure coarse {c['mlups']:.1f} MLUPS")

    particles = None
    if args.particles and args.workload == "c2":
        particles = {}
        for prec in ("fp64", "fp32"):
            log(f"bench: {args.particles} particles, {prec}")
            try:
                particles[prec] = measure_particles(prec, args.workload, args.particles, 5.0, args.steps,
                                                    args.warmup)
                log(f"bench: particles {prec} {particles[prec]['value']} MLUPS")
            except d3q19.LbmError as e:  # reported in the line, not fatal to the headline figure
                particles[prec] = {"error": str(e)}
                log(f"bench: particles {prec} failed: {e}")

    cpu = None
    if not args.no_cpu and rank == 0:
        threads = host_cores()
        log(f"bench: cpu baseline, {threads} threads")
        cpu = cpu_baseline(args.cpu_steps, threads)

    lx, ly, lz, desc = WORKLOADS[args.workload]
    out = {
        "metric": f"MLUPS ({main_prec.upper()} D3Q19 MRT channel step) + % of HBM peak",
        "value": round(main_m["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_m["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16 storage, f32 compute",
                  "fp16c": "16-bit (1-4-11) storage, f32 compute"}[main_prec],
        "data": "synthetic (seeded log-law + divergence-free perturbation)",
        "config": {"workload": desc, "lx": lx, "ly": ly, "lz": lz, "nodes": lx * ly * lz,
                   "kernel_ms": round(main_m["kernel_ms"], 4),
                   "mass_drift_per_node": main_m["mass_drift_per_node"],
                   "blocked_sweep": main_m["blocked_sweep"]},
        "roofline": main_m["roofline"],
        "cpu_baseline": cpu,
        "single_step": {"value": round(single["mlups"], 1), "unit": "MLUPS", "kernel_ms": round(single["kernel_ms"], 4),
                        "roofline_frac": single["roofline"]["frac"]},
        "per_call": per_call,
        "variants": variants,
        "closure_coarse": closure,
        "forcingp": forcingp,
        "diagnostics": diagnostics,
        "particles": particles,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
