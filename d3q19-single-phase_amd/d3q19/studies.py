"""Channel studies run on the device: reduced-precision error against FP64.

`precision_study` runs the same seeded channel (log-law mean profile plus a
divergence-free perturbation) in FP64, FP32 and FP16 storage, accumulates the
time-averaged plane statistics of Channel-Flow/saveload.f90:1348-1502
(statistc2: U+ and the Reynolds stresses normalised by u*) on the device, and
reports each variant's error against FP64:

    mean-profile error   max_i |U+_v(i) - U+_64(i)| / max_i U+_64(i)
    stress errors        max_i |<u_a u_b>_v - <u_a u_b>_64| / max_i |<u_a u_b>_64|
    mass drift           change of the mean density fluctuation per node

A perturbed flow is chaotic, so any disturbance at the round-off level of
FP32 moves the trajectory, and with it the time averages, by more than the
precision bias itself.  Two extra runs separate the two effects:

    fp64_jitter   FP64 with u_z disturbed by 1e-7 U_c at t = 0 (the size of an
                  FP32 rounding): its error against FP64 is the sampling floor
                  of the time averages that no precision can beat;
    deterministic the unperturbed log-law start (uniform in y and z, so not
                  chaotic), where a variant's error is its precision alone.

Each run also records the instantaneous U+ error at every tenth of the run.

The study runs the rate set of para.f90:131-140 (MRTtype 3) by default.  The
turbulent branch's MRTtype 1 (para.f90:70, :111-119) is linearly unstable for
a uniform stream along a lattice axis above U ~ 0.1 (growth ~7e-3 per step at
U = 0.13), which the centreline of the Re_tau = 180 channel at nu = 0.0036
reaches (U_c ~ 0.12): a run with a 3-D perturbation blows up once round-off
has grown through that mode, after ~3.7k steps in FP32 and ~7k in FP64.

Usage:  python -m d3q19.studies precision [--steps 20000] [--stats-from 10000] [--mrttype 3]
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Channel, workloads

# statistc2 columns (saveload.f90:1492): 3 = <u_y>/u* (streamwise mean, U+),
# 5, 6, 7 = u'u' variances (x, y, z), 9 = u_x'u_y' shear stress.
COLS = {"U+": 3, "uxux": 5, "uyuy": 6, "uzuz": 7, "uxuy": 9}


def run_channel(prec, lx, ly, lz, steps, stats_from, stats_every, seed=2, mrttype=3, amp=0.10,
                jitter=0.0, log=None):
    """One channel run; returns time-averaged profiles, instantaneous U+
    profiles at every tenth of the run, mass drift and timing."""
    with Channel(lx, ly, lz, prec, mrttype=mrttype) as ch:
        ux, uy, uz = workloads.loglaw_divfree(lx, ly, lz, ch.p["ustar"], ch.p["ystar"], seed=seed, amp=amp)
        if jitter:
            uc = float(np.max(uy))
            uz = uz + jitter * uc * (2.0 * workloads.uniform(99, uz.size).reshape(uz.shape) - 1.0)
        ch.set_macros(ux=ux, uy=uy, uz=uz)
        del ux, uy, uz
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        m0 = ch.mass()
        ch.stats_reset()
        t0 = time.perf_counter()
        done = 0
        inst = []
        tenth = max(steps // 10, 1)
        while done < steps:
            n = min(stats_every, steps - done)
            ch.run(n)
            done += n
            if done >= stats_from:
                ch.stats_accumulate()
            if done % tenth == 0:
                inst.append((done, ch.statistc2()[:, COLS["U+"]].copy()))
                if log:
                    log(f"{prec}{' jitter' if jitter else ''} amp {amp}: step {done}/{steps} "
                        f"mass/node {ch.mass() / (lx * ly * lz):.3e}")
        ch.sync()
        wall = time.perf_counter() - t0
        prof, ns = ch.stats_read()
        m1 = ch.mass()
        return {"profile": prof, "samples": ns, "inst": inst, "mass_drift_per_node": (m1 - m0) / (lx * ly * lz),
                "wall_s": wall, "mlups": lx * ly * lz * steps / wall / 1e6}


def compare(p, ref):
    out = {}
    for name, c in COLS.items():
        scale = np.max(np.abs(ref[:, c]))
        out[name] = float(np.max(np.abs(p[:, c] - ref[:, c])) / scale) if scale > 0 else 0.0
    return out


def _inst_errors(run, ref):
    return [[s, float(np.max(np.abs(a - b)) / np.max(np.abs(b)))] for (s, a), (_, b) in zip(run["inst"], ref["inst"])]


def precision_study(lx=192, ly=768, lz=384, steps=20000, stats_from=10000, stats_every=100,
                    precs=("fp64", "fp32", "fp16"), mrttype=3, jitter=1e-7, deterministic=True, log=None):
    cases = {}
    kw = dict(mrttype=mrttype, log=log)
    for prec in precs:
        cases[prec] = run_channel(prec, lx, ly, lz, steps, stats_from, stats_every, **kw)
    if jitter:
        cases["fp64_jitter"] = run_channel("fp64", lx, ly, lz, steps, stats_from, stats_every, jitter=jitter, **kw)
    if deterministic:
        for prec in precs:
            cases["deterministic_" + prec] = run_channel(prec, lx, ly, lz, steps, stats_from, stats_every,
                                                         amp=0.0, **kw)
    report = {"grid": [lx, ly, lz], "steps": steps, "stats_from": stats_from, "stats_every": stats_every,
              "mrttype": mrttype, "jitter": jitter, "variants": {}}
    for name, r in cases.items():
        ref = cases["deterministic_fp64" if name.startswith("deterministic_") else "fp64"]
        entry = {
            "errors_vs_fp64": compare(r["profile"], ref["profile"]) if r is not ref else None,
            "instantaneous_U+_error": _inst_errors(r, ref) if r is not ref else None,
            "mass_drift_per_node": r["mass_drift_per_node"],
            "samples": r["samples"],
            "wall_s": round(r["wall_s"], 2),
            "mlups_incl_stats": round(r["mlups"], 1),
            "centreline_U+": float(r["profile"][lx // 2, COLS["U+"]]),
        }
        if name.startswith("deterministic_"):
            # no fluctuations: only the mean profile is meaningful
            entry["errors_vs_fp64"] = None if r is ref else {"U+": entry["errors_vs_fp64"]["U+"]}
        report["variants"][name] = entry
    return report, cases


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("study", choices=["precision"])
    ap.add_argument("--grid", default="192,768,384")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--stats-from", type=int, default=10000)
    ap.add_argument("--stats-every", type=int, default=100)
    ap.add_argument("--mrttype", type=int, default=3, choices=[1, 2, 3])
    ap.add_argument("--jitter", type=float, default=1e-7, help="control run's disturbance / U_c (0: no control)")
    ap.add_argument("--no-deterministic", action="store_true", help="skip the unperturbed runs")
    ap.add_argument("--out", default=None, help="also write the report (JSON) and profiles (.npz) here")
    args = ap.parse_args(argv)
    lx, ly, lz = (int(v) for v in args.grid.split(","))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    report, res = precision_study(lx, ly, lz, args.steps, args.stats_from, args.stats_every,
                                  mrttype=args.mrttype, jitter=args.jitter,
                                  deterministic=not args.no_deterministic, log=log)
    print(json.dumps(report, indent=1))
    if args.out:
        with open(args.out + ".json", "w") as fh:
            json.dump(report, fh, indent=1)
        np.savez(args.out + ".npz", **{k: v["profile"] for k, v in res.items()})


if __name__ == "__main__":
    main()
