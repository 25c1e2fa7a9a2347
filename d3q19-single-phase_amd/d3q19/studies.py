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
from .params import para

# statistc2 columns (saveload.f90:1492): 3 = <u_y>/u* (streamwise mean, U+),
# 5, 6, 7 = u'u' variances (x, y, z), 9 = u_x'u_y' shear stress.
COLS = {"U+": 3, "uxux": 5, "uyuy": 6, "uzuz": 7, "uxuy": 9}


def run_channel(prec, lx, ly, lz, steps, stats_from, stats_every, seed=2, mrttype=3, amp=0.10,
                jitter=0.0, rstar=None, visc=None, perturb=None, log=None):
    """One channel run; returns time-averaged profiles, instantaneous U+
    profiles at every tenth of the run, mass drift and timing.  perturb: the
    keyword arguments of workloads.loglaw_divfree for the start perturbation
    (default: amplitude amp, the module's other defaults)."""
    with Channel(lx, ly, lz, prec, mrttype=mrttype, rstar=rstar, visc=visc) as ch:
        pert = dict(amp=amp) if perturb is None else dict(perturb)
        ux, uy, uz = workloads.loglaw_divfree(lx, ly, lz, ch.p["ustar"], ch.p["ystar"], seed=seed, **pert)
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
                    log(f"{prec}{' jitter' if jitter else ''} amp {pert['amp']}: step {done}/{steps} "
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
                    precs=("fp64", "fp32", "fp16"), mrttype=3, jitter=1e-7, deterministic=True, rstar=None,
                    visc=None, perturb=None, log=None):
    """Each precision on the same seeded start; errors of its time averages
    against FP64.  With perturb=TURBULENT_START the flow is turbulent
    (chaotic): trajectories of different precisions separate, and only the
    statistics compare.  The FP64 run with a round-off-sized disturbance
    (jitter) is then a second realisation of the same flow, and its distance
    from FP64 is the sampling scatter that a precision's error is measured
    against; each run's wall-to-wall asymmetry is reported beside it."""
    cases = {}
    kw = dict(mrttype=mrttype, rstar=rstar, visc=visc, perturb=perturb, log=log)
    for prec in precs:
        cases[prec] = run_channel(prec, lx, ly, lz, steps, stats_from, stats_every, **kw)
    if jitter:
        cases["fp64_jitter"] = run_channel("fp64", lx, ly, lz, steps, stats_from, stats_every, jitter=jitter, **kw)
    if deterministic:
        for prec in precs:
            cases["deterministic_" + prec] = run_channel(prec, lx, ly, lz, steps, stats_from, stats_every,
                                                         **dict(kw, perturb=dict(amp=0.0)))
    report = {"grid": [lx, ly, lz], "steps": steps, "stats_from": stats_from, "stats_every": stats_every,
              "rstar": 180.0 if rstar is None else rstar, "visc": 0.0036 if visc is None else visc,
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
        else:
            entry["wall_to_wall_asymmetry"] = _asymmetry(r["profile"])
        report["variants"][name] = entry
    if perturb is not None:
        report["perturbation"] = {k: list(v) if isinstance(v, tuple) else v for k, v in perturb.items()}
    return report, cases


# A start that becomes turbulent within ~10^4 steps on the 192 x 768 x 384
# grid: near-wall divergence-free perturbation of 30 % of U_c with spanwise
# wavelengths down to 48 nodes (~90 wall units).
TURBULENT_START = dict(amp=0.3, nmodes=16, kmax=(4, 8), near_wall=True)


def _start(ch, perturb, seed):
    ux, uy, uz = workloads.loglaw_divfree(ch.lx, ch.ly, ch.lz, ch.p["ustar"], ch.p["ystar"], seed=seed, **perturb)
    ch.set_macros(ux=ux, uy=uy, uz=uz)
    del ux, uy, uz
    ch.FORCING()
    ch.initpop()
    ch.macrovar()


# statistc2 columns that change sign under the reflection x -> lx - x (u_x -> -u_x)
_ANTISYMMETRIC = (2, 8, 9)


def fold(p):
    """Half-channel profile from both walls: plane i and plane lx-1-i averaged
    (sign flipped for <u_x>, <u_x u_z>, <u_x u_y>)."""
    h = p.shape[0] // 2
    q = 0.5 * (p[:h] + p[::-1][:h])
    for c in _ANTISYMMETRIC:
        q[:, c] = 0.5 * (p[:h, c] - p[::-1][:h, c])
    return q


def _stat_errors(p, ref):
    """Mean-profile and second-moment errors of a coarse run against the
    filtered fine statistics on the folded half channel, each as
    max over planes |v - v_ref| / max |v_ref|."""
    return compare(fold(p), fold(ref))


def _asymmetry(p):
    """Difference between the two walls' halves of one run (its own sampling
    noise), in the units of _stat_errors."""
    h = p.shape[0] // 2
    a, b = p[:h].copy(), p[::-1][:h].copy()
    for c in _ANTISYMMETRIC:
        b[:, c] = -b[:, c]
    return compare(a, b)


def _coarse_run(lx, ly, lz, nu_c, steps, stats_from, stats_every, perturb, seed, closure=None, log=None,
                half_at=None, snap_at=(), on_snap=None):
    """One coarse run with statistics after stats_from; on_snap(ch) is called
    after the steps in snap_at (its time is not counted in wall_s)."""
    with Channel(lx, ly, lz, "fp32", mrttype=3, visc=nu_c) as ch:
        _start(ch, perturb, seed)
        if closure is not None:
            ch.closure_set(closure, ystar=ch.p["ystar"])
        ch.stats_reset()
        ch.sync()
        t0 = time.perf_counter()
        done = 0
        first_half = None
        t_snap = 0.0
        while done < steps:
            ch.run(stats_every)
            done += stats_every
            if done > stats_from:
                ch.stats_accumulate()
            if done == half_at:
                first_half, _ = ch.stats_read()
            if done in snap_at:
                ch.sync()
                ts = time.perf_counter()
                on_snap(ch)
                t_snap += time.perf_counter() - ts
            if log and done % max(steps // 5, 1) == 0:
                log(f"coarse{' + closure' if closure is not None else ''}: step {done}/{steps}")
        ch.sync()
        wall = time.perf_counter() - t0 - t_snap
        stats, n = ch.stats_read()
        nut_plane = nut_plane_std = None
        if closure is not None:  # the closure's nu_t/nu by plane on the last state
            ch.macrovar()
            nut = ch.closure_nut().astype(np.float64)
            nut_plane, nut_plane_std = nut.mean(axis=(0, 1)), nut.std(axis=(0, 1))
        return {"stats": stats, "first_half": first_half, "samples": n, "wall_s": wall,
                "mlups": lx * ly * lz * steps / wall / 1e6, "nut_plane": nut_plane, "nut_plane_std": nut_plane_std}


def _fine_reference(lx, ly, lz, spinup, window, stats_every, snapshots, seed, perturb, make_trainers, say):
    """The fine run of the closure studies: plain MRT at full resolution (FP32
    storage, rate set 3) from a start that becomes turbulent; statistics and
    their 2x2x2 box-filtered counterparts averaged over the window after the
    spin-up; `snapshots` filtered snapshots (from half the spin-up on) go to
    the trainers make_trainers(input_scales, capacity) returns (a dict)."""
    from . import Trainer, closure_params
    total = spinup + window
    # end of the first half of the statistics window (a multiple of stats_every)
    half_fine = spinup + (window // 2) // stats_every * stats_every
    snap_at = sorted({int(round(s / stats_every)) * stats_every
                      for s in np.linspace(spinup // 2, total, snapshots)})
    Nc = lx * ly * lz // 8
    trainers = {}
    out = {"half_at": half_fine, "Nc": Nc}
    with Channel(lx, ly, lz, "fp32", mrttype=3) as fine:
        _start(fine, perturb, seed)
        nu_c = fine.p["visc"] / 2.0
        s9_c = 1.0 / (3.0 * nu_c + 0.5)
        ystar_c = fine.p["ystar"] / 2.0
        out.update(nu_c=nu_c, s9_c=s9_c, ystar_c=ystar_c)
        fine.stats_reset()
        fine.sync()
        t0 = time.perf_counter()
        t_snap = 0.0
        done = 0
        while done < total:
            fine.run(stats_every)
            done += stats_every
            if done > spinup:
                fine.stats_accumulate()
                fine.stats_accumulate_filtered()
            if done == half_fine:
                out["filtered_half"], _ = fine.stats_read_filtered()
            if done in snap_at:
                fine.sync()
                ts = time.perf_counter()
                if not trainers:
                    # input scales: spread of each feature over the first snapshot
                    with Trainer(Nc, closure_params(nu=nu_c)) as probe:
                        probe.add_filtered(fine, nu_c, s9_c, ystar_c)
                        f, _ = probe.samples()
                    sc = f[:, :9].astype(np.float64).std(axis=0)
                    sc[~(sc > 1e-12)] = 1.0
                    del f
                    trainers.update(make_trainers(sc, len(snap_at) * Nc))
                    out["input_scales"] = [float(v) for v in sc]
                for tr in trainers.values():
                    tr.add_filtered(fine, nu_c, s9_c, ystar_c)
                t_snap += time.perf_counter() - ts
            if done % max(total // 10, 1) == 0:
                st = fine.statistc2()
                say(f"fine: step {done}/{total} U+c {st[lx // 2, 3]:.3f} "
                    f"uu(y+~20) {st[int(20 * fine.p['ystar']), 6]:.3f}")
        fine.sync()
        out["wall_s"] = time.perf_counter() - t0 - t_snap
        out["snapshot_s"] = t_snap
        out["stats"], out["samples"] = fine.stats_read()
        out["filtered"], _ = fine.stats_read_filtered()
    return out, trainers


def _fine_report(fr, lx, ly, lz, total):
    fine_stats, fine_filt = fr["stats"], fr["filtered"]
    return {"wall_s": round(fr["wall_s"], 2), "samples": fr["samples"], "snapshot_s": round(fr["snapshot_s"], 2),
            "mlups": round(lx * ly * lz * total / fr["wall_s"] / 1e6, 1),
            # the two walls of the filtered fine statistics against each other
            "wall_to_wall_asymmetry": _asymmetry(fine_filt),
            # the first half of the window against the whole: its sampling drift
            "first_half_vs_window": _stat_errors(fr["filtered_half"], fine_filt),
            # filtered against plane-pair means of the unfiltered statistics:
            # zero for U+ (the filter is linear), the subfilter share for the rest
            "filter_effect": _stat_errors(fine_filt, 0.5 * (fine_stats[0::2] + fine_stats[1::2]))}


def _coarse_report(r, fr, wall_fine):
    fine_stats = fr["stats"]
    return {"wall_s": round(r["wall_s"], 2), "samples": r["samples"], "mlups": round(r["mlups"], 1),
            "errors_vs_filtered_fine": _stat_errors(r["stats"], fr["filtered"]),
            # the same comparison for the first half of both windows
            "first_half_errors_vs_filtered_fine": _stat_errors(r["first_half"], fr["filtered_half"]),
            "errors_vs_fine": _stat_errors(r["stats"], 0.5 * (fine_stats[0::2] + fine_stats[1::2])),
            "wall_to_wall_asymmetry": _asymmetry(r["stats"]),
            "time_to_stats_speedup": round(wall_fine / r["wall_s"], 2)}


def closure_study(lx=192, ly=768, lz=384, spinup=100000, window=100000, stats_every=100, snapshots=8,
                  iterations=4000, batch=65536, lr=2e-3, seed=2, perturb=TURBULENT_START, log=None):
    """Learned closure against plain MRT (a posteriori).

    1. Fine run, plain MRT at full resolution (FP32 storage, rate set 3), from
       a start that becomes turbulent; statistics and their 2x2x2 box-filtered
       counterparts are averaged over the window after the spin-up, and
       `snapshots` filtered snapshots (from half the spin-up on) become
       training samples.
    2. The network is trained on all snapshots but the last by the device's
       own kernels; the last one measures the a-priori gain over nu_t = 0.
    3. Coarse runs on the grid of half the size (nu/2: dx and dt double) from
       the same analytic start, plain and with the trained closure, over the
       same physical time (half the steps); their statistics are compared
       with the filtered fine ones, and their wall time with the fine run's.
    """
    from . import Trainer, closure_params
    say = log or (lambda m: None)
    total = spinup + window
    modes = ("local", "plane")
    report = {"grid": [lx, ly, lz], "coarse_grid": [lx // 2, ly // 2, lz // 2], "spinup": spinup, "window": window,
              "stats_every": stats_every, "perturbation": {k: list(v) if isinstance(v, tuple) else v
                                                          for k, v in perturb.items()}}
    trained = {}
    trainers = {}

    def make(sc, capacity):
        made = {}
        for mode in modes:
            made[mode] = Trainer(capacity, closure_params(seed=1, scales=sc, nu=nu_of[0]))
            made[mode].set_target(mode)
        return made

    # the coarse run halves the fine viscosity (dx and dt double)
    nu_of = [para(lx, mrttype=3)["visc"] / 2.0]
    try:
        fr, trainers = _fine_reference(lx, ly, lz, spinup, window, stats_every, snapshots, seed, perturb, make, say)
        report["input_scales"] = fr["input_scales"]
        report["fine"] = _fine_report(fr, lx, ly, lz, total)
        Nc = fr["Nc"]
        # mean nu_t/nu per coarse plane (samples are ordered x fastest), both walls folded
        plane = lambda v: v.reshape(lz // 2, ly // 2, lx // 2).mean(axis=(0, 1))  # noqa: E731
        half = lambda v: 0.5 * (v[:lx // 4] + v[::-1][:lx // 4])  # noqa: E731
        for mode, tr in trainers.items():
            # training on all snapshots but the last; the last one is held out
            n_train = tr.count() - Nc
            _, t_all = tr.samples()
            held = t_all[n_train:].astype(np.float64)
            del t_all
            mse_zero = float(np.mean(held ** 2))
            mse_init = float(np.mean((tr.predict(n_train, Nc) - held) ** 2))
            t1 = time.perf_counter()
            loss = tr.train(0, n_train, iterations=iterations, batch=min(batch, n_train), lr=lr)
            t_train = time.perf_counter() - t1
            pred = tr.predict(n_train, Nc).astype(np.float64)
            mse_net = float(np.mean((pred - held) ** 2))
            trained[mode] = tr.params()
            report["training_" + mode] = {
                "samples": int(n_train), "held_out": int(Nc), "iterations": iterations,
                "batch": min(batch, n_train), "lr": lr, "last_batch_loss": float(loss),
                "wall_s": round(t_train, 2),
                "held_out_mse": {"zero": mse_zero, "initial": mse_init, "trained": mse_net},
                "held_out_target_mean": float(held.mean()),
                "nu_t_over_nu_by_plane": {"target": [round(float(v), 4) for v in half(plane(held))],
                                          "predicted": [round(float(v), 4) for v in half(plane(pred))]}}
            say(f"training ({mode}): held-out mse zero {mse_zero:.3e} initial {mse_init:.3e} trained {mse_net:.3e}")
    finally:
        for tr in trainers.values():
            tr.close()
    runs = [("coarse_plain", None)] + [("coarse_closure_" + m, trained[m]) for m in modes]
    for name, closure in runs:
        r = _coarse_run(lx // 2, ly // 2, lz // 2, fr["nu_c"], total // 2, spinup // 2, stats_every // 2, perturb,
                        seed, closure=closure, log=say, half_at=fr["half_at"] // 2)
        report[name] = _coarse_report(r, fr, fr["wall_s"])
        report[name + "_profile"] = r["stats"]
    report["fine_filtered_profile"] = fr["filtered"]
    report["fine_profile"] = fr["stats"]
    return report, trained


def closure_balance_study(lx=192, ly=768, lz=384, spinup=100000, window=300000, stats_every=100, snapshots=8,
                          iterations=4000, batch=65536, lr=2e-3, seed=2, perturb=TURBULENT_START, rounds=2,
                          nut_min=-0.5, core_frac=0.05, log=None):
    """Learned closure trained on momentum-balance targets (d3q19/calibrate.py).

    1. The fine run of closure_study (filtered statistics, filtered snapshots
       whose features are the training inputs).
    2. The plain coarse run; its mean profile against the filtered fine one
       gives the per-plane target g_0 = U_plain'/U_ref' - 1 (the closure the
       mean momentum balance asks for at that run's Reynolds stress).
    3. `rounds` times: train on the current per-plane target (all snapshots,
       lbm_trainer_set_plane_targets), run the coarse channel with the fused
       closure (floor nut_min: the closure may lower the viscosity), update
       the target from that run (g_{k+1} = (1 + g_k) U_k'/U_ref' - 1, g_k the
       run's own plane-mean nu_t/nu).
    4. The last network and plain MRT, each on a second coarse realisation
       (another seed of the start perturbation): the held-out a-posteriori
       comparison with the filtered fine statistics.
    """
    from . import Trainer, closure_params
    from . import calibrate
    say = log or (lambda m: None)
    total = spinup + window
    lxc, lyc, lzc = lx // 2, ly // 2, lz // 2
    report = {"grid": [lx, ly, lz], "coarse_grid": [lxc, lyc, lzc], "spinup": spinup, "window": window,
              "stats_every": stats_every, "rounds": rounds, "nut_min": nut_min, "core_frac": core_frac,
              "seed": seed, "held_out_seed": seed + 1,
              "perturbation": {k: list(v) if isinstance(v, tuple) else v for k, v in perturb.items()}}
    nu_c = para(lx, mrttype=3)["visc"] / 2.0
    trainers = {}
    make = lambda sc, cap: {"balance": Trainer(cap, closure_params(seed=1, scales=sc, nu=nu_c,  # noqa: E731
                                                                   nut_min=nut_min))}
    arrays = {}
    coarse = lambda s, closure: _coarse_run(lxc, lyc, lzc, nu_c, total // 2, spinup // 2,  # noqa: E731
                                            stats_every // 2, perturb, s, closure=closure, log=say,
                                            half_at=fr["half_at"] // 2)
    try:
        fr, trainers = _fine_reference(lx, ly, lz, spinup, window, stats_every, snapshots, seed, perturb, make, say)
        tr = trainers["balance"]
        report["input_scales"] = fr["input_scales"]
        report["fine"] = _fine_report(fr, lx, ly, lz, total)
        U_ref = fr["filtered"][:, COLS["U+"]]
        arrays["fine_filtered_profile"] = fr["filtered"]
        arrays["fine_profile"] = fr["stats"]
        r = coarse(seed, None)
        report["coarse_plain"] = _coarse_report(r, fr, fr["wall_s"])
        arrays["coarse_plain_profile"] = r["stats"]
        say(f"plain: {report['coarse_plain']['errors_vs_filtered_fine']}")
        g = calibrate.balance_update(r["stats"][:, COLS["U+"]], U_ref, 0.0, core_frac=core_frac)
        params = None
        for k in range(rounds):
            tr.set_plane_targets(g.astype(np.float32))
            t1 = time.perf_counter()
            loss = tr.train(0, tr.count(), iterations=iterations, batch=min(batch, tr.count()), lr=lr)
            t_train = time.perf_counter() - t1
            pred = tr.predict(tr.count() - fr["Nc"], fr["Nc"]).astype(np.float64)
            pred_plane = pred.reshape(lzc, lyc, lxc).mean(axis=(0, 1))
            params = tr.params()
            r = coarse(seed, params)
            name = f"coarse_closure_round{k}"
            report[name] = _coarse_report(r, fr, fr["wall_s"])
            report[name]["target"] = [round(float(v), 4) for v in g]
            report[name]["trained_plane_mean"] = [round(float(v), 4) for v in pred_plane]
            report[name]["applied_plane_mean"] = [round(float(v), 4) for v in r["nut_plane"]]
            report[name]["training"] = {"last_batch_loss": float(loss), "wall_s": round(t_train, 2)}
            arrays[name + "_profile"] = r["stats"]
            arrays["params_round%d" % k] = params
            say(f"round {k}: {report[name]['errors_vs_filtered_fine']}")
            g = calibrate.balance_update(r["stats"][:, COLS["U+"]], U_ref, r["nut_plane"], core_frac=core_frac)
        report["next_target"] = [round(float(v), 4) for v in g]
    finally:
        for t in trainers.values():
            t.close()
    for name, closure in (("held_out_plain", None), ("held_out_closure", params)):
        r = coarse(seed + 1, closure)
        report[name] = _coarse_report(r, fr, fr["wall_s"])
        arrays[name + "_profile"] = r["stats"]
        say(f"{name}: {report[name]['errors_vs_filtered_fine']}")
    return report, arrays


# The wall-distance closure profile calibrated on the 2x-coarse Re_tau = 180
# channel (96 x 384 x 192, nu = 0.0018) against the filtered statistics of
# the 192 x 768 x 384 run: nu_t/nu = -0.03 below y+ = 10, -0.08 up to
# y+ = 30, -0.15 beyond.  Use: plane_closure_params(wall_profile(
# **WALL_CLOSURE_C5, lxc=96, ystar_c=...), nu_c, ystar_c).
WALL_CLOSURE_C5 = dict(levels=(-0.03, -0.08, -0.15), splits=(10.0, 30.0), width=5.0)


def wall_profile(levels, lxc, ystar_c, splits=(20.0,), width=5.0):
    """A piecewise-constant plane profile of nu_t/nu in wall units: levels[0]
    below y+ = splits[0], levels[k] between splits[k-1] and splits[k], the
    last level beyond the last split; neighbouring levels joined by tanh over
    `width` wall units (coarse plane centres, both walls)."""
    if len(levels) != len(splits) + 1:
        raise ValueError(f"{len(levels)} levels need {len(levels) - 1} splits, not {len(splits)}")
    i = np.arange(lxc)
    yplus = np.minimum(i + 0.5, lxc - i - 0.5) / ystar_c
    g = np.full(lxc, float(levels[0]))
    for k, ys in enumerate(splits):
        g = g + (levels[k + 1] - levels[k]) * 0.5 * (1.0 + np.tanh((yplus - ys) / width))
    return g


def plane_closure_params(g, nu_c, ystar_c, yscale=100.0, nut_min=-0.5, iterations=3000, lr=3e-3, seed=1):
    """Network parameters whose nu_t/nu is the plane profile g (coarse planes,
    x order) as a function of the wall-distance input alone: the nine
    non-equilibrium inputs are scaled out (scales 1e30) and the device's own
    kernels train on one sample per plane.  Returns the parameters and the
    largest deviation of the trained output from g."""
    from . import Trainer, closure_params
    lxc = len(g)
    i = np.arange(lxc)
    X = np.zeros((lxc, 10), dtype=np.float32)
    X[:, 9] = np.minimum(i + 0.5, lxc - i - 0.5) / ystar_c / yscale
    reps = max(1, 4096 // lxc)
    p = closure_params(seed=seed, scales=np.full(9, 1e30), yscale=yscale, nu=nu_c, nut_min=nut_min)
    with Trainer(reps * lxc, p) as tr:
        tr.add_samples(np.tile(X, (reps, 1)), np.tile(np.asarray(g, dtype=np.float32), reps))
        tr.train(0, reps * lxc, iterations=iterations, batch=reps * lxc, lr=lr)
        pred = tr.predict(0, lxc).astype(np.float64)
        out = tr.params()
    return out, float(np.max(np.abs(pred - np.asarray(g))))


def closure_plane_study(reference, lx=192, ly=768, lz=384, spinup=100000, window=300000, stats_every=100,
                        shapes=((0.0, -0.1),), seeds=(2,), plain=False, splits=(20.0,), width=5.0, nut_min=-0.5,
                        perturb=TURBULENT_START, log=None):
    """Coarse runs with wall-distance closures against a fine reference.

    reference: the .npz a closure study wrote (its fine_filtered_profile and
    fine_profile).  Each shape (a, b) is the plane profile wall_profile(a, b)
    of nu_t/nu, learned by the network from the wall-distance input
    (plane_closure_params); each is run on the coarse channel for every seed
    of the start perturbation, over the closure studies' window, and its
    statistics compared with the filtered fine ones.  plain=True adds the
    plain coarse run of every seed.  A shape is a tuple of levels of
    wall_profile for the given splits."""
    say = log or (lambda m: None)
    ref = np.load(reference)
    fine_filt, fine_stats = ref["fine_filtered_profile"], ref["fine_profile"]
    lxc, lyc, lzc = lx // 2, ly // 2, lz // 2
    if fine_filt.shape[0] != lxc:
        raise ValueError(f"reference has {fine_filt.shape[0]} planes, the coarse grid {lxc}")
    pc = para(lxc, mrttype=3, visc=para(lx, mrttype=3)["visc"] / 2.0)
    nu_c, ystar_c = pc["visc"], pc["ystar"]
    total = spinup + window
    report = {"reference": str(reference), "coarse_grid": [lxc, lyc, lzc], "spinup": spinup, "window": window,
              "splits": list(splits), "width": width, "nut_min": nut_min, "runs": []}
    cases = ([("plain", None)] if plain else []) + [("g=" + ",".join(f"{v:+.3f}" for v in sh), sh) for sh in shapes]
    profiles = {}
    for name, sh in cases:
        params, fit = (None, None)
        if sh is not None:
            g = wall_profile(sh, lxc, ystar_c, splits, width)
            params, fit = plane_closure_params(g, nu_c, ystar_c, nut_min=nut_min)
        for sd in seeds:
            r = _coarse_run(lxc, lyc, lzc, nu_c, total // 2, spinup // 2, stats_every // 2, perturb, sd,
                            closure=params, log=None)
            entry = {"case": name, "seed": sd, "fit": fit, "wall_s": round(r["wall_s"], 2),
                     "mlups": round(r["mlups"], 1),
                     "errors_vs_filtered_fine": _stat_errors(r["stats"], fine_filt),
                     "errors_vs_fine": _stat_errors(r["stats"], 0.5 * (fine_stats[0::2] + fine_stats[1::2])),
                     "wall_to_wall_asymmetry": _asymmetry(r["stats"])}
            if r["nut_plane"] is not None:
                entry["applied_plane_mean"] = [round(float(v), 4) for v in r["nut_plane"][:lxc // 2]]
            report["runs"].append(entry)
            profiles[f"{name} seed={sd}"] = r["stats"]
            say(f"{name} seed {sd}: " + " ".join(f"{k} {v * 100:.2f}%" for k, v in
                                                    entry["errors_vs_filtered_fine"].items()))
    return report, profiles


def closure_local_study(reference, lx=192, ly=768, lz=384, spinup=100000, window=300000, stats_every=100,
                        levels=WALL_CLOSURE_C5["levels"], splits=WALL_CLOSURE_C5["splits"],
                        width=WALL_CLOSURE_C5["width"], train_seed=2, seeds=(3, 4), snapshots=4, stride=2,
                        iterations=4000, batch=65536, lr=2e-3, nut_min=-0.5, perturb=TURBULENT_START, log=None):
    """The calibrated plane profile learned from all ten network inputs.

    1. A coarse run (start seed train_seed) with the wall-distance network of
       the profile; at `snapshots` times in its statistics window the nine
       non-equilibrium inputs and the wall distance of every node of every
       stride-th row (y and z) are read through lbm_closure_inputs (a probe
       closure with unit scales, set for the read only).
    2. The network is trained by the device's kernels on those samples, each
       with the profile's value at its plane as the target and the inputs
       scaled by their spread, so it may use the non-equilibrium moments as
       well as the wall distance.
    3. Coarse runs on `seeds` with that network, against the filtered fine
       statistics of `reference`; the spread of nu_t/nu within each plane
       shows how far the network's output varies from node to node.
    """
    from . import Trainer, closure_params
    say = log or (lambda m: None)
    ref = np.load(reference)
    fine_filt, fine_stats = ref["fine_filtered_profile"], ref["fine_profile"]
    lxc, lyc, lzc = lx // 2, ly // 2, lz // 2
    pc = para(lxc, mrttype=3, visc=para(lx, mrttype=3)["visc"] / 2.0)
    nu_c, ystar_c = pc["visc"], pc["ystar"]
    steps, stats_from, every = (spinup + window) // 2, spinup // 2, stats_every // 2
    g = wall_profile(levels, lxc, ystar_c, splits, width)
    params_y, fit_y = plane_closure_params(g, nu_c, ystar_c, nut_min=nut_min)
    probe = closure_params(init="zero", nu=nu_c, nut_min=nut_min)
    probe[1440] = np.log(np.expm1(-nut_min))  # nu_t = 0: only its inputs are read
    snap_at = {int(round(v / every)) * every for v in np.linspace(stats_from + (steps - stats_from) / snapshots,
                                                                     steps, snapshots)}
    feats = []

    def grab(ch):
        ch.macrovar()
        ch.closure_set(probe, ystar=ystar_c)
        feats.append(np.ascontiguousarray(ch.closure_inputs()[::stride, ::stride]).reshape(-1, 10))
        ch.closure_set(params_y, ystar=ystar_c)

    r = _coarse_run(lxc, lyc, lzc, nu_c, steps, stats_from, every, perturb, train_seed, closure=params_y,
                    snap_at=snap_at, on_snap=grab)
    report = {"reference": str(reference), "levels": list(levels), "splits": list(splits), "width": width,
              "train_seed": train_seed, "snapshots": len(feats), "stride": stride,
              "wall_distance_fit": fit_y,
              "train_run": {"errors_vs_filtered_fine": _stat_errors(r["stats"], fine_filt)}}
    X = np.concatenate(feats)
    del feats
    sc = X[:, :9].astype(np.float64).std(axis=0)
    sc[~(sc > 1e-12)] = 1.0
    X[:, :9] /= sc.astype(np.float32)
    plane = np.arange(X.shape[0]) % lxc  # rows keep every plane, x fastest
    t = g[plane].astype(np.float32)
    n = X.shape[0]
    with Trainer(n, closure_params(seed=1, scales=sc, nu=nu_c, nut_min=nut_min)) as tr:
        tr.add_samples(X, t)
        loss = tr.train(0, n, iterations=iterations, batch=min(batch, n), lr=lr)
        pred = tr.predict(0, n).astype(np.float64)
        params_local = tr.params()
    pm = np.bincount(plane, pred, lxc) / np.bincount(plane, None, lxc)
    ps = np.sqrt(np.maximum(np.bincount(plane, pred * pred, lxc) / np.bincount(plane, None, lxc) - pm * pm, 0.0))
    report["training"] = {"samples": int(n), "iterations": iterations, "last_batch_loss": float(loss),
                          "input_scales": [float(v) for v in sc],
                          "plane_mean_max_error": float(np.max(np.abs(pm - g))),
                          "plane_std_max": float(ps.max()), "rms_error": float(np.sqrt(np.mean((pred - t) ** 2)))}
    say(f"local network: plane-mean error {np.max(np.abs(pm - g)):.4f}, in-plane spread up to {ps.max():.4f}")
    del X, t, pred
    report["runs"] = []
    profiles = {}
    for sd in seeds:
        r = _coarse_run(lxc, lyc, lzc, nu_c, steps, stats_from, every, perturb, sd, closure=params_local)
        e = {"seed": sd, "wall_s": round(r["wall_s"], 2), "mlups": round(r["mlups"], 1),
             "errors_vs_filtered_fine": _stat_errors(r["stats"], fine_filt),
             "errors_vs_fine": _stat_errors(r["stats"], 0.5 * (fine_stats[0::2] + fine_stats[1::2])),
             "wall_to_wall_asymmetry": _asymmetry(r["stats"]),
             "applied_plane_mean": [round(float(v), 4) for v in r["nut_plane"][:lxc // 2]],
             "applied_plane_std": [round(float(v), 4) for v in r["nut_plane_std"][:lxc // 2]]}
        report["runs"].append(e)
        profiles[f"local_seed={sd}"] = r["stats"]
        say(f"local closure seed {sd}: " + " ".join(f"{k} {v * 100:.2f}%" for k, v in
                                                   e["errors_vs_filtered_fine"].items()))
    return report, profiles, params_local


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("study", choices=["precision", "closure", "closure-balance", "closure-plane", "closure-local"])
    ap.add_argument("--reference", default=None, help="closure-plane: .npz of a closure study (fine profiles)")
    ap.add_argument("--shapes", default="0,-0.1",
                    help="closure-plane: ';'-separated nu_t/nu levels between the --splits")
    ap.add_argument("--splits", default="20", help="closure-plane: comma list of y+ between the levels")
    ap.add_argument("--seeds", default="2", help="closure-plane: comma list of start-perturbation seeds")
    ap.add_argument("--plain", action="store_true", help="closure-plane: also the plain coarse run per seed")
    ap.add_argument("--rounds", type=int, default=2, help="closure-balance: target updates from coarse runs")
    ap.add_argument("--nut-min", type=float, default=-0.5, help="closure-balance: floor of nu_t/nu")
    ap.add_argument("--spinup", type=int, default=100000, help="closure: fine steps before the statistics window")
    ap.add_argument("--window", type=int, default=100000, help="closure: fine steps of the statistics window")
    ap.add_argument("--snapshots", type=int, default=8, help="closure: filtered fine snapshots for training")
    ap.add_argument("--iterations", type=int, default=4000, help="closure: Adam iterations")
    ap.add_argument("--batch", type=int, default=65536, help="closure: samples per iteration")
    ap.add_argument("--lr", type=float, default=2e-3, help="closure: Adam learning rate")
    ap.add_argument("--grid", default="192,768,384")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--stats-from", type=int, default=10000)
    ap.add_argument("--stats-every", type=int, default=100)
    ap.add_argument("--mrttype", type=int, default=3, choices=[1, 2, 3])
    ap.add_argument("--rstar", type=float, default=None, help="precision: Re_tau (default 180, para.f90:62)")
    ap.add_argument("--visc", type=float, default=None, help="precision: nu (default 0.0036, para.f90:61)")
    ap.add_argument("--jitter", type=float, default=1e-7, help="control run's disturbance / U_c (0: no control)")
    ap.add_argument("--no-deterministic", action="store_true", help="skip the unperturbed runs")
    ap.add_argument("--turbulent", action="store_true",
                    help="precision: start from the turbulent-start perturbation (a chaotic flow)")
    ap.add_argument("--precs", default="fp64,fp32,fp16", help="precision: comma list (fp64 first)")
    ap.add_argument("--out", default=None, help="also write the report (JSON) and profiles (.npz) here")
    args = ap.parse_args(argv)
    lx, ly, lz = (int(v) for v in args.grid.split(","))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.study == "closure-local":
        report, profiles, params = closure_local_study(args.reference, lx, ly, lz, args.spinup, args.window,
                                                       args.stats_every,
                                                       seeds=[int(v) for v in args.seeds.split(",")],
                                                       iterations=args.iterations, batch=args.batch, lr=args.lr,
                                                       nut_min=args.nut_min, log=log)
        print(json.dumps(report, indent=1))
        if args.out:
            with open(args.out + ".json", "w") as fh:
                json.dump(report, fh, indent=1)
            np.savez(args.out + ".npz", params_local=params, **profiles)
        return
    if args.study == "closure-plane":
        shapes = [tuple(float(v) for v in sh.split(",")) for sh in args.shapes.split(";") if sh]
        report, profiles = closure_plane_study(args.reference, lx, ly, lz, args.spinup, args.window,
                                               args.stats_every, shapes=shapes,
                                               seeds=[int(v) for v in args.seeds.split(",")], plain=args.plain,
                                               splits=[float(v) for v in args.splits.split(",")],
                                               nut_min=args.nut_min, log=log)
        print(json.dumps(report, indent=1))
        if args.out:
            with open(args.out + ".json", "w") as fh:
                json.dump(report, fh, indent=1)
            np.savez(args.out + ".npz", **{k.replace(" ", "_"): v for k, v in profiles.items()})
        return
    if args.study == "closure-balance":
        report, arrays = closure_balance_study(lx, ly, lz, args.spinup, args.window, args.stats_every,
                                               args.snapshots, args.iterations, args.batch, args.lr,
                                               rounds=args.rounds, nut_min=args.nut_min, log=log)
        print(json.dumps(report, indent=1))
        if args.out:
            with open(args.out + ".json", "w") as fh:
                json.dump(report, fh, indent=1)
            np.savez(args.out + ".npz", **arrays)
        return
    if args.study == "closure":
        report, trained = closure_study(lx, ly, lz, args.spinup, args.window, args.stats_every, args.snapshots,
                                        args.iterations, args.batch, args.lr, log=log)
        arrays = {k: report.pop(k) for k in [k for k in report if k.endswith("_profile")]}
        print(json.dumps(report, indent=1))
        if args.out:
            with open(args.out + ".json", "w") as fh:
                json.dump(report, fh, indent=1)
            np.savez(args.out + ".npz", **{"params_" + m: v for m, v in trained.items()}, **arrays)
        return
    report, res = precision_study(lx, ly, lz, args.steps, args.stats_from, args.stats_every,
                                  precs=tuple(args.precs.split(",")), mrttype=args.mrttype, jitter=args.jitter,
                                  rstar=args.rstar, visc=args.visc, deterministic=not args.no_deterministic,
                                  perturb=TURBULENT_START if args.turbulent else None, log=log)
    print(json.dumps(report, indent=1))
    if args.out:
        with open(args.out + ".json", "w") as fh:
            json.dump(report, fh, indent=1)
        np.savez(args.out + ".npz", **{k: v["profile"] for k, v in res.items()})


if __name__ == "__main__":
    main()
