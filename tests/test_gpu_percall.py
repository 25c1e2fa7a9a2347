"""The per-call main loop (collision_MRT then macrovar, once per step, as
Channel-Flow/main.f90:157-161 calls them): lbm_collision_mrt defers a
pairable step so that two calls of each run as one blocked pair.  Every
interleaving of other calls must see exactly the state the calls one by one
leave: compared bit for bit with a handle made with LBM_DEFER=0, and the
plain loop with the oracle."""
import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pair(monkeypatch, lx, ly, lz, prec="fp64", seed=5):
    """A deferring handle and a LBM_DEFER=0 handle on the same start."""
    monkeypatch.delenv("LBM_BLOCK_PLANES", raising=False)
    monkeypatch.delenv("LBM_BLOCK_AXIS", raising=False)
    out = []
    for d in ("1", "0"):
        monkeypatch.setenv("LBM_DEFER", d)
        out.append(d3q19.Channel(lx, ly, lz, prec))
    monkeypatch.delenv("LBM_DEFER", raising=False)
    vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=seed, uc=0.04)
    for c in out:
        c.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
        c.FORCING()
        c.initpop()
        c.macrovar()
    return out


def test_per_call_loop_equals_the_oracle():
    """32x64x32 channel, 40 per-call steps (20 pairs) against the oracle."""
    lx, ly, lz = 32, 64, 32
    ref = O.OracleChannel(lx, ly, lz)
    ref.FORCING()
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=61)
    ref.initpop()
    f0 = ref.f.copy()
    ref.macrovar()
    for _ in range(40):
        ref.collision_MRT()
        ref.macrovar()
    with d3q19.Channel(lx, ly, lz) as ch:
        assert ch.block_info()[1] > 0
        ch.upload_f(f0)
        ch.FORCING()
        ch.macrovar()
        for _ in range(40):
            ch.collision_MRT()
            ch.macrovar()
        assert np.array_equal(bits(ch.download_f()), bits(ref.f))
        for a, b in zip(ch.macrovar(download=True), (ref.rho, ref.ux, ref.uy, ref.uz)):
            assert np.array_equal(bits(a), bits(b))


OPS = ("coll", "coll", "coll", "mac", "mac", "mac", "macdl", "force", "ave", "stat", "mass", "rhoup", "dl", "run",
       "post", "sync")


def _apply(ch, op, k):
    if op == "coll":
        ch.collision_MRT()
    elif op == "mac":
        ch.macrovar()
    elif op == "macdl":
        return np.concatenate([a.ravel() for a in ch.macrovar(download=True)])
    elif op == "force":
        ch.set_force(0.0, ch.p["force_in_y"] * (1.0 + 0.01 * (k % 3)), 0.0)
    elif op == "ave":
        return np.array([ch.avedensity()])
    elif op == "stat":
        return ch.statistc2().ravel()
    elif op == "mass":
        return np.array([ch.mass()])
    elif op == "rhoup":
        return np.array([ch.rhoupdat()])
    elif op == "dl":
        return ch.download_f().ravel()
    elif op == "run":
        ch.run(1 + k % 4)
    elif op == "post":
        return ch.download_post().ravel()
    elif op == "sync":
        ch.sync()
    return None


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_call_sequences_match_undeferred_calls(monkeypatch, seed):
    """Seeded random sequences of main-loop calls (collision, macrovar with and
    without download, FORCING, avedensity, statistc2, mass, rhoupdat,
    downloads, lbm_run, sync), weighted towards the per-call loop: every
    returned value and the final state equal the LBM_DEFER=0 handle's."""
    lx, ly, lz = 24, 20, 18
    rng = np.random.default_rng(seed)
    ops = [OPS[i] for i in rng.integers(0, len(OPS), 120)]
    ops = ["coll", "mac"] * 6 + ops + ["coll", "mac", "coll", "mac", "coll"]
    a, b = _pair(monkeypatch, lx, ly, lz, seed=seed)
    with a, b:
        assert a.block_info()[1] > 0
        for k, op in enumerate(ops):
            ra, rb = _apply(a, op, k), _apply(b, op, k)
            if ra is not None:
                assert np.array_equal(bits(ra), bits(rb)), f"step {k}: {op}"
        assert np.array_equal(bits(a.download_f()), bits(b.download_f()))
        assert np.array_equal(bits(a.download_post()), bits(b.download_post()))


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16c"])
def test_per_call_loop_reduced_precision(monkeypatch, prec):
    """FP32 pairs (same node arithmetic as single steps) and FP16 storage
    (no default sweep: nothing deferred) equal the undeferred calls."""
    a, b = _pair(monkeypatch, 24, 20, 18, prec=prec)
    with a, b:
        for c in (a, b):
            for _ in range(9):
                c.collision_MRT()
                c.macrovar()
        assert np.array_equal(bits(a.download_post()), bits(b.download_post()))
