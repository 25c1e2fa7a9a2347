"""The two-step blocked sweep of lbm_run (csrc/blocked.h) against single
steps, through the C ABI.  The blocked sweep changes only where the state
between two steps lives (a ring of planes instead of the second lattice), so
FP64 results must be bitwise equal to single steps and to the oracle; the
schedule itself is checked on the host by tests/host/block_check.hip."""
import numpy as np
import pytest

import d3q19
from test_gpu_parity import assert_bitwise, bits, seeded_pair

pytestmark = pytest.mark.gpu


def _pair(monkeypatch, lx, ly, lz, prec, planes, seed, axis="z"):
    """Two channels on the same injected start: blocked (slabs per chunk, sweep
    axis) and single steps."""
    ref, _, f0 = seeded_pair(lx, ly, lz, seed=seed, noise=1e-4, prec="fp64")
    chans = []
    monkeypatch.setenv("LBM_BLOCK_AXIS", axis)
    for k in (planes, 0):
        monkeypatch.setenv("LBM_BLOCK_PLANES", str(k))
        ch = d3q19.Channel(lx, ly, lz, prec)
        ch.upload_f(f0)
        ch.FORCING()
        ch.macrovar()
        chans.append(ch)
    monkeypatch.delenv("LBM_BLOCK_PLANES")
    monkeypatch.delenv("LBM_BLOCK_AXIS")
    return ref, chans


@pytest.mark.parametrize("axis", ["z", "y"])
@pytest.mark.parametrize("shape,planes", [
    ((8, 6, 5), 2), ((12, 10, 9), 4), ((16, 12, 17), 16), ((6, 5, 1), 2), ((6, 5, 2), 2),
    ((10, 8, 7), 1), ((9, 7, 12), 5), ((32, 64, 32), 4), ((6, 1, 5), 2), ((8, 41, 45), 18),
    ((6, 70, 66), 32),
])
def test_blocked_sweep_fp64_is_bitwise_single_steps_and_oracle(monkeypatch, shape, planes, axis):
    lx, ly, lz = shape
    print(f"blocked {shape} planes={planes} axis={axis}", flush=True)
    ref, (blk, one) = _pair(monkeypatch, lx, ly, lz, "fp64", planes, seed=51, axis=axis)
    with blk, one:
        ref.macrovar()
        blk.run(21)  # ten blocked pairs and one single step
        one.run(21)
        ref.run(21)
        fb = blk.download_f()
        assert_bitwise(fb, one.download_f(), f"blocked vs single steps {shape}")
        assert_bitwise(fb, ref.f, f"blocked vs oracle {shape}")
        for name, a, b in zip(("rho", "ux", "uy", "uz"), blk.macrovar(download=True), one.macrovar(download=True)):
            assert_bitwise(a, b, name)


@pytest.mark.parametrize("axis", ["z", "y"])
@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16c"])
def test_blocked_sweep_reduced_precision_matches_single_steps(monkeypatch, prec, axis):
    lx, ly, lz = 24, 20, 18
    _, (blk, one) = _pair(monkeypatch, lx, ly, lz, prec, 4, seed=52, axis=axis)
    with blk, one:
        blk.run(40)
        one.run(40)
        a, b = blk.download_post(), one.download_post()
    # same arithmetic; the FP32 unit may contract the two kernels' FMAs differently
    same = np.array_equal(bits(a), bits(b))
    assert same or np.max(np.abs(a - b)) < 1e-6 * np.max(np.abs(b))


def test_blocked_sweep_then_other_main_loop_calls(monkeypatch):
    """After blocked pairs every other call sees the state single steps leave:
    collision without macrovar, avedensity (fused offset), statistics."""
    lx, ly, lz = 16, 12, 10
    ref, (blk, one) = _pair(monkeypatch, lx, ly, lz, "fp64", 3, seed=53, axis="y")
    with blk, one:
        for c in (blk, one):
            c.run(6)
            c.collision_MRT()
            c.collision_MRT()
            c.macrovar()
            c.avedensity()
            c.run(4)
        assert_bitwise(blk.download_f(), one.download_f(), "f")
        np.testing.assert_array_equal(blk.statistc2(), one.statistc2())
        assert blk.mass() == one.mass()


@pytest.mark.parametrize("prec,want", [("fp64", ("y", 18)), ("fp32", ("z", 18)), ("fp16", ("z", 0)),
                                       ("fp16c", ("z", 0))])
def test_default_sweep_on_the_turbulent_channel(monkeypatch, prec, want):
    """lbm_create's choice for the 192x768x384 channel: 11.2 MB slabs, 18 to
    a chunk (a 224 MB ring), along y in FP64 and z in FP32; 2-byte storage
    runs single steps.  Only the handle is made (lattices allocated and
    cleared; no step runs)."""
    monkeypatch.delenv("LBM_BLOCK_PLANES", raising=False)
    monkeypatch.delenv("LBM_BLOCK_AXIS", raising=False)
    print(f"default sweep {prec}", flush=True)
    with d3q19.Channel(192, 768, 384, prec) as ch:
        assert ch.block_info() == want
    # and the sweep stays off for the 2048x1024x512 channel's 80 MB / 40 MB slabs in FP32
    if prec == "fp32":
        with d3q19.Channel(512, 2048, 1024, prec) as ch:
            assert ch.block_info()[1] == 0


@pytest.mark.parametrize("prec,want", [("fp64", ("y", 18)), ("fp32", ("z", 18))])
def test_default_sweep_bitwise_at_full_size(monkeypatch, prec, want):
    """The defaults behind the bench numbers (FP64: y slabs, FP32: z slabs,
    18 to a chunk) on the full 192x768x384 channel from the bench's start
    (log-law + divergence-free perturbation): two blocked pairs equal four
    single steps bit for bit, populations and the FP64 mass.  (The FP32 unit
    allows FMA contraction; the sweep and the single step inline the same
    node arithmetic and contract it the same way.)"""
    from d3q19 import workloads
    lx, ly, lz = 192, 768, 384
    monkeypatch.delenv("LBM_BLOCK_AXIS", raising=False)
    print(f"full-size blocked vs single steps, {prec}", flush=True)
    chans = []
    for planes in (None, "0"):
        if planes is None:
            monkeypatch.delenv("LBM_BLOCK_PLANES", raising=False)
        else:
            monkeypatch.setenv("LBM_BLOCK_PLANES", planes)
        chans.append(d3q19.Channel(lx, ly, lz, prec))
    monkeypatch.delenv("LBM_BLOCK_PLANES", raising=False)
    blk, one = chans
    with blk, one:
        assert blk.block_info() == want and one.block_info()[1] == 0
        ux, uy, uz = workloads.loglaw_divfree(lx, ly, lz, blk.p["ustar"], blk.p["ystar"], seed=2)
        for c in (blk, one):
            c.set_macros(ux=ux, uy=uy, uz=uz)
            c.FORCING()
            c.initpop()
            c.macrovar()
        del ux, uy, uz
        blk.run(4)
        one.run(4)
        assert blk.mass() == one.mass()
        a = blk.download_post()
        b = one.download_post()
        assert np.array_equal(bits(a), bits(b)), "blocked sweep differs from single steps at full size"
