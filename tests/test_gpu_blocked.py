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
@pytest.mark.parametrize("shape,planes,rows", [
    ((8, 9, 11), 3, 2), ((6, 10, 7), 2, 3), ((5, 12, 13), 4, 1), ((7, 11, 9), 5, 4), ((6, 13, 12), 32, 5),
    ((4, 5, 6), 1, 4),
])
def test_row_windows_fp64_bitwise_single_steps_and_oracle(monkeypatch, shape, planes, rows, axis):
    """Row windows (LBM_BLOCK_ROWS): each window of `rows` rows is swept on
    its own, its two edge rows computed twice; the last window may be
    narrower.  Populations and fields bitwise those of single steps and the
    oracle over ten pairs and a single step."""
    lx, ly, lz = shape
    print(f"row windows {shape} planes={planes} rows={rows} axis={axis}", flush=True)
    monkeypatch.setenv("LBM_BLOCK_ROWS", str(rows))
    ref, (blk, one) = _pair(monkeypatch, lx, ly, lz, "fp64", planes, seed=54, axis=axis)
    monkeypatch.delenv("LBM_BLOCK_ROWS")
    with blk, one:
        ns = lz if axis == "z" else ly
        assert blk.block_info() == (axis, min(planes, ns)) and blk.block_rows() == rows
        assert one.block_info()[1] == 0 and one.block_rows() == 0
        ref.macrovar()
        for c in (blk, one, ref):
            c.run(21)
        fb = blk.download_f()
        assert_bitwise(fb, one.download_f(), f"row windows vs single steps {shape}")
        assert_bitwise(fb, ref.f, f"row windows vs oracle {shape}")
        for name, a, b in zip(("rho", "ux", "uy", "uz"), blk.macrovar(download=True), one.macrovar(download=True)):
            assert_bitwise(a, b, name)


@pytest.mark.parametrize("axis", ["z", "y"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_row_windows_with_forces_per_node(monkeypatch, prec, axis):
    """Row windows under FORCINGP's force per node (LBM_BLOCK_FORCED=1), which
    each node takes from its lattice row: the batched loop against FORCINGP,
    collision_MRT and macrovar called one step at a time."""
    lx, ly, lz = 10, 14, 12
    print(f"row windows forcingp {prec} axis={axis}", flush=True)
    _, _, f0 = seeded_pair(lx, ly, lz, seed=36, noise=1e-4)
    for k, v in (("LBM_BLOCK_AXIS", axis), ("LBM_BLOCK_PLANES", "3"), ("LBM_BLOCK_ROWS", "5"),
                 ("LBM_BLOCK_FORCED", "1")):
        monkeypatch.setenv(k, v)
    a = d3q19.Channel(lx, ly, lz, prec)
    for k in ("LBM_BLOCK_PLANES", "LBM_BLOCK_AXIS", "LBM_BLOCK_ROWS", "LBM_BLOCK_FORCED"):
        monkeypatch.delenv(k)
    b = d3q19.Channel(lx, ly, lz, prec)
    with a, b:
        assert a.block_info() == (axis, 3) and a.block_rows() == 5
        for c in (a, b):
            c.upload_f(f0)
            c.FORCING()
            c.macrovar()
        a.run_forcingp(20, 25, 35)
        b.run_forcingp_stepwise(20, 25, 35)
        pa, pb = a.download_post(), b.download_post()
        if prec == "fp64":
            assert_bitwise(pa, pb, "row windows: batched vs call-by-call FORCINGP loop")
        else:
            # the FP32 unit may contract the two kernels' FMAs differently (an
            # ulp, 6e-8 relative); a wrong row in the force moves values by O(F)
            assert np.max(np.abs(pa - pb)) <= 1e-6 * np.max(np.abs(pb))


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
    for k in ("LBM_BLOCK_PLANES", "LBM_BLOCK_AXIS", "LBM_BLOCK_ROWS"):
        monkeypatch.delenv(k, raising=False)
    print(f"default sweep {prec}", flush=True)
    with d3q19.Channel(192, 768, 384, prec) as ch:
        assert ch.block_info() == want
        assert ch.block_rows() == 0
    # the 2048x1024x512 channel's 80 MB / 40 MB slabs in FP32: y slabs in four
    # row windows of 256 rows (10 MB window slabs, 20 to a chunk)
    if prec == "fp32":
        with d3q19.Channel(512, 2048, 1024, prec) as ch:
            assert ch.block_info() == ("y", 20) and ch.block_rows() == 256


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
