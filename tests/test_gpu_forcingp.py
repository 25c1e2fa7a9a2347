"""GPU parity of the perturbation forcing FORCINGP (Channel-Flow/collision.f90:
529-602) against the CPU oracle, through the C ABI.

The device evaluates the force per node inside the step from tables of one
coordinate each; the oracle fills the reference's three force arrays.  FP64
populations and fields must agree bit for bit, including the reference's
timing: the collision reads fields that the previous macrovar formed with the
previous step's force, while its forcing term uses the current one.
"""
import numpy as np
import pytest

import d3q19
import oracle as O
from test_gpu_parity import assert_bitwise, seeded_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,istep0,nsteps,npforcing", [
    ((24, 16, 12), 0, 50, 40),       # switch-off to FORCING at step 40
    ((10, 9, 7), 470, 20, 10**9),     # amplitude near its maximum, odd sizes
    ((7, 5, 4), 100, 12, 106),        # bands of one node (lx/4 = 1)
])
def test_forcingp_main_loop_fp64_bitwise(shape, istep0, nsteps, npforcing):
    lx, ly, lz = shape
    print(f"forcingp main loop {shape} istep0={istep0}", flush=True)
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=31, noise=1e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run_forcingp(istep0, nsteps, npforcing)
        ch.run_forcingp(istep0, nsteps, npforcing)
        assert np.isfinite(ref.f).all()
        assert_bitwise(ch.download_f(), ref.f, f"f {shape}")
        fields = ch.macrovar(download=True)
        ref.macrovar()
        for name, got, want in zip(("rho", "ux", "uy", "uz"), fields, (ref.rho, ref.ux, ref.uy, ref.uz)):
            assert_bitwise(got, want, name)


def test_forcingp_state_transitions_fp64_bitwise():
    """Force changes between macrovar and collision in every field state:
    fused fields formed with another force, stale fields after a collision
    without macrovar, materialised fields (rhoupdat), back to FORCING."""
    lx, ly, lz = 12, 10, 8
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=32, noise=1e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        for r in (ref, ch):
            r.FORCINGP(300)
            r.collision_MRT()      # fields formed with FORCING, forcing term FORCINGP(300)
            r.collision_MRT()      # stale fields of the step before
        fields = ch.macrovar(download=True)
        ref.macrovar()
        for name, got, want in zip(("rho", "ux", "uy", "uz"), fields, (ref.rho, ref.ux, ref.uy, ref.uz)):
            assert_bitwise(got, want, name)
        for r in (ref, ch):
            r.FORCINGP(301)
            r.collision_MRT()
            r.macrovar()
            r.FORCINGP(302)
            r.rhoupdat()           # fields become arrays
            r.collision_MRT()
            r.macrovar()
            r.FORCING()            # fields formed with FORCINGP(302), forcing term uniform
            r.collision_MRT()
            r.macrovar()
            r.collision_MRT()
            r.macrovar()
        assert_bitwise(ch.download_f(), ref.f, "f")


def test_forcingp_statistics_and_avedensity():
    lx, ly, lz = 16, 12, 8
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=33, noise=3e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run_forcingp(200, 6, 10**9)
        ch.run_forcingp(200, 6, 10**9)
        # statistics of fields the device forms from F and the per-node force
        got = ch.statistc2()
        want = ref.statistc2()
        assert np.allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
        # avedensity (fused offset path), then a step with the offset
        m_ref = ref.avedensity()
        m = ch.avedensity()
        assert m == pytest.approx(m_ref, rel=1e-12, abs=1e-18)
        for r in (ref, ch):
            r.FORCINGP(207)
            r.collision_MRT()
            r.macrovar()
        f = ch.download_f()
        assert np.max(np.abs(f - ref.f)) < 1e-17 + 1e-12 * abs(m_ref)


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("fp16", 2e-3), ("fp16c", 2e-3)])
def test_forcingp_reduced_precision_tracks_fp64(prec, tol):
    lx, ly, lz = 32, 48, 24
    ref, _, _ = seeded_pair(lx, ly, lz, seed=34, noise=0.0, prec="fp64")
    f0 = ref.f.copy()
    outs = []
    for p in ("fp64", prec):
        with d3q19.Channel(lx, ly, lz, p) as c:
            c.upload_f(f0)
            c.FORCING()
            c.macrovar()
            c.run_forcingp(450, 100, 10**9)
            outs.append(c.macrovar(download=True)[2])
    uy_a, uy_b = outs
    rel = np.max(np.abs(uy_b - uy_a)) / np.max(np.abs(uy_a))
    assert rel < tol, rel


def test_forcingp_not_combined_with_particles():
    lx, ly, lz = 32, 24, 24
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.FORCING()
        ch.particles_create(np.array([[16.0, 12.0, 12.0]]), rad=4.0)
        with pytest.raises(d3q19.LbmError, match="particle"):
            ch.FORCINGP(10)
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.FORCINGP(10)
        with pytest.raises(d3q19.LbmError, match="uniform force"):
            ch.particles_create(np.array([[16.0, 12.0, 12.0]]), rad=4.0)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "fp16c"])
@pytest.mark.parametrize("axis,forced", [("z", True), ("y", True), ("y", False)])
def test_batched_forcingp_loop_equals_call_by_call(monkeypatch, prec, axis, forced):
    """lbm_run_forcingp against FORCINGP / FORCING, collision_MRT and macrovar
    called one step at a time: bit for bit in FP64; the same arithmetic in
    FP32/FP16 storage, whose FMA contraction may differ between kernels.  With
    LBM_BLOCK_FORCED=1 the FORCINGP steps run as blocked pairs with the force
    evaluated per node; by default they run one by one and the steps after
    npforcing in pairs.  The range starts on an odd step, switches to FORCING
    inside a pair (npforcing = 35) and ends on a single step."""
    lx, ly, lz = 20, 18, 14
    print(f"batched forcingp {prec} axis={axis} forced={forced}", flush=True)
    _, _, f0 = seeded_pair(lx, ly, lz, seed=35, noise=1e-4)
    monkeypatch.setenv("LBM_BLOCK_AXIS", axis)
    monkeypatch.setenv("LBM_BLOCK_PLANES", "4")
    if forced:
        monkeypatch.setenv("LBM_BLOCK_FORCED", "1")
    chans = [d3q19.Channel(lx, ly, lz, prec) for _ in range(2)]
    for k in ("LBM_BLOCK_PLANES", "LBM_BLOCK_AXIS", "LBM_BLOCK_FORCED"):
        monkeypatch.delenv(k, raising=False)
    a, b = chans
    with a, b:
        assert a.block_info() == (axis, 4)
        for c in (a, b):
            c.upload_f(f0)
            c.FORCING()
            c.macrovar()
        a.run_forcingp(20, 25, 35)
        b.run_forcingp_stepwise(20, 25, 35)
        pa, pb = a.download_post(), b.download_post()
        if prec == "fp64":
            assert_bitwise(pa, pb, "batched vs call-by-call FORCINGP loop")
            assert a.mass() == b.mass()
        else:
            # a contraction difference moves a stored value by an ulp: 6e-8
            # relative in FP32, 5e-4 in FP16; a wrong gather moves it by O(1)
            tol = 1e-6 if prec == "fp32" else 2e-3
            assert np.max(np.abs(pa - pb)) <= tol * np.max(np.abs(pb))
