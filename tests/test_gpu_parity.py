"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 results must be bitwise equal to the oracle (populations compared as
uint64 bit patterns): the kernels evaluate every expression in the
reference's order without FMA contraction, and streaming moves whole values.
"""
import os

import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import oracle as O

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
FIELDS = ("f", "rho", "ux", "uy", "uz", "fx", "fy", "fz")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def assert_bitwise(got, want, what):
    if not np.array_equal(bits(got), bits(want)):
        d = np.abs(got - want)
        raise AssertionError(f"{what}: {int(np.count_nonzero(bits(got) != bits(want)))} values differ, "
                             f"max |diff| = {float(np.nanmax(d)):.3e}")


def seeded_pair(lx, ly, lz, seed, laminar=False, noise=0.0, prec="fp64", vel=None, force=None):
    """Oracle and device channel holding the same injected f (loadinitflow path).
    force=None uses FORCING (force_in_y of para.f90 for this lx); a value fixes F_y."""
    ref = O.OracleChannel(lx, ly, lz, laminar=laminar)
    ref.FORCING()
    if vel is None:
        vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=seed)
    ref.ux[:], ref.uy[:], ref.uz[:] = vel
    ref.initpop()
    if noise:
        ref.f += noise * (2 * workloads.uniform(seed + 1000, ref.f.size).reshape(ref.f.shape) - 1)
    f0 = ref.f.copy()
    ch = d3q19.Channel(lx, ly, lz, prec, laminar=laminar)
    ch.upload_f(f0)
    ch.FORCING()
    ref.FORCING()
    if force is not None:
        ref.fy[:] = force
        ch.set_force(0.0, force, 0.0)
    return ref, ch, f0


def test_smoke_entry():
    import __graft_entry__
    __graft_entry__.smoke()


def test_upload_download_roundtrip_is_exact():
    lx, ly, lz = 7, 5, 3
    f = workloads.uniform(3, 19 * lx * ly * lz).reshape(lz, ly, lx, 19) - 0.5
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.upload_f(f)
        assert_bitwise(ch.download_f(), f, "roundtrip")


def test_config1_fp64_1000_steps_bitwise():
    """Channel at 64 streamwise x 32 spanwise x 32 wall-normal (lx=32, ly=64, lz=32),
    Re_tau = 180 constants, laminar parabola + 5% seeded noise, 1000 steps."""
    lx, ly, lz = 32, 64, 32
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=1)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run(1000, ab=True, nthreads=NTHREADS)
        ch.run(1000)
        assert_bitwise(ch.download_f(), ref.f, "f after 1000 steps")
        rho, ux, uy, uz = ch.macrovar(download=True)
        for name, a in (("rho", rho), ("ux", ux), ("uy", uy), ("uz", uz)):
            assert_bitwise(a, getattr(ref, name), name)


@pytest.mark.parametrize("shape", [(2, 1, 1), (3, 1, 4), (5, 7, 3), (64, 3, 2), (33, 17, 9), (130, 4, 3)])
def test_edge_shapes_fp64_bitwise(shape):
    """Ragged and degenerate blocks (one-node periodic directions, two-node
    wall-to-wall channels).  The force is fixed at the 32-node channel's value:
    force_in_y of para.f90 grows as 1/lx^3 and would be unstable at lx = 3."""
    lx, ly, lz = shape
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=5, noise=2e-4, force=1.0252e-4)
    assert np.isfinite(ref.f).all()
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run(40)
        ch.run(40)
        assert np.isfinite(ref.f).all()
        assert_bitwise(ch.download_f(), ref.f, f"f {shape}")


def test_collision_without_macrovar_uses_stale_fields():
    """collision_MRT twice with no macrovar in between reads the older fields, as
    the reference's module arrays do."""
    lx, ly, lz = 8, 6, 5
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=6, noise=1e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        for r in (ref, ch):
            r.collision_MRT()
            r.collision_MRT()
            r.macrovar()
            r.collision_MRT()
        assert_bitwise(ch.download_f(), ref.f, "f")


def test_prerelaxation_path_bitwise():
    """Fresh start of main.f90:52-102: initvel, FORCING, initpop, pre-relaxation
    with u held at the initial field, then the main loop."""
    lx, ly, lz = 16, 12, 8
    ref = O.OracleChannel(lx, ly, lz)
    ch = d3q19.Channel(lx, ly, lz)
    with ch:
        ux, uy, uz = workloads.laminar_parabola_noise(lx, ly, lz, seed=9)
        ref.ux[:], ref.uy[:], ref.uz[:] = ux, uy, uz
        ch.set_macros(ux=ux, uy=uy, uz=uz)
        ref.FORCING()
        ch.FORCING()
        ref.initpop()
        ch.initpop()
        errs_ref, errs = [], []
        for _ in range(5):  # fixed iteration count: both paths must track each other
            rhop = ref.rho.copy()
            ref.rhoupdat()
            ref.collision_MRT()
            errs_ref.append(float(np.max(np.abs(ref.rho - rhop))))
            errs.append(ch.rhoupdat())
            ch.collision_MRT()
        assert errs == errs_ref
        ref.macrovar()
        ch.macrovar()
        ref.run(20)
        ch.run(20)
        assert_bitwise(ch.download_f(), ref.f, "f")


def test_initvel_matches_reference_profile():
    lx, ly, lz = 32, 4, 4
    ref = O.OracleChannel(lx, ly, lz)
    ref.initvel()
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.initvel()
        ch.FORCING()
        ch.initpop()
        ref.FORCING()
        ref.initpop()
        assert_bitwise(ch.download_f(), ref.f, "initpop(initvel)")


def test_avedensity_fused_and_array_paths():
    lx, ly, lz = 10, 6, 4
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=12, noise=3e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run(3)
        ch.run(3)
        m_ref = ref.avedensity()
        m = ch.avedensity()  # fused (offset) path
        assert m == pytest.approx(m_ref, rel=1e-12, abs=1e-18)
        ref.collision_MRT()
        ch.collision_MRT()
        ref.macrovar()
        ch.macrovar()
        # the device sums in a different order than the sequential reference
        # sum, so the corrected density (and everything after) agrees to
        # round-off of the mean only
        f = ch.download_f()
        assert np.max(np.abs(f - ref.f)) < 1e-17 + 1e-12 * abs(m_ref)
        # array path: rhoupdat materialises the fields first
        ch2 = d3q19.Channel(lx, ly, lz)
        with ch2:
            ch2.upload_f(ref.f)
            ch2.FORCING()
            ch2.macrovar()
            ch2.collision_MRT()
            ch2.macrovar(download=True)
            ch2.rhoupdat()
            mm = ch2.avedensity()
            assert np.isfinite(mm)


def test_set_force_after_macrovar_keeps_old_fields():
    lx, ly, lz = 6, 5, 4
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=13, noise=1e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.fy[:] = 3e-4
        ch.set_force(0.0, 3e-4, 0.0)
        for r in (ref, ch):
            r.collision_MRT()
            r.macrovar()
            r.collision_MRT()
        assert_bitwise(ch.download_f(), ref.f, "f")


def test_mass_sum_and_statistics_match_oracle():
    lx, ly, lz = 24, 16, 12
    ref, ch, _ = seeded_pair(lx, ly, lz, seed=14, noise=1e-4)
    with ch:
        ref.macrovar()
        ch.macrovar()
        ref.run(30)
        ch.run(30)
        assert ch.mass() == pytest.approx(ref.f.sum(), rel=1e-13, abs=1e-15)
        s_ref = ref.statistc2()
        s = ch.statistc2()  # from the fused (virtual) fields
        assert np.allclose(s, s_ref, rtol=1e-10, atol=1e-12)
        ch.macrovar(download=True)
        s2 = ch.statistc2()
        assert np.allclose(s2, s_ref, rtol=1e-10, atol=1e-12)
        # time average of two samples
        ch.stats_reset()
        ch.stats_accumulate()
        acc = [(ref.ux.copy(), ref.uy.copy(), ref.uz.copy(), ref.rho.copy())]
        ch.run(5)
        ref.run(5)
        ch.stats_accumulate()
        acc.append((ref.ux.copy(), ref.uy.copy(), ref.uz.copy(), ref.rho.copy()))
        avg, n = ch.stats_read()
        assert n == 2
        ubar = sum(a[1] for a in acc).sum(axis=(0, 1)) / (2 * ly * lz) / ref.p.ustar
        assert np.allclose(avg[:, 3], ubar, rtol=1e-12)
        uu = sum(a[1] ** 2 for a in acc).sum(axis=(0, 1)) / (2 * ly * lz) / ref.p.ustar ** 2 - ubar ** 2
        assert np.allclose(avg[:, 6], uu, rtol=1e-8, atol=1e-12)


def test_two_runs_are_bitwise_identical():
    lx, ly, lz = 48, 40, 24
    outs = []
    for _ in range(2):
        ref, ch, _ = seeded_pair(lx, ly, lz, seed=15)
        with ch:
            ch.macrovar()
            ch.run(50)
            outs.append(ch.download_f())
    assert_bitwise(outs[0], outs[1], "determinism")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(12, 10, 6), (48, 40, 24), (64, 3, 2)])
def test_vector_run_kernel_equals_one_node_kernel(prec, shape, monkeypatch):
    """The step kernel with two nodes per thread (even lx) against the
    one-node kernel, including an avedensity (mode 1) step: bit for bit in
    FP64 (no FMA contraction); in FP32, where the compiler may contract the
    same expressions differently in the two kernels, to FP32 round-off."""
    lx, ly, lz = shape
    vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=21)
    outs = []
    for kernel in ("node", "runs"):
        monkeypatch.setenv("LBM_STEP_KERNEL", kernel)
        with d3q19.Channel(lx, ly, lz, prec, laminar=True) as ch:
            ch.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
            ch.FORCING()
            ch.initpop()
            ch.macrovar()
            ch.run(20)
            ch.avedensity()
            ch.run(20)
            outs.append(ch.download_post())
    if prec == "fp64":
        assert_bitwise(outs[0], outs[1], f"{prec} runs vs node {shape}")
    else:
        scale = np.max(np.abs(outs[0]))
        assert np.max(np.abs(outs[1] - outs[0])) < 1e-6 * scale


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("fp16", 2e-3), ("fp16c", 2e-3)])
def test_reduced_precision_tracks_fp64(prec, tol):
    lx, ly, lz = 32, 64, 32
    vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=1)
    with d3q19.Channel(lx, ly, lz, "fp64") as a, d3q19.Channel(lx, ly, lz, prec) as b:
        for c in (a, b):
            c.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
            c.FORCING()
            c.initpop()
            c.macrovar()
            c.run(200)
        _, _, uy_a, _ = a.macrovar(download=True)
        _, _, uy_b, _ = b.macrovar(download=True)
        rel = np.max(np.abs(uy_b - uy_a)) / np.max(np.abs(uy_a))
        assert rel < tol, rel
        # mean streamwise profile
        pa = uy_a.mean(axis=(0, 1))
        pb = uy_b.mean(axis=(0, 1))
        assert np.max(np.abs(pb - pa)) / np.max(np.abs(pa)) < tol


def test_fp16c_storage_is_closer_to_fp64_than_fp16():
    """The 16-bit format with one more significand bit (LBM_FP16C) tracks the
    FP64 lattice more closely than IEEE FP16 storage of the same scaled
    values, over the same 200 steps."""
    lx, ly, lz = 32, 64, 32
    vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=1)
    err = {}
    with d3q19.Channel(lx, ly, lz, "fp64") as a:
        for c in (a,):
            c.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
            c.FORCING()
            c.initpop()
            c.macrovar()
            c.run(200)
        ref = a.download_post()
    for prec in ("fp16", "fp16c"):
        with d3q19.Channel(lx, ly, lz, prec) as b:
            b.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
            b.FORCING()
            b.initpop()
            b.macrovar()
            b.run(200)
            err[prec] = np.max(np.abs(b.download_post() - ref)) / np.max(np.abs(ref))
    assert err["fp16c"] < 0.7 * err["fp16"], err


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16c"])
def test_host_widening_equals_device_gather(prec):
    """lbm_download_post_range widens the stored values on the host; the
    device gather of lbm_download_post widens them on the GPU: same values."""
    lx, ly, lz = 16, 12, 10
    vel = workloads.laminar_parabola_noise(lx, ly, lz, seed=2, amp=0.3)
    with d3q19.Channel(lx, ly, lz, prec) as c:
        c.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
        c.FORCING()
        c.initpop()
        c.macrovar()
        c.run(7)
        full = c.download_post()
        part = c.download_post_range(lx * ly * 3, lx * ly * 4)
    assert np.array_equal(part, full[:, lx * ly * 3:lx * ly * 7])


def test_rhoupdat_reports_a_nan():
    """A NaN anywhere in the lattice makes max|drho| NaN (the block and wave
    maxima order the bit patterns, where a NaN ranks above +inf), so the
    pre-relaxation loop of Channel-Flow/main.f90:70-90 can never take it for
    convergence."""
    lx, ly, lz = 24, 20, 18
    f = (workloads.uniform(71, 19 * lx * ly * lz) - 0.5).reshape(lz, ly, lx, 19) * 1e-4
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.upload_f(f)
        ch.FORCING()
        assert np.isfinite(ch.rhoupdat())
        assert ch.rhoupdat() == 0.0  # unchanged lattice: no change of rho
        f[11, 7, 5, 3] = np.nan
        ch.upload_f(f)
        assert np.isnan(ch.rhoupdat())


def test_separable_velocity_on_the_device_equals_host_fields():
    """lbm_set_velocity_separable (the log-law + divergence-free start as
    plane x profile terms, summed on the device) gives the fields the host
    evaluation of the same terms gives, bit for bit (through initpop), and
    the same populations after 20 steps."""
    lx, ly, lz = 40, 36, 28
    p = d3q19.para(lx)
    terms = workloads.loglaw_divfree_terms(lx, ly, lz, p["ustar"], p["ystar"], seed=4)
    ux, uy, uz = workloads.separable_fields(terms, lx, ly, lz)
    with d3q19.Channel(lx, ly, lz) as a, d3q19.Channel(lx, ly, lz) as b:
        a.set_velocity_separable(terms)
        b.set_macros(rho=np.zeros_like(ux), ux=ux, uy=uy, uz=uz)
        for c in (a, b):
            c.FORCING()
            c.initpop()  # the equilibrium of the fields: equal populations iff equal fields
        assert_bitwise(a.download_f(), b.download_f(), "initpop of the fields")
        for c in (a, b):
            c.macrovar()
            c.run(20)
        assert_bitwise(a.download_f(), b.download_f(), "f after 20 steps")
