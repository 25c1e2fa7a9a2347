"""Device diagnostics against the oracle's restatements on the same inputs:
diag (Channel-Flow/saveload.f90:1507-1676), vortcalc (:3929-4023) and
dissipation_calc (:4644-4882), with and without particles.  The inputs are
the device's own fields and populations after some steps, downloaded; the
oracle then evaluates the reference formulas on them."""
import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _channel(lx, ly, lz, seed, prec="fp64"):
    """Laminar start flow (U_c = 0.05) with 5% seeded noise."""
    ch = d3q19.Channel(lx, ly, lz, prec)
    ux, uy, uz = workloads.laminar_parabola_noise(lx, ly, lz, seed=seed)
    ch.set_macros(ux=ux, uy=uy, uz=uz)
    ch.FORCING()
    ch.initpop()
    ch.macrovar()
    return ch


def _oracle_with(ch, f=True):
    o = O.OracleChannel(ch.lx, ch.ly, ch.lz)
    rho, ux, uy, uz = ch.macrovar(download=True)
    o.rho[:], o.ux[:], o.uy[:], o.uz[:] = rho, ux, uy, uz
    if f:
        o.f[:] = ch.download_f()
    return o


@pytest.mark.parametrize("fields", ["fused", "arrays"])
def test_diag_matches_the_oracle(fields):
    """32x64x32 channel after 30 steps: the largest speed, its node and the
    density extrema bit for bit; means and rms to FP64 summation order.  The
    fields are the lazily fused ones, or stored arrays (set to the same
    values)."""
    with _channel(32, 64, 32, seed=7) as ch:
        ch.run(30)
        o = _oracle_with(ch, f=False)
        if fields == "arrays":
            ch.set_macros(rho=o.rho, ux=o.ux, uy=o.uy, uz=o.uz)
        got = ch.diag()
        want = o.diag()
    assert np.array_equal(bits(got[[0, 1, 2, 3, 10, 11, 12]]), bits(want[[0, 1, 2, 3, 10, 11, 12]]))
    np.testing.assert_allclose(got[4:7], want[4:7], rtol=1e-12, atol=1e-15)
    # rms = sqrt(<u^2> - <u>^2) loses (<u>/rms)^2 digits to the difference
    np.testing.assert_allclose(got[7:10], want[7:10], rtol=1e-10)


def test_diag_with_particles_masks_solid_nodes():
    """With particles the means, rms and extrema run over fluid nodes only and
    the solid fraction is reported (numpy on the downloaded fields and
    flags)."""
    lx, ly, lz = 32, 64, 32
    pos = workloads.sphere_positions(lx, ly, lz, 3, 5.0, seed=3)
    with _channel(lx, ly, lz, seed=8) as ch:
        ch.particles_create(pos, rad=5.0, visc=ch.p["visc"])
        ch.run_particles(12)
        got = ch.diag()
        rho, ux, uy, uz = ch.macrovar(download=True)
        fl = (ch.particles_solid() >= 0).reshape(lz, ly, lx) == False  # noqa: E712
    n = fl.sum()
    us = ch.p["ustar"]
    m = [a[fl].sum() / n for a in (ux, uy, uz)]
    r = [np.sqrt((a[fl] ** 2).sum() / n - mm ** 2) for a, mm in zip((ux, uy, uz), m)]
    np.testing.assert_allclose(got[4:10], np.array(m + r) / us, rtol=1e-10)
    speed = np.where(fl, np.sqrt(ux * ux + uy * uy + uz * uz), 0.0)
    k = int(np.argmax(speed.ravel()))  # first maximum in x, y, z order
    assert got[0] == speed.ravel()[k]
    assert tuple(got[1:4]) == (k % lx + 1, (k // lx) % ly + 1, k // (lx * ly) + 1)
    assert got[10] == pytest.approx(1.0 - n / (lx * ly * lz), rel=1e-15)
    assert (got[11], got[12]) == (rho[fl].max(), rho[fl].min())


def test_vortcalc_matches_the_oracle_bitwise():
    with _channel(24, 20, 18, seed=9) as ch:
        ch.run(10)
        o = _oracle_with(ch, f=False)
        got = ch.vortcalc()
    for g, w in zip(got, o.vortcalc()):
        assert np.array_equal(bits(g), bits(w))


def test_vortcalc_inside_particles():
    lx, ly, lz = 32, 64, 32
    pos = workloads.sphere_positions(lx, ly, lz, 3, 5.0, seed=3)
    omg = np.array([[0.001, 0.002, -0.003], [0.0, 0.004, 0.0], [-0.002, 0.0, 0.001]])
    with _channel(lx, ly, lz, seed=10) as ch:
        ch.particles_create(pos, omg=omg, rad=5.0)
        ch.run_particles(5)
        st = ch.particles_get()
        solid = ch.particles_solid().reshape(lz, ly, lx)
        o = _oracle_with(ch, f=False)
        got = ch.vortcalc()
    for g, w in zip(got, o.vortcalc(solid=solid, omg=st["omg"])):
        assert np.array_equal(bits(g), bits(w))
    assert np.all(got[0][solid == 1] == 2.0 * st["omg"][1, 0])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dissipation_matches_the_oracle(prec):
    """Plane means of S_ij S_ij and S_xy by both methods; FP64 per node in
    the reference's order (the plane sums in another order), FP32 storage
    to its rounding."""
    with _channel(24, 20, 18, seed=11, prec=prec) as ch:
        ch.run(10)
        got_fused = ch.dissipation()       # fields fused from the lattice
        o = _oracle_with(ch)
        ch.set_macros(rho=o.rho, ux=o.ux, uy=o.uy, uz=o.uz)
        got_arrays = ch.dissipation()      # the same fields from the arrays
        want = o.dissipation()
    tol = 1e-11 if prec == "fp64" else 1e-5
    for got in (got_fused, got_arrays):
        assert np.array_equal(got[:, :2], want[:, :2])
        scale = np.abs(want[:, 2:]).max(axis=0)
        assert np.all(np.abs(got[:, 2:] - want[:, 2:]) <= tol * scale), np.abs(got[:, 2:] - want[:, 2:]).max(0)
