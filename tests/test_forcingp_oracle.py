"""Pins of the oracle's FORCINGP (oracle/lbm_oracle.c, orc_forcingp), the
perturbation forcing of Channel-Flow/collision.f90:529-602.

The reference ships no output of this routine, so it is pinned by what the
formula fixes independently of how it is typed:

* hand-evaluated values at nodes where every angle is a multiple of pi/4, so
  each factor is 0, +-1 or +-sqrt(2)/2;
* at istep = 0 the amplitude sin(0) vanishes and the field is FORCING's
  uniform force exactly;
* outside the two bands of lx/4 nodes (x = 3 .. lx/4 + 2 and the mirror
  image, 1-based) the force is the mean force exactly;
* the perturbation is divergence-free (checked with fourth-order differences;
  a 2 % error in any one component's amplitude fails);
* the second band is the first one shifted a quarter period in y with the
  perturbation's sign reversed;
* the perturbation scales with sin(2 pi istep / 2000).
"""
import math

import numpy as np
import pytest

import oracle as O


def field(lx, ly, lz, istep):
    ch = O.OracleChannel(lx, ly, lz)
    ch.FORCINGP(istep)
    return ch, ch.p.force_in_y


def test_forcingp_hand_values_at_quarter_angles():
    # lx = 16: bands of ihh = 4 nodes at 0-based x = 2..5 and 10..13; ny = 12,
    # nz = 8, istep = 500: amplitude 40*3/12*sin(pi/2) = 10.  At the first node
    # of each band x9 = pi/4, 3*y9 = pi/4 (band 2: 7*pi/4), 2*z9 = pi/4.
    ch, fiy = field(16, 12, 8, 500)
    r = math.sqrt(0.5)
    want1 = (10 * (1 - r), 1 - 40 * r**3, 20 * r**3)
    want2 = (-10 * (1 - r), 1 - 40 * r**3, -20 * r**3)
    for ix, want in ((2, want1), (10, want2)):
        got = (ch.fx[0, 0, ix] / fiy, ch.fy[0, 0, ix] / fiy, ch.fz[0, 0, ix] / fiy)
        assert got == pytest.approx(want, rel=1e-13), (ix, got, want)
    # second node of band 1 (x9 = 3 pi/4), same y, z: 1 - cos = 1 + r, sin = r
    got = (ch.fx[0, 0, 3] / fiy, ch.fy[0, 0, 3] / fiy, ch.fz[0, 0, 3] / fiy)
    assert got == pytest.approx((10 * (1 + r), 1 - 40 * r**3, 20 * r**3), rel=1e-13)


def test_forcingp_at_step_zero_is_forcing():
    lx, ly, lz = 24, 16, 12
    ch, fiy = field(lx, ly, lz, 0)
    ref = O.OracleChannel(lx, ly, lz)
    ref.FORCING()
    assert ref.p.force_mag == 1.0
    for a, b in ((ch.fx, ref.fx), (ch.fy, ref.fy), (ch.fz, ref.fz)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("lx", [8, 16, 30, 64])
def test_forcingp_bands(lx):
    ly, lz = 12, 8
    ch, fiy = field(lx, ly, lz, 700)
    ihh = (lx // 2) // 2
    inband = np.zeros(lx, bool)
    inband[2:2 + ihh] = True
    inband[lx - 2 - ihh:lx - 2] = True
    out = ~inband
    assert np.all(ch.fx[:, :, out] == 0.0) and np.all(ch.fz[:, :, out] == 0.0)
    assert np.all(ch.fy[:, :, out] == fiy)
    # every band node carries a perturbation somewhere in its plane
    pert = np.abs(ch.fx) + np.abs(ch.fy - fiy) + np.abs(ch.fz)
    assert np.all(pert[:, :, inband].max(axis=(0, 1)) > 1e-3 * fiy)


def d4(a, axis, periodic=True):
    """Fourth-order central difference along axis (unit spacing)."""
    s = lambda k: np.roll(a, -k, axis=axis)
    return (8 * (s(1) - s(-1)) - (s(2) - s(-2))) / 12.0


def divergence_ratio(fx, fy, fz, fiy, x_lo, x_hi, scale=(1.0, 1.0, 1.0)):
    px, py, pz = scale[0] * fx, scale[1] * (fy - fiy), scale[2] * fz
    terms = [d4(px, 2), d4(py, 1), d4(pz, 0)]
    sl = (slice(None), slice(None), slice(x_lo, x_hi))
    div = (terms[0] + terms[1] + terms[2])[sl]
    big = max(np.abs(t[sl]).max() for t in terms)
    return float(np.abs(div).max() / big)


def test_forcingp_perturbation_is_divergence_free():
    lx, ly, lz = 128, 96, 64  # ihh = 32 nodes per band
    ch, fiy = field(lx, ly, lz, 500)
    ihh = lx // 4
    lo, hi = 2 + 3, 2 + ihh - 3  # band 1 interior (the x stencil stays inside)
    ratio = divergence_ratio(ch.fx, ch.fy, ch.fz, fiy, lo, hi)
    assert ratio < 2e-3, ratio
    lo2, hi2 = lx - 2 - ihh + 3, lx - 2 - 3
    assert divergence_ratio(ch.fx, ch.fy, ch.fz, fiy, lo2, hi2) < 2e-3
    # the check has teeth: a 2 % amplitude error in one component fails it
    for sc in ((1.02, 1, 1), (1, 1.02, 1), (1, 1, 1.02)):
        assert divergence_ratio(ch.fx, ch.fy, ch.fz, fiy, lo, hi, sc) > 5e-3


def test_forcingp_second_band_is_the_first_shifted():
    lx, ly, lz = 32, 24, 8
    ch, fiy = field(lx, ly, lz, 333)
    ihh = lx // 4
    b1 = slice(2, 2 + ihh)
    b2 = slice(lx - 2 - ihh, lx - 2)
    q = ly // 4
    for a, off in ((ch.fx, 0.0), (ch.fy, fiy), (ch.fz, 0.0)):
        p1 = np.roll(a[:, :, b1] - off, -q, axis=1)
        p2 = a[:, :, b2] - off
        np.testing.assert_allclose(p2, -p1, rtol=0, atol=1e-12 * np.abs(p1).max())


def test_forcingp_amplitude_follows_the_period():
    lx, ly, lz = 16, 12, 8
    ref, fiy = field(lx, ly, lz, 500)  # sin(pi/2) = 1
    for istep, factor in ((250, math.sin(math.pi / 4)), (1500, -1.0), (1900, math.sin(2 * math.pi * 0.95))):
        ch, _ = field(lx, ly, lz, istep)
        for a, b, off in ((ch.fx, ref.fx, 0.0), (ch.fy, ref.fy, fiy), (ch.fz, ref.fz, 0.0)):
            np.testing.assert_allclose(a - off, factor * (b - off), rtol=0, atol=1e-12 * np.abs(b - off).max())
