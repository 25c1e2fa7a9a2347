"""The large channel at full size: Re_tau = 395 on 2048 x 1024 x 512 nodes
(lx = 512 wall-normal, ly = 2048 streamwise, lz = 1024 spanwise; 1.07 G
nodes), FP32 shifted storage (2 x 81.6 GB of lattices) and FP16 storage
(2 x 40.8 GB).  No full download is possible, so the checks are windowed or
size-independent:

* one step of a three-slab window (z = 499..501, all x and y) against the
  FP64 oracle's pull and collision on the same stored inputs, to the
  storage precision's rounding;
* the total FP64 mass over 10 steps moves by no more than the storage
  rounding of each step allows;
* a whole-plane roll of the start in y and z commutes with 10 steps: the
  rolled run's slabs equal the plain run's slabs rolled, bit for bit, and the
  plane statistics and mass agree to FP64 summation order;
* FP32 pairs of steps run the blocked sweep in row windows (the slabs are too
  large for a whole-slab ring): two pairs equal four single steps bit for bit
  on the z planes at every window edge.

Node indices at this size exceed 2^30 and FP16 byte offsets reach 2.1 GB, so
this exercises the 32-bit node arithmetic and the addressing at their limits.
"""
import ctypes

import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import oracle as O

pytestmark = pytest.mark.gpu

LX, LY, LZ = 512, 2048, 1024
PLANE = LX * LY
RSTAR = 395.0
Z0 = 500                 # the checked window is slabs Z0 - 1, Z0, Z0 + 1
SLABS_AFTER = (0, 1, 517, LZ - 1)
DY, DZ = 300, 77         # the roll of the second run


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _start(roll=False):
    """Log-law mean profile plus the seeded divergence-free perturbation, as
    separable terms summed on the device (3 x N host arrays would be 25.8 GB)."""
    ch_p = d3q19.para(LX, rstar=RSTAR)
    terms = workloads.loglaw_divfree_terms(LX, LY, LZ, ch_p["ustar"], ch_p["ystar"], seed=2)
    if roll:
        terms = [(c, np.roll(P, (DZ, DY), axis=(0, 1)), Q) for c, P, Q in terms]
    return terms


def _channel(prec, terms):
    print(f"large channel {prec}: creating the 2048x1024x512 handle", flush=True)
    ch = d3q19.Channel(LX, LY, LZ, prec, rstar=RSTAR)
    ch.set_velocity_separable(terms)
    ch.FORCING()
    ch.initpop()
    ch.macrovar()
    return ch


def _slab(ch, z):
    return ch.download_post_range(z * PLANE, PLANE)


def _oracle_window_step(ch, win):
    """The FP64 oracle's pull, macrovar and collision of the middle slab of a
    three-slab window (z wraps inside the window only for the outer slabs,
    whose results are not used)."""
    p = O.para(LX, LY, 3)
    assert p.s9 == ch.p["s9"] and p.s1 == ch.p["s1"]
    g = O.PGeom()
    g.lx, g.ly, g.lz = LX, LY, 3
    Lin = np.ascontiguousarray(win.reshape(-1))
    Lout = np.zeros_like(Lin)
    fy = ch.p["force_in_y"] * ch.p["force_mag"]
    O._plib().orc_p_step(ctypes.byref(p), ctypes.byref(g), None, Lin.ctypes.data_as(ctypes.c_void_p),
                         Lout.ctypes.data_as(ctypes.c_void_p), 0.0, fy, 0.0, 0.0)
    return Lout.reshape(19, 3 * PLANE)[:, PLANE:2 * PLANE]


def _mass_bound(steps, max_abs, rel):
    """Storage rounding moves each of the 19 N stored values by at most half
    an ulp per step."""
    return steps * 19 * LX * LY * LZ * rel * max_abs


@pytest.fixture(scope="module")
def start():
    return _start()


@pytest.mark.timeout(900)
def test_fp32_full_size(start):
    ch = _channel("fp32", start)
    with ch:
        # 80 MB / 40 MB slabs: y slabs in row windows of 256 z rows
        assert ch.block_info() == ("y", 20) and ch.block_rows() == 256
        win = np.concatenate([_slab(ch, z) for z in (Z0 - 1, Z0, Z0 + 1)], axis=1)
        m0 = ch.mass()
        ch.run(1)
        got = _slab(ch, Z0)
        want = _oracle_window_step(ch, win)
        err = np.abs(got - want).max()
        scale = np.abs(want).max()
        print(f"fp32 window step: max |diff| {err:.3e}, max |P| {scale:.3e}", flush=True)
        assert err <= 2.0 ** -19 * scale
        ch.run(9)
        m1 = ch.mass()
        assert abs(m1 - m0) <= _mass_bound(10, scale, 2.0 ** -24)
        stats = ch.statistc2()
        slabs = {z: _slab(ch, z) for z in SLABS_AFTER}
    del win, got, want
    print("fp32 rolled run", flush=True)
    rolled = _start(roll=True)
    ch = _channel("fp32", rolled)
    del rolled
    with ch:
        ch.run(10)
        # the same values summed in another order: the error bound of the
        # device's sums (~4e4 additions deep) over 19 N values of at most scale
        assert abs(ch.mass() - m1) <= 4e4 * 2.0 ** -53 * 19 * LX * LY * LZ * scale
        # plane sums of 2 M nodes in another order; the second moments are
        # differences of nearly equal sums, so their error is absolute
        np.testing.assert_allclose(ch.statistc2(), stats, rtol=1e-10, atol=1e-11)
        for z, s in slabs.items():
            r = _slab(ch, (z + DZ) % LZ).reshape(19, LY, LX)
            assert np.array_equal(bits(np.roll(r, -DY, axis=1)), bits(s.reshape(19, LY, LX))), z


# z planes at the edges of the four row windows (z rows 0-255, 256-511,
# 512-767, 768-1023), where a window's step-t edge rows are computed twice
WINDOW_EDGES = (0, 1, 255, 256, 511, 512, 767, 768, 1022, 1023)


@pytest.mark.timeout(900)
def test_fp32_row_windows_full_size_equal_single_steps(start, monkeypatch):
    """Two blocked pairs of the default sweep (y slabs in row windows of 256 z
    rows) against four single steps from the same start, bit for bit, on the z
    planes at every window edge, and the FP64 mass.  (The FP32 unit allows FMA
    contraction; the sweep and the single step inline the same node
    arithmetic.)"""
    ch = _channel("fp32", start)
    with ch:
        assert ch.block_rows() == 256
        ch.run(4)
        blocked = {z: _slab(ch, z) for z in WINDOW_EDGES}
        m_blk = ch.mass()
    print("fp32 single steps", flush=True)
    monkeypatch.setenv("LBM_BLOCK_PLANES", "0")
    ch = _channel("fp32", start)
    monkeypatch.delenv("LBM_BLOCK_PLANES")
    with ch:
        assert ch.block_info()[1] == 0
        ch.run(4)
        for z, a in blocked.items():
            assert np.array_equal(bits(_slab(ch, z)), bits(a)), f"z plane {z}: row windows differ from single steps"
        assert ch.mass() == m_blk


@pytest.mark.timeout(900)
def test_fp16_storage_full_size(start):
    ch = _channel("fp16", start)
    with ch:
        win = np.concatenate([_slab(ch, z) for z in (Z0 - 1, Z0, Z0 + 1)], axis=1)
        m0 = ch.mass()
        ch.run(1)
        got = _slab(ch, Z0)
        want = _oracle_window_step(ch, win)
        scale = np.abs(want).max()
        # one FP16 rounding of the stored value plus FP32 arithmetic
        tol = 2.0 ** -10 * np.abs(want) + 2.0 ** -19 * scale
        excess = (np.abs(got - want) - tol).max()
        print(f"fp16 window step: max |diff| {np.abs(got - want).max():.3e}, max |P| {scale:.3e}", flush=True)
        assert excess <= 0.0
        ch.run(9)
        m1 = ch.mass()
        assert np.isfinite(m1)
        assert abs(m1 - m0) <= _mass_bound(10, scale, 2.0 ** -11)
        for z in SLABS_AFTER:
            assert np.isfinite(_slab(ch, z)).all()
