"""lbm_mass_sum_noalloc (the entry point that allocates nothing) against
lbm_mass_sum, through the C ABI: the same FP64 sum bit for bit, on two calls
in a row, the second on a smaller lattice (fewer partial sums than the first
call left in the shared static scratch), in FP64 and FP32 storage, with
particles (fluid-node mask) and after steps."""
import ctypes

import numpy as np
import pytest

import d3q19
from d3q19 import workloads

pytestmark = pytest.mark.gpu


class _DeviceDouble:
    """One double of device memory from the HIP runtime liblbm itself links
    (the test allocates; the entry point under test does not).  torch is not
    used: its wheel carries a second HIP runtime, which fails to find the GPU
    once liblbm's runtime holds it in the same process."""

    def __init__(self):
        d3q19.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")  # already loaded by liblbm: the same runtime
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8)) == 0

    def clear_nan(self):
        nan = np.array([np.nan])
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8), 1) == 0

    def read(self):
        out = np.zeros(1)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8), 2) == 0
        return float(out[0])

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture
def dev():
    d = _DeviceDouble()
    yield d
    d.free()


def _noalloc(ch, dev):
    dev.clear_nan()
    ch.mass_noalloc(dev.ptr.value)
    ch.sync()
    return dev.read()


def _same(a, b):
    return np.array([a]).view(np.uint64)[0] == np.array([b]).view(np.uint64)[0]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_mass_equals_mass_sum_on_two_calls(prec, dev):
    shapes = [(96, 64, 48), (12, 10, 6)]  # 1152 partials, then 3
    for k, (lx, ly, lz) in enumerate(shapes):
        f = (workloads.uniform(40 + k, 19 * lx * ly * lz) - 0.5).reshape(lz, ly, lx, 19) * 1e-3
        with d3q19.Channel(lx, ly, lz, prec) as ch:
            ch.upload_f(f)
            ch.FORCING()
            ch.macrovar()
            ch.run(3)
            want = ch.mass()
            got = _noalloc(ch, dev)
            assert np.isfinite(got), f"{lx}x{ly}x{lz}: no value written"
            assert _same(got, want), f"{lx}x{ly}x{lz} {prec}: {got!r} != {want!r}"


def test_noalloc_mass_with_particles(dev):
    lx, ly, lz = 24, 20, 18
    pos = workloads.sphere_positions(lx, ly, lz, 2, 3.5, seed=8)
    f = (workloads.uniform(9, 19 * lx * ly * lz) - 0.5).reshape(lz, ly, lx, 19) * 1e-3
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.upload_f(f)
        ch.FORCING()
        ch.macrovar()
        ch.particles_create(pos, rad=3.5)
        want = ch.mass()
        got = _noalloc(ch, dev)
        assert _same(got, want), f"{got!r} != {want!r}"


def test_noalloc_rejects_host_output():
    lib = d3q19.load_library()
    host = np.zeros(1)
    with d3q19.Channel(12, 10, 6) as ch:
        rc = lib.lbm_mass_sum_noalloc(ch._h, host.ctypes.data_as(ctypes.c_void_p))
        assert rc == 1  # LBM_ERR_ARG, before any launch
        assert b"not device memory" in lib.lbm_last_error()
        assert lib.lbm_mass_sum_noalloc(ch._h, None) == 1
