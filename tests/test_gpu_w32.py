"""2-byte storage (and the FP32 closure kernel) addresses the pulls with
c_x != 0 either from two 64-bit row bases or, where the lattice allows
(N <= 2^29 in 2-byte storage, 2^28 in FP32), from the lower row with a 32-bit
offset (step mode + 32, csrc/kernels_impl.h: pull).  The two forms load the
same values, so every mode that takes the 32-bit form must give the same
stored words as LBM_STEP_W32=0 does, bit for bit."""
import numpy as np
import pytest

import d3q19
from d3q19 import workloads
from test_gpu_parity import bits, seeded_pair

pytestmark = pytest.mark.gpu


def _twin(monkeypatch, make, first=None):
    """make() -> result for a handle with LBM_STEP_W32=first (unset: the
    default) and one with LBM_STEP_W32=0."""
    out = []
    for env in (first, "0"):
        if env is None:
            monkeypatch.delenv("LBM_STEP_W32", raising=False)
        else:
            monkeypatch.setenv("LBM_STEP_W32", env)
        out.append(make())
    monkeypatch.delenv("LBM_STEP_W32", raising=False)
    return out


@pytest.mark.parametrize("form", [None, "adjacent"])
@pytest.mark.parametrize("prec", ["fp16", "fp16c"])
@pytest.mark.parametrize("case", ["plain", "rho_offset", "forcingp"])
def test_w32_pulls_store_the_same_words(monkeypatch, prec, case, form):
    """form None: all walled pulls with 32-bit offsets (N <= 2^29); "adjacent":
    only those between adjacent rows, the form of lattices above 2^29 nodes."""
    lx, ly, lz = 24, 40, 30
    print(f"w32 {prec} {case} {form}", flush=True)
    _, _, f0 = seeded_pair(lx, ly, lz, seed=71, noise=1e-4)

    def make():
        with d3q19.Channel(lx, ly, lz, prec) as ch:
            ch.upload_f(f0)
            ch.FORCING()
            ch.macrovar()
            if case == "plain":
                ch.run(21)
            elif case == "rho_offset":
                for _ in range(3):
                    ch.avedensity()
                    ch.collision_MRT()
                    ch.macrovar()
            else:
                ch.run_forcingp(0, 20, 7)
            return ch.download_post()

    a, b = _twin(monkeypatch, make, form)
    assert np.array_equal(bits(a), bits(b)), f"{prec} {case} {form}: 32-bit and row-base pulls differ"


@pytest.mark.parametrize("prec", ["fp16", "fp16c"])
def test_w32_pulls_with_particles(monkeypatch, prec):
    lx, ly, lz, rad = 32, 64, 32, 5.0
    print(f"w32 {prec} particles", flush=True)
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    vel = np.array([[0.0, 0.02, 0.0]] * 3)
    vfield = workloads.laminar_parabola_noise(lx, ly, lz, seed=32, uc=0.03)

    def make():
        with d3q19.Channel(lx, ly, lz, prec) as ch:
            ch.set_macros(ux=vfield[0], uy=vfield[1], uz=vfield[2])
            ch.FORCING()
            ch.initpop()
            ch.macrovar()
            ch.particles_create(pos, vel=vel, rad=rad)
            ch.run_particles(20)
            return ch.download_post(), ch.particles_get()

    (fa, pa), (fb, pb) = _twin(monkeypatch, make)
    assert np.array_equal(bits(fa), bits(fb)), f"{prec}: lattice differs"
    for key in ("pos", "vel", "force"):
        assert np.array_equal(pa[key], pb[key]), f"{prec}: particle {key} differs"


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16c"])
@pytest.mark.parametrize("case", ["plain", "forcingp"])
def test_w32_pulls_in_the_closure_kernel(monkeypatch, prec, case):
    """The closure modes (8, and 24 under a force per node) with 32-bit pull
    offsets against the row-base form."""
    lx, ly, lz = 20, 24, 18
    print(f"w32 closure {prec} {case}", flush=True)
    _, _, f0 = seeded_pair(lx, ly, lz, seed=72, noise=1e-4)

    def make():
        with d3q19.Channel(lx, ly, lz, prec) as ch:
            p = d3q19.closure_params(seed=5, nu=ch.p["visc"])
            ch.upload_f(f0)
            ch.FORCING()
            ch.macrovar()
            ch.closure_set(p, ystar=ch.p["ystar"])
            if case == "plain":
                ch.run(12)
            else:
                ch.run_forcingp(0, 12, 10**9)
            return ch.download_post()

    a, b = _twin(monkeypatch, make)
    assert np.array_equal(bits(a), bits(b)), f"{prec} {case}: 32-bit and row-base closure pulls differ"
