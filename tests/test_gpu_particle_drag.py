"""Closed-form pin of the particle path's hydrodynamics: the Stokes drag of a
sphere translating parallel to the walls, midway between them.

A sphere of radius a moving at speed U through fluid at rest, midway between
two plane walls a distance l from its centre, feels the drag
F = 6 pi mu a U K with Faxen's wall correction
    K = 1 / (1 - 1.004 k + 0.418 k^3 + 0.21 k^4 - 0.169 k^5),   k = a / l
(Faxen 1923; Happel & Brenner, Low Reynolds Number Hydrodynamics, 1965,
sec. 7-4).  The lattice's walls are the channel's halfway bounce-back walls
(x = 0 and x = lx); the sphere is held at constant velocity by a density
ratio of 1e9, so the interpolated bounce-back with its moving-wall term, the
momentum exchange and the refill of uncovered nodes all take part.  The
reference names these particle routines but ships none of their code
(Channel-Flow/var_inc.f90:166-168); the oracle's own pins fix the formulas,
and this test fixes what they add up to.

Tolerances, written against the effects they cover:
* drag against Faxen: 4 %.  Periodic images in y and z and the mean flow the
  sphere drives through the periodic channel lower the drag by about 1 %
  (flux estimate F h^2 / (2 nu V) against U); the Oseen inertia correction at
  Re = U a / nu = 0.03 is at most 3/8 Re = 1.1 %; a radius-5 sphere's
  hydrodynamic radius differs from 5 by about a per cent with linear
  interpolated bounce-back.
* drag at 2U against twice the drag at U: 1.5 % (Stokes flow is linear; the
  inertia correction is the only difference).
* force across the walls and along z, and the torque: zero, because the
  sphere's centre lies on mirror planes of the lattice in x (x = lx/2, the
  midplane between nodes lx/2 and lx/2 + 1) and in z, and its motion along y
  keeps both mirrors; bounded by 1e-9 of F and of F a (rounding of the link
  sums only).
A wrong sign of the moving-wall term reverses the drag; a momentum exchange
that drops either population of a link roughly halves it.  Measured on
MI355X: drag / Faxen 1.0010, drag at 2U / (2 x drag at U) 1.00017.
"""
import numpy as np
import pytest

import d3q19

pytestmark = pytest.mark.gpu

NU = 1.0 / 6.0  # tau = 1
LX, LY, LZ, RAD = 40, 256, 256, 5.0


def faxen_midplane(k):
    return 1.0 / (1.0 - 1.004 * k + 0.418 * k**3 + 0.21 * k**4 - 0.169 * k**5)


def mean_force(u, nsteps=5000, window=2000, every=20):
    """Mean momentum-exchange force and torque over the last `window` steps
    of a sphere moving at (0, u, 0) from rest fluid, midway between the walls."""
    pos = np.array([[LX / 2.0, LY / 4.0, LZ / 2.0]])
    vel = np.array([[0.0, u, 0.0]])
    with d3q19.Channel(LX, LY, LZ, "fp64", visc=NU) as ch:
        ch.set_force(0.0, 0.0, 0.0)
        z = np.zeros(ch.N)
        ch.set_macros(rho=z, ux=z, uy=z, uz=z)
        ch.initpop()
        ch.macrovar()
        ch.particles_create(pos, vel=vel, rad=RAD, rhopart=1e9, pressure_force=False)
        ch.run_particles(nsteps - window)
        fs, ts = [], []
        for _ in range(window // every):
            ch.run_particles(every)
            st = ch.particles_get()
            fs.append(st["force"][0])
            ts.append(st["torque"][0])
        end = ch.particles_get()
    return np.mean(fs, axis=0), np.mean(ts, axis=0), end


def test_drag_of_a_sphere_between_two_walls():
    u = 1e-3
    print(f"drag: {LX}x{LY}x{LZ} fp64, a={RAD}, u={u}", flush=True)
    f, t, end = mean_force(u)
    assert abs(end["vel"][0][1] - u) < 1e-5 * u, "the sphere's velocity changed"
    drag = -f[1]
    want = 6.0 * np.pi * NU * RAD * u * faxen_midplane(RAD / (LX / 2.0))
    print(f"drag {drag:.6e} Faxen {want:.6e} ratio {drag / want:.4f}; force {f}, torque {t}", flush=True)
    assert abs(drag / want - 1.0) < 0.04
    assert abs(f[0]) < 1e-9 * drag and abs(f[2]) < 1e-9 * drag
    assert np.all(np.abs(t) < 1e-9 * drag * RAD)

    print(f"drag: u={2 * u}", flush=True)
    f2, _, _ = mean_force(2 * u)
    print(f"drag at 2u {-f2[1]:.6e}, ratio to twice the drag {-f2[1] / (2 * drag):.5f}", flush=True)
    assert abs(-f2[1] / (2.0 * drag) - 1.0) < 0.015
