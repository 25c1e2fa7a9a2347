"""GPU parity of the particle path (beads_links, beads_collision,
beads_move, beads_filling and the particle-aware collision_MRT) against the
particle oracle: node flags, cut links (with q) and the coupled FP64
fluid-particle state must be bitwise equal."""
import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def compare_links(ch, o):
    """Counts, nodes, directions and q bits of every particle's link list."""
    cap, counts, node, dirs, q = ch.particles_links()
    assert cap == o.g.maxlink
    assert np.array_equal(counts, o.link_count)
    valid = (np.arange(cap)[None, :] < counts[:, None]).ravel()
    assert np.array_equal(node[valid], o.link_node[valid])
    assert np.array_equal(dirs[valid], o.link_dir[valid])
    assert np.array_equal(bits(q[valid]), bits(o.link_q[valid]))


def test_classification_bitwise_config1():
    lx, ly, lz, rad = 32, 64, 32, 5.0
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.particles_create(pos, rad=rad)
        o = O.OracleParticleChannel(O.para(lx, ly, lz), O.particle_geometry(lx, ly, lz, 3, rad), None, pos)
        assert np.array_equal(ch.particles_solid(), o.solid)
        compare_links(ch, o)


def test_classification_bitwise_2000_spheres_config2_grid():
    """2000 spheres of radius 5 on the 192x768x384 grid (volume fraction 1.85%)."""
    lx, ly, lz, rad = 192, 768, 384, 5.0
    pos = workloads.sphere_positions(lx, ly, lz, 2000, rad, seed=3)
    with d3q19.Channel(lx, ly, lz, "fp32") as ch:
        ch.particles_create(pos, rad=rad)
        o = O.OracleParticleChannel(O.para(lx, ly, lz), O.particle_geometry(lx, ly, lz, 2000, rad), None, pos)
        solid = ch.particles_solid()
        assert np.array_equal(solid, o.solid)
        frac = np.count_nonzero(solid >= 0) / solid.size
        assert abs(frac - 2000 * 4 / 3 * np.pi * rad**3 / solid.size) < 2e-3
        compare_links(ch, o)


def test_coupled_particle_run_bitwise():
    """Three moving spheres in the 32x64x32 channel, 60 coupled FP64 steps."""
    lx, ly, lz, rad, nsteps = 32, 64, 32, 5.0, 60
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    vel = np.array([[0.0, 0.03, 0.0], [0.002, -0.02, 0.01], [0.0, 0.0, -0.025]])
    omg = np.array([[0.0, 0.0, 0.01], [0.005, 0.0, 0.0], [0.0, -0.004, 0.0]])
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=31, uc=0.03)
    ref.initpop()
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.upload_f(ref.f)
        ch.FORCING()
        ch.macrovar()
        L0 = ch.download_post()
        ch.particles_create(pos, vel=vel, omg=omg, rad=rad)
        g = O.particle_geometry(lx, ly, lz, 3, rad, force=(0.0, ch.p["force_in_y"], 0.0))
        o = O.OracleParticleChannel(O.para(lx, ly, lz), g, L0, pos, vel, omg)
        ch.run_particles(nsteps)
        o.step(nsteps)
        st = ch.particles_get()
        for key, arr in (("pos", o.X), ("vel", o.V), ("omg", o.W), ("force", o.FH), ("torque", o.TQ),
                         ("contact", o.FC), ("lub", o.FL)):
            assert np.array_equal(bits(st[key].ravel()), bits(arr)), key
        assert np.array_equal(ch.particles_solid(), o.solid)
        post = ch.download_post().ravel()
        assert np.array_equal(bits(post), bits(o.post))
        # the particles moved across nodes, so refill was exercised
        assert np.max(np.abs(st["pos"] - pos)) > 1.0


def test_turbulent_constants_coupled_run_bitwise_300_steps():
    """The 32x64x32 channel at the turbulent constants (tau = 0.5108, rate
    set 1, Re_tau = 180 body force) with three freely moving neutrally
    buoyant spheres of radius 5 released at rest into the laminar start flow
    (an impulsive start), lubrication on, avedensity after steps 100, 200 and
    300 (the last one due in a later call): 300 coupled FP64 steps in two
    calls, bitwise against the oracle, fields inside particles included."""
    lx, ly, lz, rad, nsteps = 32, 64, 32, 5.0, 300
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=31)
    ref.initpop()
    with d3q19.Channel(lx, ly, lz) as ch:
        assert abs(1.0 / ch.p["s9"] - 0.5108) < 1e-12
        ch.upload_f(ref.f)
        ch.FORCING()
        ch.macrovar()
        L0 = ch.download_post()
        visc = ch.p["visc"]
        ch.particles_create(pos, rad=rad, visc=visc)
        g = O.particle_geometry(lx, ly, lz, 3, rad, visc=visc, force=(0.0, ch.p["force_in_y"], 0.0))
        o = O.OracleParticleChannel(O.para(lx, ly, lz), g, L0, pos)
        ch.run_particles(100)   # the mean due after step 100 is taken by the next call
        ch.run_particles(200)
        o.step(nsteps, istep0=0)
        assert o.ave_pending
        st = ch.particles_get()
        for key, arr in (("pos", o.X), ("vel", o.V), ("omg", o.W), ("force", o.FH), ("torque", o.TQ),
                         ("contact", o.FC), ("lub", o.FL)):
            assert np.array_equal(bits(st[key].ravel()), bits(arr)), key
        assert np.array_equal(ch.particles_solid(), o.solid)
        assert np.array_equal(bits(ch.download_post().ravel()), bits(o.post))
        assert np.all(np.isfinite(o.post)) and np.max(np.abs(st["vel"])) > 1e-3
        for name, a, b in zip(("rho", "ux", "uy", "uz"), ch.macrovar(download=True), o.macrovar()):
            assert np.array_equal(bits(a), bits(b)), name


def test_avedensity_mean_over_fluid_nodes_with_particles():
    """avedensity with particles: the mean over fluid nodes only, in the
    device's summation order, equals the oracle's restatement bit for bit."""
    lx, ly, lz, rad = 32, 64, 32, 5.0
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    ref = O.OracleChannel(lx, ly, lz)
    ref.ux[:], ref.uy[:], ref.uz[:] = workloads.laminar_parabola_noise(lx, ly, lz, seed=33, uc=0.03)
    ref.initpop()
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.upload_f(ref.f)
        ch.FORCING()
        ch.macrovar()
        L0 = ch.download_post()
        ch.particles_create(pos, rad=rad)
        g = O.particle_geometry(lx, ly, lz, 3, rad, force=(0.0, ch.p["force_in_y"], 0.0))
        o = O.OracleParticleChannel(O.para(lx, ly, lz), g, L0, pos)
        ch.run_particles(7)
        o.step(7, istep0=0)
        mean = ch.avedensity()
        want = o.rho_mean()
        assert np.array_equal(bits(np.array([mean])), bits(np.array([want])))
        assert mean != 0.0


def test_particle_run_fp32_is_close_to_fp64():
    lx, ly, lz, rad, nsteps = 32, 64, 32, 5.0, 40
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    vel = np.array([[0.0, 0.02, 0.0]] * 3)
    vfield = workloads.laminar_parabola_noise(lx, ly, lz, seed=32, uc=0.03)
    out = {}
    for prec in ("fp64", "fp32"):
        with d3q19.Channel(lx, ly, lz, prec) as ch:
            ch.set_macros(ux=vfield[0], uy=vfield[1], uz=vfield[2])
            ch.FORCING()
            ch.initpop()
            ch.macrovar()
            ch.particles_create(pos, vel=vel, rad=rad)
            ch.run_particles(nsteps)
            out[prec] = ch.particles_get()
    for key in ("pos", "vel"):
        assert np.allclose(out["fp32"][key], out["fp64"][key], rtol=1e-4, atol=1e-5), key
    assert np.allclose(out["fp32"]["force"], out["fp64"]["force"], rtol=1e-2, atol=1e-5)


def test_diverged_particle_is_reported_not_indexed():
    """A particle whose state becomes non-finite (here: an infinite initial
    velocity, so its centre is infinite after the first move) yields empty
    node boxes and an error from run_particles; the device stays usable."""
    lx, ly, lz, rad = 32, 64, 32, 5.0
    pos = workloads.sphere_positions(lx, ly, lz, 3, rad, seed=3)
    vel = np.zeros((3, 3))
    vel[1, 0] = np.inf
    with d3q19.Channel(lx, ly, lz) as ch:
        ch.initvel()
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        ch.particles_create(pos, vel=vel, rad=rad)
        with pytest.raises(d3q19.LbmError, match="non-finite"):
            ch.run_particles(3)
    # the card still runs an ordinary channel
    with d3q19.Channel(12, 10, 6) as ch:
        ch.initvel()
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        ch.run(5)
        assert np.isfinite(ch.download_f()).all()
