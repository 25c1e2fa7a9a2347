"""Pins of the particle oracle (oracle/particles_oracle.c).

The reference has no particle source, so each piece is checked against what
the mathematics fixes: brute-force classification over every node and
direction, the cut-link fraction against bisection on |x + t c - X| = R, the
q = 1/2 case of the interpolated bounce-back against halfway bounce-back, the
discrete momentum balance between the fluid and a fixed particle, Newton's
third law for contacts, and the closed-form trajectory of a particle under a
constant force.
"""
import numpy as np
import pytest

from d3q19 import workloads
import oracle as O

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def make(lx, ly, lz, pos, rad, vel=None, omg=None, L0=None, force=(0.0, 0.0, 0.0), **kw):
    p = O.para(lx, ly, lz)
    g = O.particle_geometry(lx, ly, lz, len(pos), rad, force=force, **kw)
    N = lx * ly * lz
    if L0 is None:
        L0 = np.zeros(19 * N)
    return O.OracleParticleChannel(p, g, L0, pos, vel, omg)


def brute_force(lx, ly, lz, pos, rad):
    """Node flags and cut links by testing every node and direction."""
    ix, iy, iz = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="ij")
    ix, iy, iz = (a.transpose(2, 1, 0).ravel() for a in (ix, iy, iz))  # node order ix fastest
    solid = np.full(lx * ly * lz, -1)
    R2 = rad * rad
    for p in range(len(pos) - 1, -1, -1):
        X = pos[p]
        dx = (ix + 0.5) - X[0]
        dy = (iy + 0.5) - X[1]
        dy = np.where(dy > 0.5 * ly, dy - ly, np.where(dy < -0.5 * ly, dy + ly, dy))
        dz = (iz + 0.5) - X[2]
        dz = np.where(dz > 0.5 * lz, dz - lz, np.where(dz < -0.5 * lz, dz + lz, dz))
        solid[(dx * dx + dy * dy + dz * dz) < R2] = p
    links = set()
    for i in range(1, 19):
        jx = ix + CX[i]
        ok = (jx >= 0) & (jx < lx)
        jy = (iy + CY[i]) % ly
        jz = (iz + CZ[i]) % lz
        m = np.where(ok, jx + lx * (jy + ly * jz), 0)
        hit = ok & (solid < 0) & (solid[m] >= 0)
        for n in np.flatnonzero(hit):
            links.add((int(solid[m[n]]), int(n), i))
    return solid, links


def oracle_links(o):
    out = {}
    for p in range(o.g.np):
        for k in range(o.link_count[p]):
            s = p * o.g.maxlink + k
            out[(p, int(o.link_node[s]), int(o.link_dir[s]))] = float(o.link_q[s])
    return out


@pytest.mark.parametrize("case", ["interior", "periodic", "wall"])
def test_classification_matches_brute_force(case):
    lx, ly, lz, rad = 32, 24, 20, 4.3
    pos = {"interior": np.array([[16.2, 11.7, 9.9], [9.0, 3.1, 15.4]]),
           "periodic": np.array([[15.6, 0.4, 19.8], [20.0, 23.5, 2.2]]),
           "wall": np.array([[5.0, 12.0, 10.0], [27.2, 5.5, 5.5]])}[case]
    o = make(lx, ly, lz, pos, rad)
    solid, links = brute_force(lx, ly, lz, pos, rad)
    assert np.array_equal(o.solid, solid)
    got = oracle_links(o)
    assert set(got) == links
    # q against bisection on the sphere equation
    for (p, n, i), q in list(got.items())[::7]:
        x = np.array([n % lx + 0.5, (n // lx) % ly + 0.5, n // (lx * ly) + 0.5])
        d = x - pos[p]
        d[1] -= ly * np.round(d[1] / ly)
        d[2] -= lz * np.round(d[2] / lz)
        c = np.array([CX[i], CY[i], CZ[i]], dtype=float)
        lo, hi = 0.0, 1.0
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if np.dot(d + mid * c, d + mid * c) < rad * rad:
                hi = mid
            else:
                lo = mid
        assert abs(q - lo) < 1e-12


def test_links_cover_every_pull_from_a_solid_node():
    lx, ly, lz = 24, 20, 18
    pos = workloads.sphere_positions(lx, ly, lz, 2, 3.5, seed=8)
    o = make(lx, ly, lz, pos, 3.5)
    links = oracle_links(o)
    # every fluid node's pull from a solid neighbour must be a listed link
    for n in np.flatnonzero(o.solid < 0):
        ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
        for i in range(1, 19):
            jx = ix + CX[i]
            if 0 <= jx < lx:
                m = jx + lx * ((iy + CY[i]) % ly + ly * ((iz + CZ[i]) % lz))
                if o.solid[m] >= 0:
                    assert (int(o.solid[m]), int(n), i) in links


def test_half_link_is_halfway_bounce_back():
    """Centre on a node x-line, R = 3.5: the +x link from the node 4 below has q = 1/2."""
    lx, ly, lz = 20, 12, 12
    pos = np.array([[8.5, 6.5, 6.5]])
    rng = np.random.default_rng(1)
    L0 = 1e-3 * rng.standard_normal(19 * lx * ly * lz)
    vel = np.array([[0.01, -0.02, 0.005]])
    o = make(lx, ly, lz, pos, 3.5, vel=vel, L0=L0)
    N = lx * ly * lz
    n = 4 + lx * (6 + ly * 6)  # node (4, 6, 6): x = 4.5, 4 below the centre
    links = oracle_links(o)
    assert links[(0, n, 1)] == 0.5
    fi = o.post[1 * N + n]
    o.beads_collision()
    ns = n + 1
    cu = 1.0 * vel[0, 0]
    assert o.post[2 * N + ns] == fi - 6.0 * (1 / 18) * cu


def _fluid_momentum(o, L):
    N = o.N
    fluid = o.solid < 0
    Lr = L.reshape(19, N)
    return np.array([(CX[:, None] * Lr)[:, fluid].sum(), (CY[:, None] * Lr)[:, fluid].sum(),
                     (CZ[:, None] * Lr)[:, fluid].sum()])


def test_momentum_exchange_balances_the_fluid():
    """Fixed particle at rest, no body force, perturbation away from the walls:
    over one step the fluid's momentum changes by minus the exchanged force."""
    lx, ly, lz = 24, 20, 20
    pos = np.array([[12.3, 9.8, 10.4]])
    rng = np.random.default_rng(2)
    N = lx * ly * lz
    L0 = np.zeros((19, lz, ly, lx))
    L0[:, :, :, 4:lx - 4] = 1e-3 * rng.standard_normal((19, lz, ly, lx - 8))
    o = make(lx, ly, lz, pos, 4.0, L0=L0.reshape(-1))
    m0 = _fluid_momentum(o, o.post)
    o.beads_collision()
    o.collision_MRT()
    m1 = _fluid_momentum(o, o.post)
    assert np.max(np.abs((m1 - m0) + o.FH)) < 1e-14 * max(1.0, np.abs(o.FH).max())
    assert np.abs(o.FH).max() > 1e-6  # a real force was exchanged


def test_contact_forces_obey_newtons_third_law():
    lx, ly, lz = 40, 30, 30
    pos = np.array([[20.0, 10.0, 10.0], [20.5, 18.9, 11.3]])  # surface gap 0.33 < mingap
    o = make(lx, ly, lz, pos, 4.0)
    o.beads_move()
    assert np.any(o.FC != 0)
    assert np.array_equal(o.FC[0:3], -o.FC[3:6])


def test_constant_force_trajectory():
    """With a constant force the Adams-Bashforth-2/trapezoid update reproduces
    x_n = x_0 + a n^2 / 2 (Euler start v_1 = a, then v_n = n a)."""
    lx, ly, lz = 40, 40, 40
    pos = np.array([[20.0, 20.0, 20.0]])
    o = make(lx, ly, lz, pos, 3.0, pressure_force=0)
    a = np.array([1e-4, -2e-4, 5e-5])
    for n in range(1, 21):
        o.FH[:] = a * o.g.mass
        o.TQ[:] = 0.0
        o.beads_move()
        want = pos[0] + a * n * n / 2
        want[1] %= ly
        assert np.max(np.abs(o.X - want)) < 1e-12


def test_refill_sets_wall_velocity_and_neighbour_density():
    lx, ly, lz = 24, 20, 20
    pos = np.array([[12.0, 10.0, 10.0]])
    rng = np.random.default_rng(4)
    L0 = 1e-3 * rng.standard_normal(19 * lx * ly * lz)
    vel = np.array([[0.0, 0.3, 0.0]])  # moves 0.3 per step: uncovers nodes
    o = make(lx, ly, lz, pos, 4.0, vel=vel, L0=L0, pressure_force=0)
    before = o.solid.copy()
    o.beads_collision()
    o.collision_MRT()
    o.FH[:] = 0.0
    o.TQ[:] = 0.0
    o.beads_move()
    o.beads_links()
    fills = [int(o.fill_node[k]) for k in range(o.fill_count[0])]
    assert fills and all(before[n] == 0 for n in fills)
    o.beads_filling()
    N = o.N
    c = np.stack([CX, CY, CZ], 1)
    for n in fills:
        f = o.post.reshape(19, N)[:, n]
        u = c.T @ f
        assert np.allclose(u, o.V, atol=1e-15)
        assert o.solid[n] == -1
