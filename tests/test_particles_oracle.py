"""Pins of the particle oracle (oracle/particles_oracle.c).

The reference has no particle source, so each piece is checked against what
the mathematics fixes:
* classification: brute force over every node and direction; the cut-link
  fraction against bisection on |x + t c - X| = R;
* interpolated bounce-back: on a planar wall, populations that are linear
  along each link and mirrored at the wall (the exact reflection of a linear
  field) must be reproduced exactly for q < 1/2 and q >= 1/2, for a fixed and
  a moving wall; the fallback where the second node is missing is halfway
  bounce-back;
* Galilean invariance: a sphere moving with a uniform flow feels no force
  and keeps its velocity over steps that cover and uncover nodes;
* momentum exchange: the discrete momentum balance between the fluid and a
  fixed particle, and, for a moving particle, the balance in the particle's
  frame (force = -(fluid momentum change) + U (fluid mass change));
* refill: the refilled node carries the mean density of its fluid neighbours
  and the rigid-body velocity;
* contact (Feng & Michaelides 2005, eq. 28) and lubrication (Nguyen & Ladd
  2002) forces against hand-evaluated values; Newton's third law;
* integration: the closed-form trajectory under a constant force; the
  two-step force average removing a force that alternates each step;
* avedensity's device-order sum against exact sums; macrovar's solid-node
  velocity against the rigid-body formula.
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
    """A node uncovered by the move gets the equilibrium of the rigid-body
    velocity V + W x r at the node and of the mean density of its settled
    fluid neighbours (read before any refill writes)."""
    lx, ly, lz = 24, 20, 20
    pos = np.array([[12.0, 10.0, 10.0]])
    rng = np.random.default_rng(4)
    L0 = 1e-3 * rng.standard_normal(19 * lx * ly * lz)
    vel = np.array([[0.05, 0.3, -0.1]])  # moves 0.3 per step: uncovers nodes
    omg = np.array([[0.01, -0.02, 0.015]])
    o = make(lx, ly, lz, pos, 4.0, vel=vel, omg=omg, L0=L0, pressure_force=0)
    before = o.solid.copy()
    o.beads_collision()
    o.collision_MRT()
    o.FH[:] = 0.0
    o.TQ[:] = 0.0
    o.beads_move()
    o.beads_links()
    fills = [int(o.fill_node[k]) for k in range(o.fill_count[0])]
    assert fills and all(before[n] == 0 for n in fills)
    N = o.N
    P = o.post.reshape(19, N).copy()
    rho_node = P.sum(axis=0)
    want_rho = {}
    for n in fills:
        ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
        nb = []
        for j in range(1, 19):
            jx = ix + CX[j]
            if 0 <= jx < lx:
                m = jx + lx * ((iy + CY[j]) % ly + ly * ((iz + CZ[j]) % lz))
                if o.solid[m] == -1:
                    nb.append(rho_node[m])
        want_rho[n] = np.mean(nb)
    o.beads_filling()
    c = np.stack([CX, CY, CZ], 1)
    for n in fills:
        f = o.post.reshape(19, N)[:, n]
        ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
        r = np.array([ix + 0.5, iy + 0.5, iz + 0.5]) - o.X
        r[1:] -= np.array([ly, lz]) * np.round(r[1:] / np.array([ly, lz]))
        uw = o.V + np.cross(o.W, r)
        assert np.allclose(c.T @ f, uw, rtol=0, atol=1e-15)
        assert abs(f.sum() - want_rho[n]) < 1e-15
        assert o.solid[n] == -1


def test_moving_particle_momentum_balance_in_its_frame():
    """A translating particle (velocity U, no rotation, no body force): over one
    bounce-back and collision the exchanged force (Wen et al. 2014) equals
    minus the fluid's momentum change plus U times the fluid's mass change,
    i.e. momentum balances in the particle's frame.  Plain momentum exchange
    (no U terms) or a sign error in them fails by |U dM|."""
    lx, ly, lz = 24, 20, 20
    pos = np.array([[12.3, 9.8, 10.4]])
    U = np.array([0.011, -0.023, 0.017])
    rng = np.random.default_rng(12)
    N = lx * ly * lz
    L0 = np.zeros((19, lz, ly, lx))
    L0[:, :, :, 4:lx - 4] = 1e-3 * rng.standard_normal((19, lz, ly, lx - 8))
    o = make(lx, ly, lz, pos, 4.0, vel=np.array([U]), L0=L0.reshape(-1))
    fluid = o.solid < 0
    m0 = _fluid_momentum(o, o.post)
    M0 = o.post.reshape(19, N)[:, fluid].sum()
    o.beads_collision()
    o.collision_MRT()
    m1 = _fluid_momentum(o, o.post)
    M1 = o.post.reshape(19, N)[:, fluid].sum()
    want = -(m1 - m0) + U * (M1 - M0)
    assert np.max(np.abs(o.FH - want)) < 1e-13
    assert np.max(np.abs(U * (M1 - M0))) > 1e-6  # the mass term matters here


# --- interpolated bounce-back on a planar wall --------------------------------

PLUS_X = [i for i in range(1, 19) if CX[i] == 1]  # directions cut by a wall on the +x side


def _planar_wall(q, vel, behind_solid=()):
    """Oracle state with a planar wall at x = 5.5 + q: nodes ix >= 6 are solid
    (particle 0, whose sphere lies far outside the grid), and each fluid node
    (5, iy, iz), 1 <= iy, iz <= 6, has a cut link with fraction q in every
    direction with c_x = +1.  Nodes (4, iy, iz) in behind_solid are solid too,
    so links that would interpolate from them take the fallback."""
    lx, ly, lz = 12, 8, 8
    N = lx * ly * lz
    o = make(lx, ly, lz, np.array([[lx + 100.0, 4.0, 4.0]]), 1.0, vel=np.array([vel]))
    solid = np.full((lz, ly, lx), -1, dtype=np.int32)
    solid[:, :, 6:] = 0
    for iy, iz in behind_solid:
        solid[iz, iy, 4] = 0
    o.solid[:] = solid.ravel()
    k = 0
    for iz in range(1, 7):
        for iy in range(1, 7):
            for i in PLUS_X:
                o.link_node[k] = 5 + lx * (iy + ly * iz)
                o.link_dir[k] = i
                o.link_q[k] = q
                k += 1
    o.link_count[0] = k
    return o, (lx, ly, lz, N)


def _mirrored_linear_field(q, vel, shape, seed):
    """Populations linear in position along every link direction with c_x = 1
    and mirrored at the wall plane X_w = 5.5 + q: for each such i, with
    P_i(x) = a + b (x - X_w) + g . (y, z), the opposite population is
    P_opp(x) = (a - 6 w_i c_i . U) + (-b - 2 g . (c_y, c_z)) (x - X_w) + g . (y, z).
    That is the exact reflection: the opposite population at a point a
    distance s in front of the wall equals the incoming one a distance s
    behind it, less the moving-wall term, for every s along the link line."""
    lx, ly, lz, N = shape
    rng = np.random.default_rng(seed)
    P = 1e-2 * rng.standard_normal((19, lz, ly, lx))
    Z, Y, X = np.meshgrid(np.arange(lz) + 0.5, np.arange(ly) + 0.5, np.arange(lx) + 0.5, indexing="ij")
    Xw = 5.5 + q
    coef = {}
    for i in PLUS_X:
        a, b = 1e-2 * rng.standard_normal(2)
        g = 1e-3 * rng.standard_normal(2)
        wall = 6.0 * W[i] * (CX[i] * vel[0] + CY[i] * vel[1] + CZ[i] * vel[2])
        P[i] = a + b * (X - Xw) + g[0] * Y + g[1] * Z
        bo = -b - 2.0 * (g[0] * CY[i] + g[1] * CZ[i])
        P[OPP[i]] = (a - wall) + bo * (X - Xw) + g[0] * Y + g[1] * Z
        coef[i] = (a, b, g, wall)
    return P.reshape(-1), coef, Xw


def _departure_value(coef, i, Xw, x):
    """The incoming population P_i at point x, less the moving-wall term."""
    a, b, g, wall = coef[i]
    return a + b * (x[0] - Xw) + g[0] * x[1] + g[1] * x[2] - wall


@pytest.mark.parametrize("q", [0.2, 0.35, 0.5, 0.65, 0.8])
@pytest.mark.parametrize("vel", [(0.0, 0.0, 0.0), (0.012, -0.03, 0.021)])
def test_bounce_back_reproduces_a_mirrored_linear_field(q, vel):
    """The value arriving at x_f in direction opp(i) after one step left the
    departure point x_f + (2q - 1) c_i (travel q to the wall, 1 - q back).
    Linear interpolation (q < 1/2: from x_f and x_f - c_i before streaming;
    q >= 1/2: from the reflected value and P_opp(x_f) after streaming) must
    reproduce it exactly when the populations are the mirrored linear field."""
    o, (lx, ly, lz, N) = _planar_wall(q, vel)
    P, coef, Xw = _mirrored_linear_field(q, vel, (lx, ly, lz, N), seed=int(q * 100) + 7)
    o.L[o.cur][:] = P
    o.beads_collision()
    out = o.post
    worst = 0.0
    margin = []
    for k in range(o.link_count[0]):
        n, i = int(o.link_node[k]), int(o.link_dir[k])
        ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
        xf = np.array([ix + 0.5, iy + 0.5, iz + 0.5])
        c = np.array([CX[i], CY[i], CZ[i]], dtype=float)
        want = _departure_value(coef, i, Xw, xf + (2 * q - 1) * c)
        ns = (ix + 1) + lx * ((iy + CY[i]) % ly + ly * ((iz + CZ[i]) % lz))
        got = out[OPP[i] * N + ns]
        worst = max(worst, abs(got - want))
        # the plausible mistakes this data separates from the method
        fi = P[i * N + n]
        if q < 0.5:
            fb = P[i * N + (ix - 1) + lx * ((iy - CY[i]) + ly * (iz - CZ[i]))]
            swapped = (2 * q) * fb + (1 - 2 * q) * fi - coef[i][3]
        else:
            fib = P[OPP[i] * N + n]
            swapped = (1 / (2 * q)) * fib + ((2 * q - 1) / (2 * q)) * fi - coef[i][3] / (2 * q)
        margin.append(abs(swapped - want))
    assert worst < 1e-15
    if q != 0.5:
        assert min(margin) > 1e-6


def test_bounce_back_fallback_is_halfway_bounce_back():
    """q < 1/2 with the second node x_f - c_i solid: no interpolation point, so
    the value is the halfway bounce-back P_i(x_f) - 6 w c.U (the q = 1/2 case
    of the departure-point rule); links whose second node is fluid keep the
    interpolation."""
    q, vel = 0.3, (0.01, 0.02, -0.015)
    behind = [(3, 3), (4, 2)]
    o, (lx, ly, lz, N) = _planar_wall(q, vel, behind_solid=behind)
    P, coef, Xw = _mirrored_linear_field(q, vel, (lx, ly, lz, N), seed=5)
    o.L[o.cur][:] = P
    o.beads_collision()
    nfall = 0
    for k in range(o.link_count[0]):
        n, i = int(o.link_node[k]), int(o.link_dir[k])
        ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
        xf = np.array([ix + 0.5, iy + 0.5, iz + 0.5])
        c = np.array([CX[i], CY[i], CZ[i]], dtype=float)
        fallback = (iy - CY[i], iz - CZ[i]) in behind
        nfall += fallback
        want = _departure_value(coef, i, Xw, xf if fallback else xf + (2 * q - 1) * c)
        ns = (ix + 1) + lx * ((iy + CY[i]) % ly + ly * ((iz + CZ[i]) % lz))
        assert abs(o.post[OPP[i] * N + ns] - want) < 1e-15
    assert nfall == 2 * len(PLUS_X)


def _uniform_flow(lx, ly, lz, rho, U):
    """Equilibrium populations (initpop form, Channel-Flow/initial.f90:29-44) of a uniform state."""
    cu = CX * U[0] + CY * U[1] + CZ * U[2]
    usq = 1.5 * (U[0] ** 2 + U[1] ** 2 + U[2] ** 2)
    feq = W * (rho + 3.0 * cu + 4.5 * cu * cu - usq)
    return np.repeat(feq, lx * ly * lz), feq


def test_sphere_moving_with_a_uniform_flow_feels_no_force():
    """Galilean invariance (Wen et al. 2014): a sphere translating with the
    velocity of a uniform flow (BGK rates, no body force) exchanges no net
    momentum, so it keeps its velocity and the flow stays uniform, over steps
    that cover and uncover nodes (refill) and use every bounce-back branch.
    The rotation is held fixed (infinite inertia): the moving-wall part of
    Wen's exchange leaves a small torque of lattice origin.  The channel walls
    (at rest, halfway bounce-back) disturb the flow by at most one node per
    step, so the channel is wide enough that no disturbance reaches the
    sphere or the checked band in the 30 steps."""
    lx, ly, lz = 96, 18, 16
    U = np.array([0.02, 0.05, -0.03])
    L0, feq = _uniform_flow(lx, ly, lz, 0.003, U)
    p = O.para(lx, ly, lz, mrttype=2)
    g = O.particle_geometry(lx, ly, lz, 1, 3.7, pressure_force=0)
    g.inertia = 1e300
    pos = np.array([[48.3, 8.1, 7.6]])
    o = O.OracleParticleChannel(p, g, L0, pos, np.array([U]))
    N = lx * ly * lz
    band = np.zeros((lz, ly, lx), dtype=bool)
    band[:, :, 31:65] = True  # more than 30 nodes from both walls
    band = band.ravel()
    branches = set()
    for step in range(30):
        q = o.link_q[:o.link_count[0]]
        branches |= {bool(v) for v in (q < 0.5)}
        o.step(1)
        assert np.max(np.abs(o.FH)) < 1e-13, step
        assert np.max(np.abs(o.V - U)) < 1e-14, step
        fluid = (o.solid < 0) & band
        dev = np.abs(o.post.reshape(19, N)[:, fluid] - feq[:, None]).max()
        assert dev < 1e-15, step
    assert branches == {True, False}
    assert abs(o.X[1] - (pos[0, 1] + 30 * U[1])) < 1e-12  # moved 1.5 nodes: covered and uncovered nodes


# --- contact, lubrication, integration -----------------------------------------

@pytest.mark.parametrize("d, want", [
    (11.5, 0.0),                    # R = 4: gap d - 8 = 3.5 > mingap 3, no force
    (10.5, 0.025 / 36.0),           # gap 2.5: 0.025 (0.5/3)^2
    (10.0, 0.025 / 9.0),            # gap 2: 0.025 ((3 - 2)/3)^2
    (8.5, 0.025 * 25.0 / 36.0),     # gap 0.5: 0.025 (2.5/3)^2
    (7.7, 0.025 * 1.21 + 0.002 * 0.1),  # overlap 0.3: 0.025 (3.3/3)^2 + 0.002 (0.3/3)
])
def test_contact_force_feng_michaelides(d, want):
    """Feng & Michaelides (2005) eq. 28 with stf0 = 0.025, stf1 = 0.002,
    mingap = 3 (Channel-Flow/para.f90:357-360, var_inc.f90:67), R = 4:
    hand-evaluated magnitudes along the line of centres, equal and opposite."""
    lx, ly, lz = 40, 40, 40
    n = np.array([0.6, 0.0, 0.8])
    pos = np.array([[20.0, 20.0, 20.0], [20.0, 20.0, 20.0]])
    pos[1] = pos[0] - d * n
    o = make(lx, ly, lz, pos, 4.0, pressure_force=0)
    o.beads_move()
    assert np.allclose(o.FC[0:3], want * n, rtol=1e-14, atol=1e-17)
    assert np.array_equal(o.FC[3:6], -o.FC[0:3])


@pytest.mark.parametrize("x, want", [
    (5.6, 0.0),                                 # gap to the image 3.2 > 3
    (4.6, 0.025 * 0.36),                        # image gap 1.2: 0.025 (1.8/3)^2
    (3.9, 0.025 * (3.2 / 3.0) ** 2 + 0.002 * (0.2 / 3.0)),  # image overlap 0.2
])
def test_wall_contact_uses_the_mirror_image(x, want):
    """Walls act as mirror-image particles (Feng & Michaelides 2005): the gap
    is 2 (distance to the wall) - 2R, with stf0_w = 0.025, stf1_w = 0.002."""
    lx, ly, lz = 30, 30, 30
    pos = np.array([[x, 15.0, 15.0], [lx - x, 5.0, 5.0]])
    o = make(lx, ly, lz, pos, 4.0, pressure_force=0)
    o.beads_move()
    assert abs(o.FC[0] - want) < 1e-16
    assert abs(o.FC[3] + want) < 1e-16
    assert np.all(o.FC[[1, 2, 4, 5]] == 0.0)


def test_lubrication_sphere_wall_approach():
    """Nguyen & Ladd (2002): below the cut-off gap hc the normal force on a
    sphere approaching a plane wall at speed U is -6 pi mu R^2 (1/h - 1/hc) U
    (mu = nu, rho0 = 1).  R = 4, nu = 0.0036, hc = 2/3, gap h = 0.25,
    U_x = -0.01: F_x = 6 pi 0.0036 16 (4 - 1.5) 0.01 = 0.0271433605..."""
    lx, ly, lz = 30, 30, 30
    pos = np.array([[4.25, 15.0, 15.0], [lx - 4.25, 5.0, 5.0], [15.0, 25.0, 25.0]])
    vel = np.array([[-0.01, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.0, 0.0]])
    o = make(lx, ly, lz, pos, 4.0, vel=vel, pressure_force=0, visc=0.0036, mingap_w=0.1)
    o.beads_move()
    want = 0.027143360527015815
    assert abs(o.FL[0] - want) < 1e-16
    assert abs(o.FL[3] + want) < 1e-16  # upper wall, receding mirror of the same motion
    assert np.all(o.FL[[1, 2, 4, 5, 6, 7, 8]] == 0.0)
    assert np.all(o.FC == 0.0)


def test_lubrication_pair_and_cutoffs():
    """Two equal spheres (a = R/2) approaching along the line of centres:
    -6 pi mu (R/2)^2 (1/h - 1/hc) U_n on each, opposite; zero beyond hc; the
    gap saturates at lub_min.  R = 4, nu = 0.0036, closing speed 0.02."""
    lx, ly, lz = 40, 40, 40
    base = np.array([20.0, 20.0, 20.0])
    n = np.array([0.0, 0.6, 0.8])
    vel = np.array([-0.01 * n, 0.01 * n])
    cases = [(0.4, 6 * np.pi * 0.0036 * 4.0 * (2.5 - 1.5) * 0.02),  # h = 0.4
             (0.7, 0.0),                                          # beyond hc = 2/3
             (0.001, 6 * np.pi * 0.0036 * 4.0 * (100.0 - 1.5) * 0.02)]  # saturated at lub_min 0.01
    for h, want in cases:
        pos = np.array([base, base - (8.0 + h) * n])
        o = make(lx, ly, lz, pos, 4.0, vel=vel, pressure_force=0, visc=0.0036, mingap=1e-9)
        o.beads_move()
        assert np.allclose(o.FL[0:3], want * n, rtol=1e-13, atol=1e-18), h
        assert np.array_equal(o.FL[3:6], -o.FL[0:3])


def test_force_average_removes_an_alternating_force():
    """The hydrodynamic force enters the velocity update averaged over the
    last two steps: a force that alternates +F, -F moves the particle by the
    first step's Euler kick a = F/m only (v_1 = a, v_2 = v_1 - a/2 from the
    AB2 history, then v_n = a/2); plain AB2 would oscillate."""
    lx, ly, lz = 40, 40, 40
    pos = np.array([[20.0, 20.0, 20.0]])
    o = make(lx, ly, lz, pos, 3.0, pressure_force=0)
    F = np.array([2e-3, -1e-3, 5e-4])
    a = F / o.g.mass
    vs = []
    for n in range(8):
        o.FH[:] = F if n % 2 == 0 else -F
        o.TQ[:] = 0.0
        o.beads_move()
        vs.append(o.V.copy())
    assert np.allclose(vs[0], a, rtol=1e-15, atol=0)
    for v in vs[1:]:
        assert np.allclose(v, 0.5 * a, rtol=1e-14, atol=0)


# --- reductions and fields -----------------------------------------------------

def test_device_order_sum_is_exact_on_exact_data():
    """orc_device_sum restates the device's block/wave/lane order.  On data
    whose partial sums are all exact (small integers) every order gives the
    true sum; on random data it agrees with a correctly rounded sum to a few
    ulps; the mask drops exactly the masked terms."""
    import ctypes
    L = O._plib()
    rng = np.random.default_rng(3)
    for n in (1, 255, 256, 1000, 70_000, 600_000):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        m = (rng.random(n) < 0.7).astype(np.uint8)
        got = L.orc_device_sum(v.ctypes.data_as(ctypes.c_void_p), n, m.ctypes.data_as(ctypes.c_void_p))
        assert got == float(v[m.astype(bool)].sum())
        r = rng.standard_normal(n)
        got = L.orc_device_sum(r.ctypes.data_as(ctypes.c_void_p), n, None)
        import math
        assert abs(got - math.fsum(r)) <= 1e-13 * max(1.0, np.abs(r).sum())


def test_particle_macrovar_and_rho_mean():
    """Fields of the particle-laden channel: fluid nodes equal the pulled
    populations' sums (np.roll streaming with halfway bounce-back at the
    walls), nodes inside a particle carry V + W x r (np.cross) and rhopart,
    and avedensity's mean is the fluid-masked mean density."""
    lx, ly, lz = 20, 16, 14
    pos = np.array([[9.7, 7.2, 13.1]])  # wraps across z
    vel = np.array([[0.01, -0.02, 0.03]])
    omg = np.array([[0.002, 0.004, -0.003]])
    rng = np.random.default_rng(9)
    N = lx * ly * lz
    L0 = 1e-3 * rng.standard_normal(19 * N)
    o = make(lx, ly, lz, pos, 4.0, vel=vel, omg=omg, L0=L0, force=(0.0, 1e-5, 0.0), rhopart=1.3)
    rho, ux, uy, uz = o.macrovar()
    P = L0.reshape(19, lz, ly, lx)
    F = np.empty_like(P)
    for i in range(19):
        s = np.roll(P[i], (CZ[i], CY[i], CX[i]), axis=(0, 1, 2))
        # halfway bounce-back: what would come from beyond a wall is the node's own opposite value
        if CX[i] == 1:
            s[:, :, 0] = P[OPP[i]][:, :, 0]
        elif CX[i] == -1:
            s[:, :, -1] = P[OPP[i]][:, :, -1]
        F[i] = s
    fluid = (o.solid < 0).reshape(lz, ly, lx)
    assert np.allclose(rho[fluid], F.sum(axis=0)[fluid], rtol=0, atol=1e-16)
    assert np.allclose(uy[fluid], (CY[:, None, None, None] * F).sum(axis=0)[fluid] + 0.5e-5, rtol=0, atol=1e-16)
    Z, Y, X = np.meshgrid(np.arange(lz) + 0.5, np.arange(ly) + 0.5, np.arange(lx) + 0.5, indexing="ij")
    r = np.stack([X - pos[0, 0], Y - pos[0, 1], Z - pos[0, 2]], -1)
    r[..., 1] -= ly * np.round(r[..., 1] / ly)
    r[..., 2] -= lz * np.round(r[..., 2] / lz)
    u = vel[0] + np.cross(omg[0], r)
    sol = ~fluid
    assert sol.sum() > 200
    for k, a in enumerate((ux, uy, uz)):
        assert np.allclose(a[sol], u[..., k][sol], rtol=0, atol=1e-17)
    assert np.all(rho[sol] == 1.3)
    mean = o.rho_mean()
    assert abs(mean - F.sum(axis=0)[fluid].mean()) < 1e-17
