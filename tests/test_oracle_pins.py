"""Pins of the CPU oracle (oracle/lbm_oracle.c) against what the mathematics fixes.

The reference (Fortran + MPI) cannot be built in this image, so the oracle is
checked against independent formulations of every piece of the path:

* the per-node MRT collision against the matrix form m* = m - S(m - m_eq),
  f* = M^-1 m* + Fbar/2, with the moment matrix M built here from the
  polynomial definitions of d'Humieres et al. (2002), Phil. Trans. R. Soc. A
  360:437, eq. (A1)-(A2), and inverted with numpy.linalg.inv;
* MRTtype 2 (para.f90:121-130) against textbook single-relaxation BGK with
  Guo forcing, f* = f9 - (f9 - f_eq)/tau + Fbar/2;
* the in-place swap streaming plus single-rank exchange
  (collision.f90:224-372) against push streaming written with np.roll and
  halfway bounce-back;
* macrovar/rhoupdat/initpop against their defining sums;
* invariants: mass, y/z translation, x-reflection;
* the laminar start-up against the analytic Poiseuille series the reference
  itself prints (saveload.f90:920-936), reproducing the errors measured on the
  compiled reference.
"""
import math

import numpy as np
import pytest

import oracle as O

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def moment_matrix():
    """Rows in d'Humieres (2002) order: rho, e, eps, jx, qx, jy, qy, jz, qz,
    3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz."""
    c2 = CX**2 + CY**2 + CZ**2
    rows = [np.ones(19), 19 * c2 - 30, (21 * c2**2 - 53 * c2 + 24) / 2,
            CX, (5 * c2 - 9) * CX, CY, (5 * c2 - 9) * CY, CZ, (5 * c2 - 9) * CZ,
            3 * CX**2 - c2, (3 * c2 - 5) * (3 * CX**2 - c2),
            CY**2 - CZ**2, (3 * c2 - 5) * (CY**2 - CZ**2),
            CX * CY, CY * CZ, CX * CZ,
            (CY**2 - CZ**2) * CX, (CZ**2 - CX**2) * CY, (CX**2 - CY**2) * CZ]
    return np.array(rows, dtype=np.float64)


def guo_fbar(u, F):
    """Fbar_i = w_i (3 c.F + 9 (c.F)(c.u) - 3 u.F), with Fbar_0 = -u.F (collision.f90:68-82)."""
    c = np.stack([CX, CY, CZ], 1).astype(float)
    cF, cu, uF = c @ F, c @ u, u @ F
    fb = np.where(np.arange(19) <= 6, (1 / 18) * (3 * cF + 9 * cF * cu - 3 * uF),
                  (1 / 36) * (3 * cF + 9 * cF * cu - 3 * uF))
    fb[0] = -uF
    return fb


def matrix_collide(p, fin, rho, u, F):
    M = moment_matrix()
    fb = guo_fbar(u, F)
    f9 = fin + 0.5 * fb
    m = M @ f9
    ux, uy, uz = u
    t1 = ux * ux + uy * uy + uz * uz
    meq = np.zeros(19)
    meq[0], meq[3], meq[5], meq[7] = rho, ux, uy, uz
    meq[1] = -11 * rho + 19 * t1
    meq[2] = p.omegepsl * rho + p.omegepslj * t1
    meq[4], meq[6], meq[8] = -2 / 3 * ux, -2 / 3 * uy, -2 / 3 * uz
    meq[9] = 2 * ux * ux - uy * uy - uz * uz
    meq[10] = p.omegxx * meq[9]
    meq[11] = uy * uy - uz * uz
    meq[12] = p.omegxx * meq[11]
    meq[13], meq[14], meq[15] = ux * uy, uy * uz, ux * uz
    s = np.array([0, p.s1, p.s2, 0, p.s4, 0, p.s4, 0, p.s4, p.s9, p.s10, p.s9, p.s10,
                  p.s13, p.s13, p.s13, p.s16, p.s16, p.s16])
    mstar = m - s * (m - meq)
    # conserved moments are taken from the supplied fields (collision.f90:157-166)
    mstar[0], mstar[3], mstar[5], mstar[7] = rho, ux, uy, uz
    return np.linalg.inv(M) @ mstar + 0.5 * fb


def test_moment_matrix_is_orthogonal_with_reference_norms():
    M = moment_matrix()
    G = M @ M.T
    assert np.allclose(G, np.diag(np.diag(G)))
    # para.f90:152-160: val1..val9 and the 1/coef4, 1/coef3 norms
    assert list(np.diag(G)[[0, 1, 2, 3, 4, 9, 10, 11, 12, 13, 16]]) == [19, 2394, 252, 10, 40, 36, 72, 12, 24, 4, 8]


@pytest.mark.parametrize("mrt", [1, 2, 3])
def test_collide_node_matches_matrix_form(mrt):
    p = O.para(32, 64, 32, laminar=False, mrttype=mrt)
    rng = np.random.default_rng(100 + mrt)
    worst = 0.0
    for _ in range(200):
        fin = rng.standard_normal(19) * 0.05
        rho = rng.standard_normal() * 0.01
        u = rng.standard_normal(3) * 0.1
        F = rng.standard_normal(3) * 1e-3
        got = O.collide_node(p, fin, rho, *u, *F)
        want = matrix_collide(p, fin, rho, u, F)
        worst = max(worst, float(np.max(np.abs(got - want))))
    assert worst < 1e-15, worst


def test_collide_node_matrix_form_detects_planted_error():
    """The matrix-form check is sharp enough to see a single wrong relaxation rate."""
    p = O.para(32, 64, 32, laminar=False, mrttype=1)
    q = O.para(32, 64, 32, laminar=False, mrttype=1)
    q.s10 = q.s10 * (1 + 1e-6)
    rng = np.random.default_rng(5)
    fin, rho, u, F = rng.standard_normal(19) * 0.05, 0.01, rng.standard_normal(3) * 0.1, np.zeros(3)
    got = O.collide_node(q, fin, rho, *u, *F)
    want = matrix_collide(p, fin, rho, u, F)
    assert np.max(np.abs(got - want)) > 1e-12


def test_mrt_type2_is_bgk_with_guo_forcing():
    p = O.para(32, 8, 8, laminar=True)  # MRTtype 2, para.f90:84
    assert p.mrttype == 2
    rng = np.random.default_rng(3)
    c = np.stack([CX, CY, CZ], 1).astype(float)
    M = moment_matrix()
    Minv = np.linalg.inv(M)
    for _ in range(100):
        # fields first, then populations whose f9 = f + Fbar/2 has exactly
        # those conserved moments (as macrovar guarantees in the reference)
        rho9 = rng.standard_normal() * 0.01
        u9 = rng.standard_normal(3) * 0.1
        F = rng.standard_normal(3) * 1e-3
        cu = c @ u9
        feq = W * (rho9 + 3 * cu + 4.5 * cu * cu - 1.5 * (u9 @ u9))
        neq = M @ (rng.standard_normal(19) * 0.01)
        neq[[0, 3, 5, 7]] = 0.0
        f9 = feq + Minv @ neq
        fb = guo_fbar(u9, F)
        fin = f9 - 0.5 * fb
        want = f9 - p.s9 * (f9 - feq) + 0.5 * fb
        got = O.collide_node(p, fin, rho9, *u9, *F)
        assert np.max(np.abs(got - want)) < 1e-15


def _random_state(o, seed, amp=0.05, noise=1e-3):
    """Seeded state: u = amp*(0.2 N, 1 + 0.2 N, 0.2 N), f = equilibrium plus
    noise*N.  At the turbulent tau = 0.511 such node-scale noise is only
    stable for a few tens of steps; long runs use noise=0 or laminar constants."""
    rng = np.random.default_rng(seed)
    o.FORCING()
    o.ux[:] = amp * 0.2 * rng.standard_normal(o.ux.shape)
    o.uy[:] = amp * (1 + 0.2 * rng.standard_normal(o.uy.shape))
    o.uz[:] = amp * 0.2 * rng.standard_normal(o.uz.shape)
    o.initpop()
    o.f += noise * rng.standard_normal(o.f.shape)
    o.macrovar()


def _numpy_push_step(o):
    """One collision_MRT: node collisions via orc_collide_node, then np.roll
    push streaming in y/z and halfway bounce-back in x."""
    lz, ly, lx = o.rho.shape
    post = np.empty_like(o.f)
    for iz in range(lz):
        for iy in range(ly):
            for ix in range(lx):
                post[iz, iy, ix] = O.collide_node(o.p, o.f[iz, iy, ix], o.rho[iz, iy, ix], o.ux[iz, iy, ix],
                                                  o.uy[iz, iy, ix], o.uz[iz, iy, ix], o.fx[iz, iy, ix],
                                                  o.fy[iz, iy, ix], o.fz[iz, iy, ix])
    out = np.full_like(o.f, np.nan)
    for i in range(19):
        a = post[..., i]
        a = np.roll(a, CY[i], axis=1)
        a = np.roll(a, CZ[i], axis=0)
        if CX[i] == 0:
            out[..., i] = a
        elif CX[i] == 1:
            out[:, :, 1:, i] = a[:, :, :-1]
            out[:, :, lx - 1, OPP[i]] = post[:, :, lx - 1, i]
        else:
            out[:, :, :-1, i] = a[:, :, 1:]
            out[:, :, 0, OPP[i]] = post[:, :, 0, i]
    return out


@pytest.mark.parametrize("shape", [(5, 4, 3), (4, 1, 3), (3, 5, 1), (2, 2, 2)])
def test_swap_streaming_equals_numpy_push_streaming(shape):
    lx, ly, lz = shape
    o = O.OracleChannel(lx, ly, lz)
    _random_state(o, 11)
    want = _numpy_push_step(o)
    assert not np.isnan(want).any()
    o.collision_MRT()
    assert np.array_equal(o.f.view(np.uint64), want.view(np.uint64))


def test_swap_and_ab_schedules_are_bitwise_equal():
    a = O.OracleChannel(9, 7, 5)
    _random_state(a, 12)
    b = O.OracleChannel(9, 7, 5)
    for k in ("f", "rho", "ux", "uy", "uz", "fx", "fy", "fz"):
        setattr(b, k, getattr(a, k).copy())
    a.run(30)
    b.run(30, ab=True, nthreads=4)
    assert np.array_equal(a.f.view(np.uint64), b.f.view(np.uint64))


def test_macrovar_rhoupdat_initpop_definitions():
    o = O.OracleChannel(6, 5, 4)
    _random_state(o, 13)
    o.macrovar()
    c = np.stack([CX, CY, CZ], 1).astype(float)
    assert np.allclose(o.rho, o.f.sum(-1), rtol=0, atol=1e-16)
    j = o.f @ c
    assert np.allclose(o.ux, j[..., 0] + o.fx / 2, rtol=0, atol=1e-16)
    assert np.allclose(o.uy, j[..., 1] + o.fy / 2, rtol=0, atol=1e-16)
    assert np.allclose(o.uz, j[..., 2] + o.fz / 2, rtol=0, atol=1e-16)
    rho = np.empty_like(o.rho)
    O.lib().orc_rhoupdat(O.ctypes.byref(o.p), O._ptr(o.f), O._ptr(rho))
    assert np.array_equal(rho, o.rho)  # same summation order as macrovar
    u = np.stack([o.ux, o.uy, o.uz], -1)
    o.initpop()
    cu = u @ c.T
    feq = W * (3 * cu + 4.5 * cu**2 - 1.5 * (u**2).sum(-1, keepdims=True))
    assert np.allclose(o.f, feq, rtol=0, atol=1e-17)
    assert np.all(o.rho == 0)


def test_mass_is_conserved():
    o = O.OracleChannel(8, 6, 5, laminar=True)
    _random_state(o, 14)
    m0 = o.f.sum()
    o.run(100)
    assert abs(o.f.sum() - m0) < 1e-13


def test_translation_in_y_and_z_commutes_with_a_step():
    a = O.OracleChannel(6, 5, 4)
    _random_state(a, 15)
    b = O.OracleChannel(6, 5, 4)
    for k in ("f", "rho", "ux", "uy", "uz", "fx", "fy", "fz"):
        setattr(b, k, np.ascontiguousarray(np.roll(np.roll(getattr(a, k), 2, axis=1), 3, axis=0)))
    a.run(10)
    b.run(10)
    want = np.roll(np.roll(a.f, 2, axis=1), 3, axis=0)
    assert np.array_equal(b.f.view(np.uint64), want.view(np.uint64))


def test_reflection_across_the_channel_centre():
    """x -> lx+1-x with u_x -> -u_x maps the solution onto itself (to round-off)."""
    a = O.OracleChannel(6, 5, 4)
    _random_state(a, 16)
    b = O.OracleChannel(6, 5, 4)
    perm = np.array([int(np.flatnonzero((CX == -CX[i]) & (CY == CY[i]) & (CZ == CZ[i]))[0]) for i in range(19)])
    b.f = np.ascontiguousarray(a.f[:, :, ::-1, :][..., perm])
    for k in ("rho", "uy", "uz", "fx", "fy", "fz"):
        setattr(b, k, np.ascontiguousarray(getattr(a, k)[:, :, ::-1]))
    b.ux = np.ascontiguousarray(-a.ux[:, :, ::-1])
    b.fx = np.ascontiguousarray(-b.fx)
    a.run(20)
    b.run(20)
    want = a.f[:, :, ::-1, :][..., perm]
    assert np.max(np.abs(b.f - want)) < 1e-15


def poiseuille_error(nx, ny, nz, steps_out):
    """Laminar start-up of para.f90:74-88 (LBGK, fluid at rest), driven as
    main.f90 drives it (initvel, FORCING, initpop, one pre-relaxation
    iteration, macrovar, main loop); returns max|u/u* - analytic| / max(analytic)
    over the profile written by outputuy (saveload.f90:920-936) at each step."""
    o = O.OracleChannel(nx, ny, nz, laminar=True)
    p = o.p
    o.FORCING()
    o.initpop()
    it, err = o.prerelax()
    assert it == 0
    o.macrovar()
    out = {}
    step = 0
    for target in steps_out:
        o.run(target - step)
        step = target
        time1 = step * p.visc / ((nx / 2.0) ** 2)
        errs, amax = [], 0.0
        for i in range(1, nx // 2 + 1):
            xx0 = abs(i - 0.5 - nx / 2.0)
            uut = 0.0
            for n in range(26):
                uut += (4 * (-1) ** n) / (math.pi * (n + 0.5)) ** 3 * math.exp(
                    -((n + 0.5) ** 2) * math.pi**2 * time1) * math.cos((n + 0.5) * math.pi * (xx0 / (nx / 2.0)))
            uut = (1 - (xx0 / (nx / 2.0)) ** 2) - uut
            sim = o.uy[nz // 2 - 1, :, i - 1] / p.ustar
            errs.append(np.max(np.abs(sim - uut)))
            amax = max(amax, abs(uut))
        out[step] = max(errs) / amax
    return out


def test_poiseuille_startup_matches_reference_run():
    """Errors measured on the compiled reference at nx=32, 8x8 periodic
    (t* = 0.0625, 0.1875, 0.625): 9.95e-3, 3.01e-3, 6.73e-4."""
    e = poiseuille_error(32, 8, 8, [100, 300, 1000])
    assert round(e[100], 5) == pytest.approx(9.95e-3, abs=5e-6)
    assert round(e[300], 5) == pytest.approx(3.01e-3, abs=5e-6)
    assert round(e[1000], 6) == pytest.approx(6.73e-4, abs=5e-7)


def _trig_fields(lx, ly, lz, seed):
    """Fields whose plane moments are known in closed form: with
    t_j = 2 pi j / ly and p_k = 2 pi k / lz, the plane means of cos t_j,
    sin p_k and their products vanish and the means of cos^2, sin^2 are 1/2."""
    rng = np.random.default_rng(seed)
    A, B, C, D, E, F, R, G = (rng.uniform(-1e-2, 1e-2, lx) for _ in range(8))
    t = 2 * np.pi * np.arange(ly) / ly
    p = 2 * np.pi * np.arange(lz) / lz
    ct = np.cos(t)[None, :, None]
    sp = np.sin(p)[:, None, None]
    shape = (lz, ly, lx)
    ux = np.broadcast_to(A + B * ct, shape).copy()
    uy = np.broadcast_to(C + D * sp, shape).copy()
    uz = np.broadcast_to(E + F * ct, shape).copy()
    rho = np.broadcast_to(R + G * ct, shape).copy()
    return (A, B, C, D, E, F, R, G), (rho, ux, uy, uz)


def test_statistc2_restatement_matches_closed_forms():
    """saveload.f90:1382-1470: plane means, variances, covariances, pressure
    moments and the y+ column, on fields whose moments are known exactly."""
    lx, ly, lz = 10, 6, 8
    o = O.OracleChannel(lx, ly, lz)
    (A, B, C, D, E, F, R, G), (rho, ux, uy, uz) = _trig_fields(lx, ly, lz, 5)
    o.rho[:], o.ux[:], o.uy[:], o.uz[:] = rho, ux, uy, uz
    out = o.statistc2()
    us, ys = o.p.ustar, o.p.ystar
    i = np.arange(1, lx + 1)
    yplus = np.where(i > lx // 2, (lx - i + 0.5) / ys, (i - 0.5) / ys)
    want = np.stack([i - 0.5, yplus, A / us, C / us, E / us, B * B / 2 / us**2, D * D / 2 / us**2,
                     F * F / 2 / us**2, B * F / 2 / us**2, 0 * A, 0 * A, R / us**2, G * G / 2 / us**4,
                     np.ones(lx), np.zeros(lx)], axis=1)
    scale = np.abs(want).max(axis=0)
    scale[[9, 10]] = np.abs(want[:, 5:9]).max()  # vanishing covariances: the second moments' size
    scale[14] = 1.0
    assert np.all(np.abs(out - want) <= 1e-11 * scale), np.abs(out - want).max(axis=0) / scale


def test_diag_restatement_matches_closed_forms():
    """saveload.f90:1535-1660: volume means and rms of the velocity, the
    largest speed and its node, and the density extrema."""
    lx, ly, lz = 10, 6, 8
    o = O.OracleChannel(lx, ly, lz)
    (A, B, C, D, E, F, R, G), (rho, ux, uy, uz) = _trig_fields(lx, ly, lz, 6)
    ux[3, 2, 7] += 0.5  # one fast node, well above every other speed
    o.rho[:], o.ux[:], o.uy[:], o.uz[:] = rho, ux, uy, uz
    out = o.diag()
    us = o.p.ustar
    n = lx * ly * lz
    um = (ux.sum()) / n
    ur = np.sqrt((ux**2).sum() / n - um**2)
    vm, wm = C.mean(), E.mean()
    vr = np.sqrt((C**2 + D**2 / 2).mean() - vm**2)
    wr = np.sqrt((E**2 + F**2 / 2).mean() - wm**2)
    speed = np.sqrt(ux**2 + uy**2 + uz**2)
    assert out[0] == pytest.approx(speed[3, 2, 7], rel=1e-15)
    assert tuple(out[1:4]) == (8, 3, 4)  # 1-based (i, j, k) of node (ix=7, iy=2, iz=3)
    np.testing.assert_allclose(out[4:10], [um / us, vm / us, wm / us, ur / us, vr / us, wr / us], rtol=1e-10)
    assert out[10] == 0.0
    assert (out[11], out[12]) == (rho.max(), rho.min())


def test_vortcalc_restatement_matches_closed_forms():
    """saveload.f90:3929-4023 on ux = b sin(2 pi z/lz), uy = c x (L - x),
    uz = a sin(2 pi y/ly) (x = i - 1/2 from the lower wall, y = j - 1/2,
    z = k - 1/2): the central differences of a sine are cos times
    sin(2 pi / L) exactly; the wall forms (3 u_1 + u_2)/3 and
    -(3 u_lx + u_lx-1)/3 are the exact derivative of a parabola that vanishes
    at the walls, and give 4u/3 and -4u/3 for a field constant in x.
    Nodes inside a particle carry twice its angular velocity."""
    lx, ly, lz = 12, 10, 8
    a, b, c = 0.013, -0.021, 0.0007
    o = O.OracleChannel(lx, ly, lz)
    x = np.arange(lx) + 0.5
    y = np.arange(ly) + 0.5
    z = np.arange(lz) + 0.5
    shape = (lz, ly, lx)
    o.ux[:] = np.broadcast_to(b * np.sin(2 * np.pi * z / lz)[:, None, None], shape)
    o.uy[:] = np.broadcast_to(c * x * (lx - x), shape)
    o.uz[:] = np.broadcast_to(a * np.sin(2 * np.pi * y / ly)[None, :, None], shape)
    solid = np.full(shape, -1, dtype=np.int32)
    solid[3, 4, 5] = 1
    omg = np.array([[9.0, 9.0, 9.0], [0.01, -0.02, 0.03]])
    ox, oy, oz = o.vortcalc(solid=solid, omg=omg)
    pwy = a * np.cos(2 * np.pi * y / ly)[None, :, None] * np.sin(2 * np.pi / ly)
    puz = b * np.cos(2 * np.pi * z / lz)[:, None, None] * np.sin(2 * np.pi / lz)
    uz = np.broadcast_to(o.uz, shape)
    pwx = np.zeros(shape)
    pwx[:, :, 0] = 4.0 * uz[:, :, 0] / 3.0
    pwx[:, :, -1] = -4.0 * uz[:, :, -1] / 3.0
    pvx = np.broadcast_to(c * (lx - 2 * x), shape)
    want = [np.broadcast_to(pwy, shape).copy(), np.broadcast_to(puz, shape) - pwx, pvx.copy()]
    for w, k in zip(want, range(3)):
        w[3, 4, 5] = 2.0 * omg[1, k]
    for got, w in zip((ox, oy, oz), want):
        assert np.abs(got - w).max() < 1e-15


def _strained_state(lx, ly, lz, seed):
    """f = equilibrium (initpop form) + w_i (9/2) Q : (c_i c_i - I/3) with a
    symmetric traceless Q that grows across the channel, Q(x) = Q0 (1 + x/lx):
    the non-equilibrium second moment is exactly Q, its trace and the
    non-equilibrium energy and density vanish."""
    rng = np.random.default_rng(seed)
    Q0 = rng.uniform(-1e-4, 1e-4, (3, 3))
    Q0 = 0.5 * (Q0 + Q0.T)
    Q0 -= np.trace(Q0) / 3 * np.eye(3)
    c = np.stack([CX, CY, CZ], 1).astype(float)
    rho = rng.uniform(-1e-3, 1e-3, (lz, ly, lx))
    u = rng.uniform(-0.03, 0.03, (3, lz, ly, lx))
    cu = np.einsum("ia,azyx->zyxi", c, u)
    usq = 1.5 * (u ** 2).sum(0)
    feq = W * (rho[..., None] + 3 * cu + 4.5 * cu * cu - usq[..., None])
    ccq = np.einsum("ia,ab,ib->i", c, Q0, c) - np.trace(Q0) / 3 * (c * c).sum(1)
    scale = 1.0 + (np.arange(lx) + 0.5) / lx
    fneq = W * 4.5 * ccq * scale[None, None, :, None]
    return feq + fneq, rho, u, Q0, scale


def test_dissipation_restatement_matches_closed_forms():
    """saveload.f90:4644-4882 on a state whose non-equilibrium stress is a
    known traceless Q: the moment method gives S = -(3 s9 / 2) Q and the
    population method S = (3 / (2 tau)) Q, so with s9 = 1/tau both give
    S_ij S_ij = (1.5/tau)^2 Q:Q and S_xy of opposite signs; the plane means
    follow Q's growth across the channel; x and y+ columns as statistc2's."""
    lx, ly, lz = 10, 6, 4
    o = O.OracleChannel(lx, ly, lz)
    f, rho, u, Q0, scale = _strained_state(lx, ly, lz, 3)
    o.f[:] = f
    o.rho[:], o.ux[:], o.uy[:], o.uz[:] = rho, u[0], u[1], u[2]
    out = o.dissipation()
    k = 1.5 / o.p.tau
    ss = k * k * (Q0 * Q0).sum() * scale ** 2
    sxy = k * Q0[0, 1] * scale
    i = np.arange(1, lx + 1)
    ys = o.p.ystar
    assert np.array_equal(out[:, 0], i - 0.5)
    np.testing.assert_allclose(out[:, 1], np.where(i > lx // 2, (lx - i + 0.5) / ys, (i - 0.5) / ys), rtol=1e-15)
    for col, want in ((2, ss), (3, ss), (4, -sxy), (5, sxy)):
        np.testing.assert_allclose(out[:, col], want, rtol=1e-9, atol=0, err_msg=f"column {col}")
