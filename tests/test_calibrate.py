"""The momentum-balance targets of the learned closure (d3q19/calibrate.py)
against closed forms: laminar channel profiles whose plane gradients and
viscosity ratios are known exactly."""
import numpy as np
import pytest

from d3q19 import calibrate as C


def _centres(nx):
    return np.arange(nx) + 0.5


def test_plane_gradient_is_exact_for_quadratics_through_the_walls():
    nx = 24
    x = _centres(nx)
    U = x * (nx - x)                        # zero at both halfway walls
    np.testing.assert_allclose(C.plane_gradient(U), nx - 2 * x, rtol=0, atol=1e-12)
    # a cubic with U(0) = 0 is not exact at the wall plane: the one-sided form
    # is second order, error U'''/6 * (1/2 * 3/2) ... nonzero
    V = x ** 3
    assert abs(C.plane_gradient(V)[0] - 3 * 0.25) > 1e-3


def _laminar(nx, g, F=1e-5, nu=0.05, n=4000):
    """Plane-centre velocities of the laminar channel with viscosity nu (1 + g(x)):
    nu (1 + g) U' = F (h - x), U(0) = 0 (composite Simpson on n intervals per cell)."""
    h = nx / 2.0
    out = np.empty(nx)
    for i, xi in enumerate(_centres(nx)):
        s = np.linspace(0.0, xi, 2 * n + 1)
        f = F * (h - s) / (nu * (1.0 + g(s)))
        out[i] = (s[1] - s[0]) / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    return out


def test_balance_update_recovers_a_prescribed_viscosity_profile():
    """A run with uniform viscosity against the profile of nu (1 + g*(x)): one
    update gives g* at the planes where the gradient is resolved, to the
    second-order error of the plane gradients."""
    nx = 64
    gstar = lambda s: -0.3 + 0.6 * np.sin(np.pi * s / nx) ** 4  # noqa: E731
    U_ref = _laminar(nx, gstar)
    U_run = _laminar(nx, lambda s: 0.0 * s)
    g = C.balance_update(U_run, U_ref, 0.0, core_frac=0.05)
    x = _centres(nx)
    d = C.plane_gradient(U_ref)
    use = np.abs(d) > 0.05 * abs(d[0])
    err = np.abs(g - gstar(x))[use]
    assert err.max() < 5e-3, err.max()
    # the core holds the value of the nearest resolved plane
    core = np.flatnonzero(~use)
    assert core.size and np.all(np.diff(core) == 1)
    lo, hi = core[0] - 1, core[-1] + 1
    half = core.size // 2
    np.testing.assert_allclose(g[core[:half]], g[lo], rtol=0, atol=1e-15)
    np.testing.assert_allclose(g[core[core.size - half:]], g[hi], rtol=0, atol=1e-15)


def test_balance_update_from_a_closure_run():
    """A run made with closure g_k: the update is multiplicative in 1 + g."""
    nx = 48
    gk = lambda s: 0.2 + 0.0 * s  # noqa: E731
    gstar = lambda s: 0.5 * np.cos(np.pi * (s - nx / 2) / nx) ** 2 - 0.2  # noqa: E731
    U_ref = _laminar(nx, gstar)
    U_run = _laminar(nx, gk)
    g = C.balance_update(U_run, U_ref, np.full(nx, 0.2), core_frac=0.05)
    d = C.plane_gradient(U_ref)
    use = np.abs(d) > 0.05 * abs(d[0])
    assert np.abs(g - gstar(_centres(nx)))[use].max() < 5e-3


def test_balance_update_is_a_fixed_point_at_the_reference():
    nx = 32
    U = _laminar(nx, lambda s: 0.1 * np.cos(2 * np.pi * s / nx))
    gk = np.linspace(-0.2, 0.3, nx)
    g = C.balance_update(U, U, gk)
    np.testing.assert_allclose(g, np.clip(C.symmetrize(gk), -0.5, 1.0), rtol=0, atol=1e-12)


def test_balance_update_keeps_the_run_where_nothing_is_resolved():
    nx = 8
    U = np.ones(nx)
    gk = np.full(nx, 0.25)
    np.testing.assert_allclose(C.balance_update(U, U, gk, core_frac=1e9), gk)


def test_balance_update_clips_and_checks_shapes():
    nx = 16
    U_ref = _laminar(nx, lambda s: 0.0 * s)
    g = C.balance_update(10.0 * U_ref, U_ref, 0.0, hi=1.0)
    assert g.max() == 1.0
    with pytest.raises(ValueError):
        C.balance_update(U_ref[:-1], U_ref, 0.0)


def test_wall_profile_levels_and_symmetry():
    """The piecewise wall-distance profile of the closure studies: its levels
    far from the splits, mirror symmetry about the centre plane, and the
    calibrated profile's values at the walls and in the core."""
    from d3q19 import studies
    lxc, ystar = 96, 0.26666666666666666  # Re_tau = 180 on the 2x-coarse grid
    g = studies.wall_profile((-0.03, -0.08, -0.15), lxc, ystar, splits=(10.0, 30.0), width=5.0)
    np.testing.assert_allclose(g, g[::-1], rtol=0, atol=1e-15)
    yplus = (np.arange(lxc) + 0.5) / ystar
    assert abs(g[0] + 0.03) < 2e-3                      # y+ = 1.9
    assert abs(g[np.argmin(abs(yplus - 20))] + 0.08) < 2e-3
    assert abs(g[lxc // 2 - 1] + 0.15) < 1e-9           # core
    assert np.all(np.diff(g[:lxc // 2]) <= 0)           # monotone towards the centre
    same = studies.wall_profile(lxc=lxc, ystar_c=ystar, **studies.WALL_CLOSURE_C5)
    np.testing.assert_array_equal(g, same)
    with pytest.raises(ValueError):
        studies.wall_profile((0.0, 0.1), lxc, ystar, splits=(10.0, 30.0))
