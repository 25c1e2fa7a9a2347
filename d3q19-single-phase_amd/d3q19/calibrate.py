"""Wall-normal targets for the learned closure from the mean momentum balance.

A channel at statistical equilibrium carries the total shear stress
tau(x) = F (h - x) (driving force F, half width h) as viscous plus Reynolds
stress:

    nu (1 + g(x)) dU/dx - <u'_x u'_y> = tau(x),

with g the plane-mean eddy viscosity of the closure in units of nu.  A coarse
run with closure g_k produces the profile U_k.  The closure g for which the
same balance holds with the reference profile U_ref (the box-filtered fine
statistics), at the Reynolds stress of that run, is

    1 + g = (1 + g_k) U_k' / U_ref'.

So each coarse run yields a better per-plane target; the network is trained
on it (lbm_trainer_set_plane_targets) and the next coarse run tests it.  In
the core dU/dx vanishes and the ratio carries no information (the viscous
stress is a vanishing share of tau there): planes where |U_ref'| is below
`core_frac` of its wall value take the value of the nearest plane where it is
not, so the target stays continuous.

The gradients are central differences on the plane centres x_i = i + 1/2
(cell units); at the planes next to the halfway walls (x = 0 and x = nx,
where U = 0) the second-order one-sided form through the wall value,
U'(1/2) = U_0 + U_1 / 3, and its mirror image.
"""
import numpy as np


def plane_gradient(U):
    """dU/dx at the plane centres of a wall-bounded profile U (nx planes,
    halfway walls at both ends where U = 0)."""
    U = np.asarray(U, dtype=np.float64)
    nx = U.size
    if nx < 3:
        raise ValueError("need at least three planes")
    g = np.empty(nx)
    g[1:-1] = 0.5 * (U[2:] - U[:-2])
    g[0] = U[0] + U[1] / 3.0
    g[-1] = -(U[-1] + U[-2] / 3.0)
    return g


def symmetrize(g):
    """The mean of a plane profile and its mirror image (x -> nx - 1 - x)."""
    g = np.asarray(g, dtype=np.float64)
    return 0.5 * (g + g[::-1])


def balance_update(U_run, U_ref, g_run, core_frac=0.05, lo=-0.5, hi=1.0):
    """Next per-plane closure target g from a run's mean profile U_run (made
    with the plane closure g_run) and the reference U_ref: the balance
    relation above where the reference gradient is resolved, held at the
    nearest resolved plane's value elsewhere (g_run if no plane is resolved);
    symmetrized about the centre plane and clipped to [lo, hi]."""
    U_run = np.asarray(U_run, dtype=np.float64)
    U_ref = np.asarray(U_ref, dtype=np.float64)
    g_run = np.broadcast_to(np.asarray(g_run, dtype=np.float64), U_ref.shape)
    if U_run.shape != U_ref.shape:
        raise ValueError(f"profiles of {U_run.shape} and {U_ref.shape} planes")
    d_run, d_ref = plane_gradient(U_run), plane_gradient(U_ref)
    wall = 0.5 * (abs(d_ref[0]) + abs(d_ref[-1]))
    use = np.abs(d_ref) > core_frac * wall
    g = np.array(g_run, dtype=np.float64)
    g[use] = (1.0 + g_run[use]) * d_run[use] / d_ref[use] - 1.0
    idx = np.flatnonzero(use)
    if idx.size:
        nearest = idx[np.abs(np.arange(g.size)[:, None] - idx[None, :]).argmin(axis=1)]
        g = g[nearest]
    return np.clip(symmetrize(g), lo, hi)
