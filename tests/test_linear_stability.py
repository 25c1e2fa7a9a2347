"""Linear stability of the reference's three MRT rate sets for a uniform stream.

One step (macroscopic fields from f, collision of Channel-Flow/collision.f90:
56-220, streaming) linearised about the equilibrium of a uniform velocity U:
a Fourier mode exp(i k.x) of the populations is multiplied by
diag(exp(-i k.c)) J, with J the Jacobian of the node collision.  The collision
is quadratic in f, so central differences give J to round-off.

Rate set 1 (para.f90:111-119, the turbulent branch's choice) has a growing
mode once U along a lattice axis exceeds ~0.1; the Re_tau = 180 channel at
nu = 0.0036 reaches U ~ 0.12 on its centreline, so any faithful run of it with
a three-dimensional disturbance diverges after a few thousand steps.  Rate
set 3 (para.f90:131-140) stays neutral at the same speed.
"""
import numpy as np
import pytest

import oracle as O

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _equilibrium(u):
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    return W * (3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


def _node_step(p, f):
    rho = f.sum()
    u = np.array([CX @ f, CY @ f, CZ @ f])
    return O.collide_node(p, f, rho, *u, 0.0, 0.0, 0.0)


def _jacobian(p, u):
    f0 = _equilibrium(u)
    J = np.empty((19, 19))
    h = 1e-6
    for k in range(19):
        e = np.zeros(19)
        e[k] = h
        J[:, k] = (_node_step(p, f0 + e) - _node_step(p, f0 - e)) / (2 * h)
    return J


def _max_amplification(p, u, nk):
    J = _jacobian(p, u)
    ks = np.linspace(-np.pi, np.pi, nk + 1)
    KX, KY, KZ = np.meshgrid(ks, ks, ks[ks >= 0], indexing="ij")
    K = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1)
    ph = np.exp(-1j * (K @ np.stack([CX, CY, CZ]).astype(float)))
    return float(np.abs(np.linalg.eigvals(ph[:, :, None] * J[None])).max())


def test_jacobian_depends_only_on_the_conserved_moments():
    """The base state is the BGK equilibrium, which rate set 1 does not keep
    fixed (its epsilon and pi equilibria differ); the collision is quadratic
    in f through (rho, j) only, so any state with the same (rho, j) linearises
    to the same J."""
    p = O.para(192, 768, 384, laminar=False, mrttype=1)
    u = np.array([0.0, 0.13, 0.0])
    f0 = _equilibrium(u)
    assert np.max(np.abs(_node_step(p, f0) - f0)) > 1e-4
    ghost = np.random.default_rng(0).standard_normal(19) * 1e-3
    M = np.stack([np.ones(19), CX, CY, CZ]).astype(float)
    ghost -= M.T @ np.linalg.solve(M @ M.T, M @ ghost)  # no mass or momentum
    J0 = _jacobian(p, u)
    J1 = np.empty((19, 19))
    for k in range(19):
        e = np.zeros(19)
        e[k] = 1e-6
        J1[:, k] = (_node_step(p, f0 + ghost + e) - _node_step(p, f0 + ghost - e)) / 2e-6
    assert np.max(np.abs(J1 - J0)) < 1e-8


def test_conserved_modes_have_unit_amplification_at_k0():
    p = O.para(192, 768, 384, laminar=False, mrttype=1)
    lam = np.abs(np.linalg.eigvals(_jacobian(p, np.array([0.0, 0.05, 0.0]))))
    # rho, jx, jy, jz are collision invariants: four eigenvalues of exactly 1
    assert np.sum(np.abs(lam - 1.0) < 1e-9) == 4
    assert lam.max() < 1.0 + 1e-9


def test_rate_set_1_grows_at_channel_centreline_speed():
    p = O.para(192, 768, 384, laminar=False, mrttype=1)
    lam = _max_amplification(p, np.array([0.0, 0.13, 0.0]), nk=8)
    assert lam > 1.005, lam


@pytest.mark.parametrize("mrt", [2, 3])
def test_rate_sets_2_and_3_are_neutral_at_the_same_speed(mrt):
    p = O.para(192, 768, 384, laminar=False, mrttype=mrt)
    lam = _max_amplification(p, np.array([0.0, 0.13, 0.0]), nk=16)
    assert lam < 1.0 + 1e-9, lam


def test_rate_set_1_is_neutral_at_low_speed():
    p = O.para(192, 768, 384, laminar=False, mrttype=1)
    lam = _max_amplification(p, np.array([0.0, 0.05, 0.0]), nk=16)
    assert lam < 1.0 + 1e-9, lam
