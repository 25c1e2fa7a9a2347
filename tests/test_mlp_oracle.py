"""Pins of the closure oracle (oracle/mlp_oracle.py): the analytic gradient
against central finite differences, and the feature/target restatement
against closed forms (a pure shear flow has no subgrid stress and one
non-zero shear feature; a solid-body mean gives zero targets)."""
import numpy as np
import pytest

import mlp_oracle as M
from d3q19 import closure_params


@pytest.mark.parametrize("floor", [0.0, -0.4])
def test_gradient_matches_finite_differences(floor):
    rng = np.random.default_rng(0)
    p = closure_params(seed=3, nut_min=floor).astype(np.float64)
    p[M.B3] = 0.3
    X = rng.standard_normal((37, 10))
    y = np.abs(rng.standard_normal(37))
    loss, g = M.loss_and_grad(p, X, y)
    idx = np.concatenate([rng.choice(M.NPARAM, 60, replace=False), [M.B3, M.W3, M.B1, M.W1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, X, y)[0] - M.loss_and_grad(pm, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_linear_shear_has_no_subgrid_stress():
    lz, ly, lx = 6, 8, 10
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    uy = 1e-3 * (x + 0.5)  # du_y/dx = 1e-3, constant: box filter is exact
    F = M.box_filter(np.zeros_like(uy), np.zeros_like(uy), uy, np.zeros_like(uy))
    feat, t = M.features_targets(F, nu=0.0018, s9=1.9, inv_ystar=1.0, scales=np.ones(9), yscale=100.0)
    f = feat.reshape(lz // 2, ly // 2, lx // 2, 10)
    inner = f[:, :, 1:-1]
    # coarse spacing 2 fine nodes: du_y/dx = 2e-3 per coarse cell; S_xy = 1e-3
    assert np.allclose(inner[..., 6], -2.0 / (3.0 * 1.9) * 1e-3)
    assert np.allclose(inner[..., [0, 1, 2, 3, 4, 5, 7, 8]], 0.0)
    # tau_xy of a linear profile over a box = variance term (1/4 of (du)^2 /... ) : nonnegative nu_t
    assert np.all(t >= 0.0) and np.all(t <= 10.0)


def _eddy_field(c):
    """Filtered fields on a 8 x 6 x 10 coarse grid whose subgrid stress is
    tau = -2 c S exactly (products = <u><u> - 2 c S), S from the same central
    differences the features use, and a wall-bounded, y- and z-varying u."""
    lz, ly, lx = 8, 6, 10
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    xs = x + 0.5
    ux = 1e-4 * np.sin(np.pi * xs / lx) ** 2 * np.cos(2 * np.pi * y / ly)
    uy = 1e-3 * xs * (lx - xs) * (1.0 + 0.1 * np.sin(2 * np.pi * z / lz))
    uz = 2e-4 * np.sin(np.pi * xs / lx) * np.sin(2 * np.pi * y / ly)
    u = [ux, uy, uz]

    def ddx(a):
        pad = np.concatenate([-a[:, :, :1], a, -a[:, :, -1:]], axis=2)
        return 0.5 * (pad[:, :, 2:] - pad[:, :, :-2])

    def ddy(a):
        return 0.5 * (np.roll(a, -1, axis=1) - np.roll(a, 1, axis=1))

    def ddz(a):
        return 0.5 * (np.roll(a, -1, axis=0) - np.roll(a, 1, axis=0))

    g = [[ddx(q), ddy(q), ddz(q)] for q in u]
    S = [[0.5 * (g[a][b] + g[b][a]) for b in range(3)] for a in range(3)]
    F = {"rho": np.zeros_like(ux), "ux": ux, "uy": uy, "uz": uz}
    names = {(0, 0): "uxux", (1, 1): "uyuy", (2, 2): "uzuz", (0, 1): "uxuy", (1, 2): "uyuz", (0, 2): "uxuz"}
    for (a, b), name in names.items():
        F[name] = u[a] * u[b] - 2.0 * c * S[a][b]
    return F


def test_targets_recover_a_prescribed_eddy_viscosity():
    """tau = -2 c S gives nu_t = c at every node and on every plane, so both
    target modes return c / nu; a single wrong factor (the 2, the sign, the
    nu) moves them off."""
    nu, c = 0.0018, 0.0027
    F = _eddy_field(c)
    for plane in (False, True):
        _, t = M.features_targets(F, nu=nu, s9=1.9, inv_ystar=1.0, scales=np.ones(9), yscale=100.0, plane=plane)
        assert np.allclose(t, c / nu, rtol=1e-9), plane


def test_plane_target_is_the_least_squares_ratio_of_its_plane():
    """A mix of two eddy viscosities over a plane: the per-plane target is
    -sum(tau:S)/(2 sum(S:S)), i.e. the S:S-weighted mean of the local ones."""
    nu = 0.0018
    F1, F2 = _eddy_field(0.001), _eddy_field(0.004)
    F = {k: np.where(np.arange(F1["ux"].shape[1])[None, :, None] < 3, F1[k], F2[k]) for k in F1}
    _, tl = M.features_targets(F, nu=nu, s9=1.9, inv_ystar=1.0, scales=np.ones(9), yscale=100.0)
    _, tp = M.features_targets(F, nu=nu, s9=1.9, inv_ystar=1.0, scales=np.ones(9), yscale=100.0, plane=True)
    lz, ly, lx = F["ux"].shape
    tl, tp = tl.reshape(lz, ly, lx), tp.reshape(lz, ly, lx)
    assert np.allclose(tp, tp[:1, :1, :])  # one value per x plane
    assert np.all((tp[0, 0] > 0.001 / nu) & (tp[0, 0] < 0.004 / nu))
    assert not np.allclose(tl, tp)
