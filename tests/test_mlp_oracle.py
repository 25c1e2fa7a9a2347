"""Pins of the closure oracle (oracle/mlp_oracle.py): the analytic gradient
against central finite differences, and the feature/target restatement
against closed forms (a pure shear flow has no subgrid stress and one
non-zero shear feature; a solid-body mean gives zero targets)."""
import numpy as np

import mlp_oracle as M
from d3q19 import closure_params


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = closure_params(seed=3).astype(np.float64)
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
