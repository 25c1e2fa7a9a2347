"""CPU oracle for the learned closure -- TEST INFRASTRUCTURE ONLY (tests/ check
the GPU results against it).

Float64 numpy restatements of:
  * the 10-32-32-1 tanh network, nu_t/nu = softplus(o) + m (m: the floor
    stored after nu) and the loss mean((softplus(o) + m - y)^2),
    with the analytic gradient (backpropagation), pinned against central
    finite differences in tests/test_mlp_oracle.py;
  * the 2x2x2 box filter of a fine channel (means of rho, u, u_a u_b from the
    reference's macrovar, Channel-Flow/collision.f90:398-417);
  * the coarse features and targets: strain by central differences with
    no-slip walls half a cell outside, subgrid stress, eddy viscosity
    nu_t = -tau:S/(2 S:S) as nu_t/nu clipped to [-10, 10], first-order
    non-equilibrium moments -(2/(3 s9)) x (strain combination).
"""
import numpy as np

NPARAM, NIN, NH = 1441, 10, 32
W1, B1 = 0, 320
W2, B2 = 352, 1376
W3, B3 = 1408, 1440
SCALE, YSCALE, NU, NUTMIN = 1441, 1450, 1451, 1452


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(NH, NIN), p[B1:W2], p[W2:B2].reshape(NH, NH), p[B2:W3], p[W3:B3], p[B3])


def softplus(o):
    return np.where(o > 20, o, np.log1p(np.exp(np.minimum(o, 20))))


def floor(p):
    """The floor m of nu_t/nu (0 for a vector without one)."""
    p = np.asarray(p)
    return float(p[NUTMIN]) if p.size > NUTMIN else 0.0


def nut(p, X):
    """nu_t/nu of the network for inputs X."""
    return softplus(forward(p, X)) + floor(p)


def forward(p, X):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    h1 = np.tanh(X @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    return h2 @ w3 + b3


def loss_and_grad(p, X, y):
    """Mean squared error of softplus(o) + m against y and its gradient (NPARAM)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(p)
    n = len(y)
    a1 = X @ w1.T + b1
    h1 = np.tanh(a1)
    a2 = h1 @ w2.T + b2
    h2 = np.tanh(a2)
    o = h2 @ w3 + b3
    pred = softplus(o) + floor(p)
    err = pred - y
    loss = np.mean(err * err)
    g_o = 2.0 * err / n / (1.0 + np.exp(-o))
    g = np.zeros(NPARAM)
    g[W3:B3] = g_o @ h2
    g[B3] = g_o.sum()
    g_a2 = np.outer(g_o, w3) * (1.0 - h2 * h2)
    g[W2:B2] = (g_a2.T @ h1).ravel()
    g[B2:W3] = g_a2.sum(0)
    g_a1 = (g_a2 @ w2) * (1.0 - h1 * h1)
    g[W1:B1] = (g_a1.T @ X).ravel()
    g[B1:W2] = g_a1.sum(0)
    return loss, g


def box_filter(rho, ux, uy, uz):
    """2x2x2 means of rho, u and u_a u_b; fields (lz, ly, lx) -> (lz/2, ly/2, lx/2)."""
    def avg(a):
        lz, ly, lx = a.shape
        return a.reshape(lz // 2, 2, ly // 2, 2, lx // 2, 2).mean(axis=(1, 3, 5))
    return {"rho": avg(rho), "ux": avg(ux), "uy": avg(uy), "uz": avg(uz), "uxux": avg(ux * ux),
            "uyuy": avg(uy * uy), "uzuz": avg(uz * uz), "uxuy": avg(ux * uy), "uyuz": avg(uy * uz),
            "uxuz": avg(ux * uz)}


def features_targets(F, nu, s9, inv_ystar, scales, yscale, plane=False):
    """Features (N, 10) and targets (N,) of the coarse nodes, node order ix fastest.
    plane=True: every node's target is its x plane's least-squares eddy
    viscosity -sum(tau:S) / (2 sum(S:S)) / nu (clipped to [-10, 10])."""
    u = [F["ux"], F["uy"], F["uz"]]  # (lz, ly, lx)
    lz, ly, lx = u[0].shape

    def ddx(a):  # no-slip walls: ghost = -a beyond both ends
        pad = np.concatenate([-a[:, :, :1], a, -a[:, :, -1:]], axis=2)
        return 0.5 * (pad[:, :, 2:] - pad[:, :, :-2])

    def ddy(a):
        return 0.5 * (np.roll(a, -1, axis=1) - np.roll(a, 1, axis=1))

    def ddz(a):
        return 0.5 * (np.roll(a, -1, axis=0) - np.roll(a, 1, axis=0))

    g = [[ddx(c), ddy(c), ddz(c)] for c in u]
    S = [[0.5 * (g[c][d] + g[d][c]) for d in range(3)] for c in range(3)]
    tau = [[None] * 3 for _ in range(3)]
    tau[0][0] = F["uxux"] - u[0] * u[0]
    tau[1][1] = F["uyuy"] - u[1] * u[1]
    tau[2][2] = F["uzuz"] - u[2] * u[2]
    tau[0][1] = tau[1][0] = F["uxuy"] - u[0] * u[1]
    tau[1][2] = tau[2][1] = F["uyuz"] - u[1] * u[2]
    tau[0][2] = tau[2][0] = F["uxuz"] - u[0] * u[2]
    ts = sum(tau[c][d] * S[c][d] for c in range(3) for d in range(3))
    ss = sum(S[c][d] * S[c][d] for c in range(3) for d in range(3))
    if plane:
        a, b = ts.sum(axis=(0, 1)), ss.sum(axis=(0, 1))
        yp = np.where(b > 1e-30, -a / (2.0 * np.where(b > 1e-30, b, 1.0)) / nu, 0.0)
        y = np.broadcast_to(np.clip(yp, -10.0, 10.0), ts.shape)
    else:
        y = np.where(ss > 1e-30, -ts / (2.0 * np.where(ss > 1e-30, ss, 1.0)) / nu, 0.0)
        y = np.clip(y, -10.0, 10.0)
    k = -2.0 / (3.0 * s9)
    neq = [np.zeros_like(u[0]), np.zeros_like(u[0]), k * (2 * S[0][0] - S[1][1] - S[2][2]), np.zeros_like(u[0]),
           k * (S[1][1] - S[2][2]), np.zeros_like(u[0]), k * S[0][1], k * S[1][2], k * S[0][2]]
    ix = np.arange(lx, dtype=np.float64)
    wd = np.minimum(ix + 0.5, lx - ix - 0.5)
    feat = np.stack([neq[q] / scales[q] for q in range(9)] + [np.broadcast_to(wd * inv_ystar / yscale, u[0].shape)],
                    axis=-1)
    return feat.reshape(-1, 10), y.reshape(-1)
