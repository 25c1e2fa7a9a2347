"""The learned closure inside the collision (matrix-core network, FP32/FP16
storage) against the CPU oracle, through the C ABI.

* The network inputs the fused collision forms at each node are the
  non-equilibrium parts of the moments it relaxes.  The oracle gives them
  independently: with no body force, M (post - f) = -S (m - m_eq) for the
  oracle's collision of that node, with M built from the moment definitions
  of d'Humieres et al. (2002).
* nu_t/nu from the matrix-core network (FP16 operands, FP32 sums) against the
  float64 network of oracle/mlp_oracle.py on the same inputs.
* One collision with the closure equals the oracle's collision with each
  node's shear rates set from that nu_t, followed by streaming.
"""
import ctypes

import numpy as np
import pytest

import d3q19
import mlp_oracle as M
import oracle as O
from test_oracle_pins import CX, CY, CZ, OPP, moment_matrix

pytestmark = pytest.mark.gpu

ROWS = [1, 2, 9, 10, 11, 12, 13, 14, 15]  # e, eps, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz


def _state(lx, ly, lz, seed):
    """Seeded populations (FP32-representable) and the oracle for the same grid."""
    o = O.OracleChannel(lx, ly, lz)
    rng = np.random.default_rng(seed)
    o.ux[:] = 0.01 * rng.standard_normal(o.ux.shape)
    o.uy[:] = 0.05 * (1 + 0.2 * rng.standard_normal(o.uy.shape))
    o.uz[:] = 0.01 * rng.standard_normal(o.uz.shape)
    o.initpop()
    o.f += 1e-3 * rng.standard_normal(o.f.shape)
    f0 = o.f.astype(np.float32).astype(np.float64)
    return o, f0


def _fields(f):
    c = np.stack([CX, CY, CZ], 1).astype(float)
    return f.sum(-1), f @ c[:, 0], f @ c[:, 1], f @ c[:, 2]


def _rates(p):
    return np.array([p.s1, p.s2, p.s9, p.s10, p.s9, p.s10, p.s13, p.s13, p.s13])


def _oracle_neq(o, f0):
    """Non-equilibrium moments from the oracle's collision (no force)."""
    Mm = moment_matrix()
    rho, ux, uy, uz = _fields(f0)
    s = _rates(o.p)
    out = np.empty(f0.shape[:-1] + (9,))
    for idx in np.ndindex(f0.shape[:-1]):
        post = O.collide_node(o.p, f0[idx], rho[idx], ux[idx], uy[idx], uz[idx], 0.0, 0.0, 0.0)
        out[idx] = -(Mm @ (post - f0[idx]))[ROWS] / s
    return out


def _channel(lx, ly, lz, f0, params, prec="fp32"):
    ch = d3q19.Channel(lx, ly, lz, prec)
    ch.upload_f(f0)
    ch.set_force(0.0, 0.0, 0.0)
    ch.macrovar()
    ch.closure_set(params)
    return ch


def _params(scales, nu, seed=3):
    p = d3q19.closure_params(seed=seed, scales=scales, nu=nu)
    p[M.B3] = 0.0  # nu_t ~ 0.7 nu: the network output matters
    return p


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16c"])
def test_fused_closure_inputs_are_the_oracle_nonequilibrium_moments(prec):
    lx, ly, lz = 12, 10, 7  # 840 nodes: the last block is partly past the lattice
    o, f0 = _state(lx, ly, lz, 41)
    scales = np.abs(_oracle_neq(o, f0)).reshape(-1, 9).std(axis=0) + 1e-30
    p = _params(scales, o.p.visc)
    with _channel(lx, ly, lz, f0, p, prec) as ch:
        feat = ch.closure_inputs().astype(np.float64)
        ystar = ch.p["ystar"]
        fdev = ch.download_f()  # the stored values (FP16 storage rounds f0)
    neq = _oracle_neq(o, fdev)
    got = feat[..., :9] * scales
    tol = 1e-4 * np.abs(neq).reshape(-1, 9).max(axis=0)
    err = np.abs(got - neq).reshape(-1, 9).max(axis=0)
    assert np.all(err <= tol), (err, tol)
    ix = np.arange(lx)
    wd = np.minimum(ix + 0.5, lx - ix - 0.5)
    np.testing.assert_allclose(feat[..., 9], np.broadcast_to(wd / ystar / 100.0, feat.shape[:-1]), rtol=1e-6)


def test_matrix_core_network_matches_float64_network():
    lx, ly, lz = 12, 10, 7
    o, f0 = _state(lx, ly, lz, 42)
    neq = _oracle_neq(o, f0)
    scales = np.abs(neq).reshape(-1, 9).std(axis=0) + 1e-30
    p = _params(scales, o.p.visc)
    with _channel(lx, ly, lz, f0, p) as ch:
        feat = ch.closure_inputs().astype(np.float64).reshape(-1, 10)
        nut = ch.closure_nut().astype(np.float64).reshape(-1)
    want = M.nut(p, feat)
    assert np.std(want) > 0.05  # the inputs move the output
    # FP16 operands: inputs and hidden values carry 2^-11 relative rounding
    assert np.max(np.abs(nut - want)) < 5e-3 * np.max(np.abs(want)), np.max(np.abs(nut - want))


def test_large_hidden_biases_saturate_as_the_float64_network():
    """Hidden biases near the +-32 bound: the fused form carries each bias as
    the factor 2^(c b) on 2^(c a) and relies on IEEE overflow (a product of
    inf) or a vanishing product for the saturated units; the float64 network
    gives +-1 there, and the units with small biases still move the output."""
    lx, ly, lz = 12, 10, 7
    o, f0 = _state(lx, ly, lz, 44)
    neq = _oracle_neq(o, f0)
    scales = np.abs(neq).reshape(-1, 9).std(axis=0) + 1e-30
    p = _params(scales, o.p.visc)
    rng = np.random.default_rng(6)
    p[M.B1:M.W2] = rng.choice([-31.9, -20.0, 0.2, 20.0, 31.9], M.NH)
    p[M.B2:M.W3] = rng.choice([-31.9, 0.1, 31.9], M.NH)
    p[M.B2:M.B2 + 4] = 0.0  # four units of layer 2 stay unsaturated
    p[M.W1:M.B1] *= 4.0     # larger pre-activations for the small-bias units
    with _channel(lx, ly, lz, f0, p) as ch:
        feat = ch.closure_inputs().astype(np.float64).reshape(-1, 10)
        nut = ch.closure_nut().astype(np.float64).reshape(-1)
    want = M.nut(p, feat)
    assert np.all(np.isfinite(nut))
    assert np.std(want) > 0.01
    assert np.max(np.abs(nut - want)) < 5e-3 * np.max(np.abs(want)), np.max(np.abs(nut - want))


def test_closure_set_rejects_biases_beyond_the_bound():
    with d3q19.Channel(8, 6, 5, "fp32") as ch:
        p = d3q19.closure_params(seed=3, nu=ch.p["visc"])
        ch.closure_set(p)
        for where, bad in ((M.B1 + 3, 32.5), (M.B2 + 7, -33.0), (M.B2, np.nan)):
            q = p.copy()
            q[where] = bad
            with pytest.raises(d3q19.LbmError, match="bias"):
                ch.closure_set(q)


def test_collision_applies_the_reported_eddy_viscosity():
    lx, ly, lz = 8, 6, 5
    o, f0 = _state(lx, ly, lz, 43)
    neq = _oracle_neq(o, f0)
    scales = np.abs(neq).reshape(-1, 9).std(axis=0) + 1e-30
    p = _params(scales, o.p.visc)
    with _channel(lx, ly, lz, f0, p) as ch:
        nut = ch.closure_nut().astype(np.float64)
        ch.collision_MRT()
        got = ch.download_f()
    rho, ux, uy, uz = _fields(f0)
    post = np.empty_like(f0)
    for idx in np.ndindex(f0.shape[:-1]):
        q = type(o.p).from_buffer_copy(o.p)
        q.s9 = q.s13 = 1.0 / (3.0 * o.p.visc * (1.0 + nut[idx]) + 0.5)
        post[idx] = O.collide_node(q, f0[idx], rho[idx], ux[idx], uy[idx], uz[idx], 0.0, 0.0, 0.0)
    want = np.empty_like(post)
    for i in range(19):  # push streaming, periodic y/z, halfway bounce-back in x
        a = np.roll(np.roll(post[..., i], CY[i], axis=1), CZ[i], axis=0)
        if CX[i] == 0:
            want[..., i] = a
        elif CX[i] == 1:
            want[:, :, 1:, i] = a[:, :, :-1]
            want[:, :, lx - 1, OPP[i]] = post[:, :, lx - 1, i]
        else:
            want[:, :, :-1, i] = a[:, :, 1:]
            want[:, :, 0, OPP[i]] = post[:, :, 0, i]
    # FP32 arithmetic on FP32 storage against the FP64 oracle
    assert np.max(np.abs(got - want)) < 2e-6 * np.max(np.abs(want)), np.max(np.abs(got - want))
