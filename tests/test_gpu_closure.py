"""GPU tests of the learned closure: training kernels against the float64
oracle, training-data generation against its numpy restatement, the fused
inference inside the collision against closed forms, and what a trained
network gains over the plain method on held-out data."""
import numpy as np
import pytest

import d3q19
from d3q19 import workloads
import mlp_oracle as M

pytestmark = pytest.mark.gpu


def test_gradient_matches_oracle():
    rng = np.random.default_rng(1)
    p = d3q19.closure_params(seed=5)
    p[M.B3] = 0.2
    X = rng.standard_normal((1000, 10)).astype(np.float32)
    y = np.abs(rng.standard_normal(1000)).astype(np.float32)
    with d3q19.Trainer(2000, p) as tr:
        tr.add_samples(X, y)
        loss, g = tr.grad(0, 1000)
    loss_ref, g_ref = M.loss_and_grad(p, X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    scale = np.abs(g_ref).max()
    assert np.max(np.abs(g - g_ref)) < 2e-4 * scale


def test_training_recovers_a_teacher_network():
    rng = np.random.default_rng(2)
    teacher = d3q19.closure_params(seed=11)
    teacher[M.B3] = 0.5
    X = rng.standard_normal((32768, 10))
    y = M.softplus(M.forward(teacher, X))
    student = d3q19.closure_params(seed=12)
    with d3q19.Trainer(40000, student) as tr:
        tr.add_samples(X.astype(np.float32), y.astype(np.float32))
        l0, _ = tr.grad(0, 32768)
        tr.train(0, 32768, iterations=600, batch=4096, lr=3e-3)
        l1, _ = tr.grad(0, 32768)
        pred = tr.predict(0, 1000)
    assert l1 < 0.1 * l0, (l0, l1)
    assert np.isfinite(pred).all()


def _fine_flow(lx, ly, lz, seed):
    ux, uy, uz = workloads.laminar_parabola_noise(lx, ly, lz, seed=seed, uc=0.05, amp=0.2)
    return ux, uy, uz


@pytest.mark.parametrize("mode", ["local", "plane"])
def test_filtered_samples_match_restatement(mode):
    lx, ly, lz = 24, 16, 12
    ch = d3q19.Channel(lx, ly, lz, "fp64")
    p = d3q19.closure_params(scales=np.linspace(0.5, 2.0, 9))
    with ch, d3q19.Trainer(10 * lx * ly * lz, p) as tr:
        tr.set_target(mode)
        ux, uy, uz = _fine_flow(lx, ly, lz, 3)
        ch.set_macros(ux=ux, uy=uy, uz=uz)
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        ch.run(5)
        rho, ux, uy, uz = ch.macrovar(download=True)
        nu_c, s9_c, ystar_c = 0.0018, 1.0 / (3 * 0.0018 + 0.5), 0.5
        n = tr.add_filtered(ch, nu_c, s9_c, ystar_c)
        assert n == lx * ly * lz // 8
        f, t = tr.samples()
    F = M.box_filter(rho, ux, uy, uz)
    f_ref, t_ref = M.features_targets(F, nu_c, s9_c, 1.0 / ystar_c, p[M.SCALE:M.SCALE + 9].astype(np.float64),
                                      float(p[M.YSCALE]), plane=(mode == "plane"))
    assert np.allclose(f, f_ref, rtol=1e-4, atol=1e-7)
    if mode == "plane":
        # plane sums average the FP32 cancellation in tau = <uu> - <u><u>
        assert np.allclose(t, t_ref, rtol=2e-2, atol=2e-3)
        return
    # local targets are ratios of small differences: compare where the strain is resolved
    ok = (t_ref > 0.05) & (t_ref < 9.5)
    assert ok.sum() > 10
    assert np.allclose(t[ok], t_ref[ok], rtol=2e-2, atol=2e-3)


def _plane_columns(rho, ux, uy, uz, ustar):
    """statistc2 columns 2..12 of one sample from (z, y, x) field arrays."""
    m = lambda a: a.mean(axis=(0, 1))  # noqa: E731
    vx, vy, vz = m(ux) / ustar, m(uy) / ustar, m(uz) / ustar
    us2 = ustar * ustar
    pr = m(rho) / us2
    return np.stack([vx, vy, vz, m(ux * ux) / us2 - vx * vx, m(uy * uy) / us2 - vy * vy,
                     m(uz * uz) / us2 - vz * vz, m(ux * uz) / us2 - vx * vz, m(ux * uy) / us2 - vx * vy,
                     m(uy * uz) / us2 - vy * vz, pr, m(rho * rho) / us2 ** 2 - pr * pr], axis=1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_filtered_statistics_match_box_filter_of_fields(prec):
    lx, ly, lz = 24, 16, 12
    with d3q19.Channel(lx, ly, lz, prec, laminar=True) as ch:
        ux, uy, uz = _fine_flow(lx, ly, lz, 6)
        ch.set_macros(ux=ux, uy=uy, uz=uz)
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        ch.run(7)
        rho, ux, uy, uz = ch.macrovar(download=True)
        ch.stats_reset()
        ch.stats_accumulate_filtered()
        ch.stats_accumulate_filtered()
        out, n = ch.stats_read_filtered()
        ustar = ch.p["ustar"]
    assert n == 2 and out.shape == (lx // 2, 15)
    box = lambda a: a.reshape(lz // 2, 2, ly // 2, 2, lx // 2, 2).mean(axis=(1, 3, 5))  # noqa: E731
    want = _plane_columns(box(rho), box(ux), box(uy), box(uz), ustar)
    # columns 2..12 of saveload.f90:1492; the fields are FP64 in both cases
    assert np.allclose(out[:, 2:13], want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    assert np.allclose(out[:, 13], 1.0) and np.allclose(out[:, 14], 0.0)
    assert np.allclose(out[:, 0], np.arange(lx // 2) + 0.5)


def _run_fp32(lx, ly, lz, vel, steps, visc=None, closure=None, ystar=None, forcingp=False):
    with d3q19.Channel(lx, ly, lz, "fp32", visc=visc) as ch:
        ch.set_macros(ux=vel[0], uy=vel[1], uz=vel[2])
        ch.set_force(0.0, 2e-5, 0.0)
        ch.initpop()
        ch.macrovar()
        if closure is not None:
            ch.closure_set(closure, ystar=ystar)
        if forcingp:  # perturbation forcing (force varying over the nodes) from step 400 on
            ch.run_forcingp(400, steps, npforcing=10**9, force_in_y=2e-5)
        else:
            ch.run(steps)
        return ch.download_f()


@pytest.mark.parametrize("forcingp", [False, True])
def test_constant_eddy_viscosity_equals_doubled_viscosity(forcingp):
    """Zero weights and b3 = log(e - 1): nu_t = nu everywhere, i.e. plain MRT
    with viscosity 2 nu (also under the perturbation forcing)."""
    lx, ly, lz = 32, 24, 16
    vel = _fine_flow(lx, ly, lz, 4)
    nu = 0.0036
    p = d3q19.closure_params(init="zero", nu=nu)
    p[M.B3] = np.log(np.e - 1.0)
    a = _run_fp32(lx, ly, lz, vel, 100, visc=None, closure=p, ystar=1.0, forcingp=forcingp)
    b = _run_fp32(lx, ly, lz, vel, 100, visc=2 * nu, forcingp=forcingp)
    # the closure forms its rate in FP32, plain MRT rounds the FP64 rate once
    assert np.max(np.abs(a - b)) < 1e-5 * np.max(np.abs(b))


def test_vanishing_closure_is_plain_mrt():
    lx, ly, lz = 32, 24, 16
    vel = _fine_flow(lx, ly, lz, 5)
    p = d3q19.closure_params(seed=7, nu=0.0036)
    p[M.W3:M.B3] = 0.0
    p[M.B3] = -40.0  # softplus(-40) ~ 4e-18
    a = _run_fp32(lx, ly, lz, vel, 100, closure=p, ystar=1.0)
    b = _run_fp32(lx, ly, lz, vel, 100)
    assert np.max(np.abs(a - b)) < 1e-5 * np.max(np.abs(b))


def test_trained_closure_beats_no_closure_on_held_out_snapshots():
    """A priori: trained on filtered snapshots of a developing fine channel, the
    network predicts the subgrid eddy viscosity of later snapshots better than
    the plain method (nu_t = 0)."""
    lx, ly, lz = 64, 64, 32
    nu_f = 0.0036
    nu_c = nu_f / 2  # acoustic scaling: dx and dt double
    s9_c = 1.0 / (3 * nu_c + 0.5)
    # Re_tau = 60 keeps the centreline at U ~ 0.1 on 64 nodes; rate set 3 is
    # linearly stable there (rate set 1 is not above U ~ 0.1)
    with d3q19.Channel(lx, ly, lz, "fp32", mrttype=3, rstar=60) as fine:
        ux, uy, uz = workloads.loglaw_divfree(lx, ly, lz, fine.p["ustar"], fine.p["ystar"], seed=2, amp=0.3)
        fine.set_macros(ux=ux, uy=uy, uz=uz)
        fine.FORCING()
        fine.initpop()
        fine.macrovar()
        ystar_c = fine.p["ystar"] / 2
        Nc = lx * ly * lz // 8
        p = d3q19.closure_params(seed=1, scales=[1.0, 1.0, 1e-3, 1.0, 1e-3, 1.0, 1e-3, 1e-3, 1e-3], nu=nu_c)
        with d3q19.Trainer(12 * Nc, p) as tr:
            for _ in range(12):
                fine.run(50)
                tr.add_filtered(fine, nu_c, s9_c, ystar_c)
            feats, t = tr.samples()
            # the targets clip at zero, which would hide a diverged fine flow
            assert np.isfinite(feats).all()
            assert (t > 0).mean() > 0.1
            train_n = 8 * Nc
            tr.train(0, train_n, iterations=1500, batch=8192, lr=2e-3)
            pred = tr.predict(train_n, 4 * Nc)
            held = t[train_n:]
    mse_net = float(np.mean((pred - held) ** 2))
    mse_zero = float(np.mean(held ** 2))
    assert mse_net < 0.7 * mse_zero, (mse_net, mse_zero)
