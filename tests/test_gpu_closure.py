"""GPU tests of the learned closure: training kernels against the float64
oracle, training-data generation against its numpy restatement, the fused
inference inside the collision against closed forms, and what a trained
network gains over the plain method on held-out data."""
import numpy as np
import pytest

import d3q19
from d3q19 import calibrate, workloads
import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("floor", [0.0, -0.5])
def test_gradient_matches_oracle(floor):
    rng = np.random.default_rng(1)
    p = d3q19.closure_params(seed=5, nut_min=floor)
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
    y = M.nut(teacher, X)
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


def test_negative_floor_lowers_the_viscosity():
    """Zero weights and softplus(b3) = 0.2 under the floor -0.5: nu_t = -0.3 nu
    everywhere, i.e. plain MRT with viscosity 0.7 nu."""
    lx, ly, lz = 32, 24, 16
    vel = _fine_flow(lx, ly, lz, 6)
    nu = 0.0036
    p = d3q19.closure_params(init="zero", nu=nu, nut_min=-0.5)
    p[M.B3] = np.log(np.expm1(0.2))
    a = _run_fp32(lx, ly, lz, vel, 100, closure=p, ystar=1.0)
    b = _run_fp32(lx, ly, lz, vel, 100, visc=0.7 * nu)
    assert np.max(np.abs(a - b)) < 1e-5 * np.max(np.abs(b))


def test_closure_set_checks_the_floor():
    with d3q19.Channel(8, 6, 5, "fp32") as ch:
        p = d3q19.closure_params(seed=3, nu=ch.p["visc"])
        for bad in (-1.0, -1.5, 0.1, np.nan):
            q = p.copy()
            q[M.NUTMIN] = bad
            with pytest.raises(d3q19.LbmError, match="floor"):
                ch.closure_set(q)


def test_plane_targets_fill_by_sample_index():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((100, 10)).astype(np.float32)
    y = rng.standard_normal(100).astype(np.float32)
    vals = np.array([0.1, -0.2, 0.3, 0.4, -0.5], dtype=np.float32)
    with d3q19.Trainer(200, d3q19.closure_params(seed=2)) as tr:
        tr.add_samples(X, y)
        tr.set_plane_targets(vals, first=20, count=60)
        f, t = tr.samples()
        with pytest.raises(d3q19.LbmError):
            tr.set_plane_targets(vals, first=90, count=20)  # past the samples held
    assert np.array_equal(f, X)
    want = y.copy()
    want[20:80] = vals[np.arange(60) % 5]
    assert np.array_equal(t, want)


def test_balance_calibrated_closure_reproduces_a_laminar_reference():
    """The closure loop on a laminar channel: the reference is the profile of
    a prescribed wall-distance-dependent viscosity nu (1 + g*(x)) (closed form
    by quadrature); the plain run (uniform nu) misses it by ~25 %.  One
    momentum-balance update gives the per-plane target, the network is
    trained on the run's own closure inputs by the device's kernels, and the
    run with the fused closure (negative floor: g* < 0 near the walls) lands
    on the reference.  The non-equilibrium inputs of a laminar flow are
    fixed functions of the wall distance that scale with the viscosity the
    closure changes, so their scales are set large here: the network learns
    the target from the wall-distance input alone."""
    lx, ly, lz = 32, 4, 4
    nu, F = 0.05, 2e-5
    h = lx / 2.0
    gstar = lambda s: -0.3 + 0.6 * np.sin(np.pi * s / lx) ** 2  # noqa: E731
    x = np.arange(lx) + 0.5
    U_ref = np.empty(lx)
    for i, xi in enumerate(x):  # Simpson on 2000 intervals
        s = np.linspace(0.0, xi, 2001)
        f = F * (h - s) / (nu * (1.0 + gstar(s)))
        U_ref[i] = (s[1] - s[0]) / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    ystar = h / 100.0  # wall-distance feature from 0 to 1
    zero = np.zeros((lz, ly, lx))

    def profile(ch):
        return ch.macrovar(download=True)[2].mean(axis=(0, 1))

    with d3q19.Channel(lx, ly, lz, "fp32", visc=nu) as ch:
        ch.set_macros(ux=zero, uy=zero, uz=zero)
        ch.set_force(0.0, F, 0.0)
        ch.initpop()
        ch.macrovar()
        ch.run(40000)
        U0 = profile(ch)
        g0 = calibrate.balance_update(U0, U_ref, 0.0, core_frac=0.05)
        # the run's own closure inputs, from a closure that applies nu_t = 0
        p0 = d3q19.closure_params(seed=4, nu=nu, nut_min=-0.5)
        p0[M.W3:M.B3] = 0.0
        p0[M.B3] = np.log(np.expm1(0.5))
        ch.closure_set(p0, ystar=ystar)
        feat = ch.closure_inputs().reshape(-1, 10)
        sc = 1e6 * np.abs(feat[:, :9]).max() * np.ones(9)
        p = d3q19.closure_params(seed=4, nu=nu, nut_min=-0.5, scales=sc)
        with d3q19.Trainer(feat.shape[0], p) as tr:
            tr.add_samples(feat / np.r_[sc, 1.0].astype(np.float32), np.zeros(feat.shape[0], np.float32))
            tr.set_plane_targets(g0.astype(np.float32))
            tr.train(0, feat.shape[0], iterations=3000, batch=feat.shape[0], lr=3e-3)
            pred = tr.predict(0, lx)
            trained = tr.params()
        ch.closure_set(trained, ystar=ystar)
        ch.run(40000)
        U1 = profile(ch)
        nut = ch.closure_nut().mean(axis=(0, 1))
    e0 = np.max(np.abs(U0 - U_ref)) / U_ref.max()
    e1 = np.max(np.abs(U1 - U_ref)) / U_ref.max()
    print(f"plain {e0:.4f} closure {e1:.4f} fit {np.max(np.abs(pred - g0)):.4f} "
          f"applied {np.max(np.abs(nut - g0)):.4f}")
    assert np.max(np.abs(pred - g0)) < 0.03, np.max(np.abs(pred - g0))
    assert np.max(np.abs(nut - g0)) < 0.03, np.max(np.abs(nut - g0))
    assert e0 > 0.1, e0
    assert e1 < 0.1 * e0, (e0, e1)


def test_wall_distance_closure_applies_its_profile():
    """The calibrated wall-distance closure of the coarse channel: the
    network the device's kernels fit to the plane profile (inputs other than
    the wall distance scaled out) applies that profile at every node of the
    fused collision, whatever the flow."""
    from d3q19 import studies
    lxc, ly, lz = 96, 8, 6
    nu_c = 0.0018
    with d3q19.Channel(lxc, ly, lz, "fp32", mrttype=3, visc=nu_c) as ch:
        ystar_c = ch.p["ystar"]
        g = studies.wall_profile(lxc=lxc, ystar_c=ystar_c, **studies.WALL_CLOSURE_C5)
        params, fit = studies.plane_closure_params(g, nu_c, ystar_c)
        ux, uy, uz = workloads.laminar_parabola_noise(lxc, ly, lz, seed=8, uc=0.1, amp=0.3)
        ch.set_macros(ux=ux, uy=uy, uz=uz)
        ch.FORCING()
        ch.initpop()
        ch.macrovar()
        ch.closure_set(params, ystar=ystar_c)
        nut = ch.closure_nut().astype(np.float64)
    assert fit < 0.02, fit
    assert g.min() < -0.14 and g.max() > -0.04
    assert np.max(np.abs(nut - g[None, None, :])) < 0.02, np.max(np.abs(nut - g[None, None, :]))


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
