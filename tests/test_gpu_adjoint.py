"""Discrete adjoint of the bounce-back step (lbm_adjoint_forward/backward).

The reference has no adjoint; the gradients are pinned by central finite
differences of the forward solver in FP64 (h = 1e-4 relative), by the FP64
GPU adjoint for the FP32 one, and by the checkpoint schedule not changing
them. Objective: time-averaged squared wall shear stress over a wall segment
plus the inlet-outlet pressure drop; parameters: inlet-profile coefficients
and a per-node Brinkman porosity field. Every test runs on D3Q19 and D3Q27.
"""
import numpy as np
import pytest

import lbm
import workloads

pytestmark = pytest.mark.gpu

TAU, DT = 0.8, 1e-5
T, K = 60, 16  # 3 full checkpoint intervals and a partial one
M = 3


@pytest.fixture(scope="module", params=[19, 27], ids=["D3Q19", "D3Q27"])
def setup(tmp_path_factory, request):
    wd = str(tmp_path_factory.mktemp("adj"))
    cfg = workloads.straight_tube(wd, radius=5.0, length=20.0, tau=TAU, cte_vel=0.03, max_iter=T)
    geom = lbm.Geometry.from_config(cfg, q=request.param)
    ijk, kind, _ = geom.nodes()
    inl = np.where(kind == lbm.KIND_INLET)[0]
    x = ijk[inl, 0] - ijk[inl, 0].mean()
    y = ijk[inl, 1] - ijk[inl, 1].mean()
    r2 = (x * x + y * y) / max((x * x + y * y).max(), 1)
    basis = np.stack([np.ones_like(r2), x / max(np.abs(x).max(), 1), 1.0 - r2])
    nw, _ = geom.nearwall()
    z = ijk[nw, 2]
    mask = ((z > z.min() + 0.3 * np.ptp(z)) & (z < z.min() + 0.7 * np.ptp(z))).astype(np.uint8)
    rng = np.random.default_rng(4)
    alpha = (1e-3 * np.abs(rng.standard_normal(geom.n))).astype(np.float32)
    theta = np.array([0.03, 0.004, 0.01])
    s = 1.0 + 0.3 * np.sin(2 * np.pi * np.arange(T) / 40.0)
    st = dict(geom=geom, basis=basis, mask=mask, alpha=alpha, theta=theta, s=s, wss_w=1.0, dp_w=0.0)
    # Weights that give both objective terms the same size.
    Jw = objective(st, theta)[1]
    st.update(wss_w=0.0, dp_w=1.0)
    Jp = objective(st, theta)[1]
    st.update(wss_w=abs(Jp / Jw), dp_w=1.0, Jw=Jw, Jp=Jp)
    return st


def make(st, prec=lbm.PREC_F64, alpha=None):
    sol = lbm.Solver(st["geom"], precision=prec, tau=TAU, dt=DT, wall_mode=lbm.WALL_HALFWAY_BB,
                     inlet_mode=lbm.INLET_NEE_VELOCITY, outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0)
    sol.set_inlet_profile(st["basis"])
    sol.set_porosity(st["alpha"] if alpha is None else alpha)
    return sol


def objective(st, theta, alpha=None, prec=lbm.PREC_F64):
    sol = make(st, prec, alpha)
    coef = st["s"][:, None] * theta[None, :]
    J = sol.adjoint_forward(T, coef, wss_weight=st["wss_w"], dp_weight=st["dp_w"], wss_mask=st["mask"],
                            checkpoint_every=K)
    return sol, J


def gradients(st, prec=lbm.PREC_F64, slots=0, tape=True):
    sol, J = objective(st, st["theta"], prec=prec)
    gc, ga = sol.adjoint_backward(slots, tape=tape)
    gtheta = (st["s"][:, None] * gc).sum(0)
    return J, gtheta, ga, sol


def test_objective_terms_are_both_active(setup):
    print(f"stress term (unit weight) {setup['Jw']:.6e}, pressure-drop term {setup['Jp']:.6e}")
    assert setup["Jw"] > 0 and setup["Jp"] != 0
    _, J = objective(setup, setup["theta"])
    # J is linear in the two weights.
    assert J == pytest.approx(setup["wss_w"] * setup["Jw"] + setup["Jp"], rel=1e-9)


def test_inlet_gradient_matches_central_differences(setup):
    J, gtheta, _, _ = gradients(setup)
    for m in range(M):
        h = 1e-4 * abs(setup["theta"][m])
        tp = setup["theta"].copy()
        tm = setup["theta"].copy()
        tp[m] += h
        tm[m] -= h
        fd = (objective(setup, tp)[1] - objective(setup, tm)[1]) / (2 * h)
        rel = abs(gtheta[m] - fd) / max(abs(fd), 1e-300)
        print(f"theta[{m}]: adjoint {gtheta[m]:.9e}  FD {fd:.9e}  rel {rel:.2e}", flush=True)
        assert rel <= 1e-3


def test_porosity_gradient_matches_central_differences(setup):
    J, _, ga, _ = gradients(setup)
    geom = setup["geom"]
    ijk, kind, _ = geom.nodes()
    nw, _ = geom.nearwall()
    picks = [int(np.argmax(np.abs(ga))), int(nw[np.argmax(setup["mask"])]), int(np.where(kind == lbm.KIND_INLET)[0][0]),
             geom.n // 2]
    for n in picks:
        h = 1e-4
        ap = setup["alpha"].astype(np.float64).copy()
        am = ap.copy()
        ap[n] += h
        am[n] -= h
        # FP32 porosity input: perturb by exactly representable amounts.
        ap32, am32 = ap.astype(np.float32), am.astype(np.float32)
        hh = (float(ap32[n]) - float(am32[n])) / 2
        fd = (objective(setup, setup["theta"], ap32)[1] - objective(setup, setup["theta"], am32)[1]) / (2 * hh)
        rel = abs(ga[n] - fd) / max(abs(fd), 1e-300)
        print(f"alpha[{n}] (kind {kind[n]}): adjoint {ga[n]:.9e}  FD {fd:.9e}  rel {rel:.2e}", flush=True)
        assert rel <= 1e-3 or abs(ga[n] - fd) <= 1e-9 * np.max(np.abs(ga))
    # Directional derivative over the whole field.
    rng = np.random.default_rng(9)
    v = rng.standard_normal(geom.n)
    eps = 1e-4 / np.max(np.abs(v))
    ap = (setup["alpha"] + eps * v).astype(np.float32)
    am = (setup["alpha"] - eps * v).astype(np.float32)
    dv = (ap.astype(np.float64) - am.astype(np.float64)) / 2
    fd = (objective(setup, setup["theta"], ap)[1] - objective(setup, setup["theta"], am)[1])
    ad = 2 * float(ga @ dv)
    print(f"directional: adjoint {ad:.9e}  FD {fd:.9e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd)


def test_checkpoint_slots_do_not_change_gradients(setup):
    base = gradients(setup, slots=K - 1)
    for slots in (0, 1, 3):
        J, gt, ga, _ = gradients(setup, slots=slots)
        assert J == base[0]
        assert np.array_equal(gt, base[1]) and np.array_equal(ga, base[2]), slots


@pytest.mark.parametrize("prec", [lbm.PREC_F64, lbm.PREC_F32])
def test_moments_tape_does_not_change_gradients(setup, prec):
    # The reverse step reads the recomputation's moments tape where it can;
    # re-pulling the moments from the recomputed state must agree bit for bit
    # (with several slot counts, so that tape and re-pull steps mix
    # differently).
    for slots in (K - 1, 1):
        _, gt, ga, _ = gradients(setup, prec, slots=slots, tape=True)
        _, gt0, ga0, _ = gradients(setup, prec, slots=slots, tape=False)
        assert np.array_equal(gt, gt0) and np.array_equal(ga, ga0), slots


def test_fp32_adjoint_tracks_fp64(setup):
    J64, gt64, ga64, _ = gradients(setup, lbm.PREC_F64)
    J32, gt32, ga32, _ = gradients(setup, lbm.PREC_F32)
    rt = np.linalg.norm(gt32 - gt64) / np.linalg.norm(gt64)
    ra = np.linalg.norm(ga32 - ga64) / np.linalg.norm(ga64)
    print(f"J rel {abs(J32 - J64) / abs(J64):.2e}, theta-gradient rel L2 {rt:.2e}, porosity-gradient rel L2 {ra:.2e}")
    assert abs(J32 - J64) <= 1e-4 * abs(J64)
    assert rt <= 1e-3 and ra <= 1e-3


def test_forward_pass_is_the_plain_forward_and_state_is_restored(setup):
    coef = setup["s"][:, None] * setup["theta"][None, :]
    a = make(setup)
    a.step_coefficients(coef)
    b, _ = objective(setup, setup["theta"])
    fb = b.populations()
    assert np.array_equal(a.populations(), fb)
    b.adjoint_backward(2)
    assert np.array_equal(b.populations(), fb)


def test_torch_autograd_function(setup):
    torch = pytest.importorskip("torch")

    class Objective(torch.autograd.Function):
        """J(theta, alpha) through lbm_adjoint_forward / lbm_adjoint_backward."""

        @staticmethod
        def forward(ctx, theta, alpha):
            sol, J = objective(setup, theta.detach().numpy(), alpha.detach().numpy().astype(np.float32))
            gc, ga = sol.adjoint_backward()
            ctx.save_for_backward(torch.from_numpy((setup["s"][:, None] * gc).sum(0)), torch.from_numpy(ga))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            gt, ga = ctx.saved_tensors
            return gJ * gt, gJ * ga

    theta = torch.tensor(setup["theta"], dtype=torch.float64, requires_grad=True)
    alpha = torch.tensor(setup["alpha"].astype(np.float64), requires_grad=True)
    J = Objective.apply(theta, alpha)
    (2.0 * J).backward()
    _, gt, ga, _ = gradients(setup)
    assert np.allclose(theta.grad.numpy(), 2 * gt, rtol=1e-12, atol=0)
    assert np.allclose(alpha.grad.numpy(), 2 * ga, rtol=1e-12, atol=0)
    # torch's own finite-difference check on the inlet coefficients.
    f = lambda th: Objective.apply(th, alpha.detach())  # noqa: E731
    assert torch.autograd.gradcheck(f, (theta.detach().clone().requires_grad_(True),), eps=1e-7, atol=1e-8,
                                    rtol=1e-3)


def test_backward_refuses_a_changed_solver(setup):
    # The reverse pass belongs to the recorded forward pass: a later step,
    # new populations, a new inlet profile or a new porosity field ends the
    # record, and the backward call then reports LBM_ERR_STATE instead of
    # returning gradients of another problem.
    st = setup
    coef = st["s"][:, None] * st["theta"][None, :]
    geom = st["geom"]
    _, _, w = lbm.lattice(geom.q)
    changes = {
        "step": lambda s: s.step(1),
        "populations": lambda s: s.set_populations(np.tile(w, (geom.n, 1))),
        "inlet profile": lambda s: s.set_inlet_profile(np.vstack([st["basis"], st["basis"][:1]])),
        "porosity": lambda s: s.set_porosity(st["alpha"]),
    }
    for name, change in changes.items():
        sol = make(st)
        sol.adjoint_forward(T, coef=coef, wss_weight=st["wss_w"], dp_weight=st["dp_w"], wss_mask=st["mask"],
                            checkpoint_every=K)
        g1, _ = sol.adjoint_backward()
        g2, _ = sol.adjoint_backward()  # the record survives a backward pass
        assert np.array_equal(g1, g2)
        change(sol)
        with pytest.raises(lbm.LbmError) as e:
            sol.adjoint_backward()
        assert e.value.code == lbm.ERR_STATE, name


@pytest.mark.parametrize("prec", [lbm.PREC_F64, lbm.PREC_F32], ids=["f64", "f32"])
def test_second_pass_after_reset_repeats_the_first(setup, prec):
    # lbm_set_populations(NULL) returns the solver to the rest state it was
    # created in; a second forward/backward pair on the same solver (its
    # checkpoint arena and reverse-pass workspace kept from the first) gives
    # the first pair's objective and gradients bit for bit.
    st = setup
    coef = st["s"][:, None] * st["theta"][None, :]
    obj = dict(wss_weight=st["wss_w"], dp_weight=st["dp_w"], wss_mask=st["mask"], checkpoint_every=K)
    sol = make(st, prec)
    J1 = sol.adjoint_forward(T, coef, **obj)
    gc1, ga1 = sol.adjoint_backward()
    sol.set_populations(None)
    _, _, w = lbm.lattice(st["geom"].q)
    assert np.array_equal(sol.populations(), np.tile(w, (st["geom"].n, 1)))
    J2 = sol.adjoint_forward(T, coef, **obj)
    gc2, ga2 = sol.adjoint_backward()
    assert J1 == J2
    assert np.array_equal(gc1, gc2) and np.array_equal(ga1, ga2)
