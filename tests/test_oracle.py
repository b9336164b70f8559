"""Pins of the oracle itself (CPU only).

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong index set or sign) fails it: closed forms of the collision, Zou-He
identities, conservation in closed boxes, the plane-shear closed form of the
WSS formula and spot values of the inlet waveform.
"""
import numpy as np
import pytest

import lbm
import oracle


def _feq(c, w, rho, u):
    cu = u @ c.T
    usq = (u * u).sum(-1, keepdims=True)
    return w[None, :] * rho[:, None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)


def _box_reference(n=(6, 5, 7), periodic=(1, 1, 1), q=19, tau=0.9, wall_mode=1, force=(0, 0, 0)):
    nx, ny, nz = n
    ijk = np.array([(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)], np.int32)
    kind = np.zeros(len(ijk), np.uint8)
    oid = -np.ones(len(ijk), np.int32)
    return oracle.Reference(q, n, ijk, kind, oid, np.zeros((0, 3)), (0, 0, 0), np.zeros(0, np.int32),
                            np.zeros((0, 3)), tau, 1.0, wall_mode=wall_mode, inlet_mode=2, outlet_mode=2,
                            periodic=periodic, force=force)


@pytest.mark.parametrize("q", [19, 27])
def test_equilibrium_is_a_fixed_point_and_moments(q):
    rng = np.random.default_rng(7)
    ref = _box_reference(q=q)
    c, _, w = oracle.lattice(q)
    rho = 1.0 + 0.1 * rng.uniform(-1, 1, ref.n)
    u = 0.05 * rng.uniform(-1, 1, (ref.n, 3))
    f = _feq(c.astype(float), w, rho, u)
    ref.set_f(f)
    ref.collide()
    np.testing.assert_allclose(ref.fs(), f, rtol=0, atol=1e-15)
    feq = ref.feq()
    np.testing.assert_allclose(feq.sum(1), rho, rtol=1e-14)
    np.testing.assert_allclose(feq @ c, rho[:, None] * u, atol=1e-15)


def test_bgk_relaxes_non_equilibrium_by_one_minus_omega():
    rng = np.random.default_rng(8)
    ref = _box_reference(tau=0.7)
    c, _, w = oracle.lattice(19)
    f = w[None, :] * (1 + 0.1 * rng.uniform(-1, 1, (ref.n, 19)))
    ref.set_f(f)
    ref.collide()
    feq = ref.feq()
    np.testing.assert_allclose(ref.fs() - feq, (1 - 1 / 0.7) * (f - feq), atol=1e-15)
    rho = f.sum(1)
    u = (f @ c) / rho[:, None]
    np.testing.assert_allclose(feq, _feq(c.astype(float), w, rho, u), atol=1e-15)


@pytest.mark.parametrize("q,wall_mode", [(19, 1), (27, 1), (19, 0)])
def test_closed_periodic_box_conserves_mass_and_momentum(q, wall_mode):
    rng = np.random.default_rng(9)
    ref = _box_reference(q=q, wall_mode=wall_mode)
    _, _, w = oracle.lattice(q)
    f = w[None, :] * (1 + 0.2 * rng.uniform(-1, 1, (ref.n, q)))
    c, _, _ = oracle.lattice(q)
    ref.set_f(f)
    m0, p0 = f.sum(), (f @ c).sum(0)
    ref.step(np.zeros(200))
    f1 = ref.f()
    assert abs(f1.sum() - m0) < 1e-11 * m0
    np.testing.assert_allclose((f1 @ c).sum(0), p0, atol=1e-11)


def test_bounce_back_box_conserves_mass():
    rng = np.random.default_rng(10)
    ref = _box_reference(periodic=(0, 0, 0), wall_mode=1)
    _, _, w = oracle.lattice(19)
    f = w[None, :] * (1 + 0.2 * rng.uniform(-1, 1, (ref.n, 19)))
    ref.set_f(f)
    ref.step(np.zeros(100))
    assert abs(ref.f().sum() - f.sum()) < 1e-11 * f.sum()


def test_guo_force_adds_momentum_linearly():
    g = np.array([1e-5, -2e-5, 3e-5])
    ref = _box_reference(force=g)
    c, _, _ = oracle.lattice(19)
    ref.step(np.zeros(50))
    f = ref.f()
    # each step adds F per node to the momentum (Guo scheme, uniform state)
    np.testing.assert_allclose((f @ c).sum(0), 50 * g * ref.n, rtol=1e-9)


@pytest.mark.parametrize("normal", [(0, 0, -1), (1, 0, 0), (-1, 0, 0)])
def test_zou_he_identities(normal):
    rng = np.random.default_rng(11)
    c, _, w = oracle.lattice(19)
    for _ in range(20):
        f = w * (1 + 0.1 * rng.uniform(-1, 1, 19))
        U = 0.05 * rng.uniform(-1, 1)
        g = oracle.zouhe(f, U, normal)
        a = int(np.nonzero(normal)[0][0])
        s = normal[a]
        tangential = [q for q in range(19) if c[q, a] == 0]
        outgoing = [q for q in range(19) if c[q, a] == s]
        incoming = [q for q in range(19) if c[q, a] == -s]
        # rho = (sum tangential + 2 sum outgoing) / (1 - u_a) on a low face,
        # / (1 + u_a) on a high face, u_a = U the velocity component imposed
        rhoN = (f[tangential].sum() + 2 * f[outgoing].sum()) / (1 - U if s < 0 else 1 + U)
        # only the populations entering the domain were rewritten
        untouched = [q for q in range(19) if q not in incoming]
        assert (g[untouched] == f[untouched]).all()
        assert abs(g.sum() - rhoN) < 1e-14
        mom = g @ c
        if a == 2:
            assert abs(mom[2] - rhoN * U) < 1e-14 and abs(mom[0]) < 1e-14 and abs(mom[1]) < 1e-14
        else:
            assert abs(mom[0] - rhoN * U) < 1e-14 and abs(mom[1]) < 1e-14 and abs(mom[2]) < 1e-14


def test_zou_he_reference_quirks_are_kept():
    # +y/-y branches use a hard-coded velocity of 1.0 (Boundary.cpp:178,206)
    c, _, w = oracle.lattice(19)
    f = w.copy()
    g1 = oracle.zouhe(f, 0.3, (0, 1, 0))
    g2 = oracle.zouhe(f, 0.7, (0, 1, 0))
    assert (g1 == g2).all()


def test_wss_plane_shear_closed_form():
    # f_neq = d w c_x c_y with normal +y: tau = C d sum(w c_x^2 c_y^2) e_x = C d / 9 e_x
    c, _, w = oracle.lattice(19)
    d, C = 1e-3, 2.5
    feq = w.copy()
    f = feq + d * w * c[:, 0] * c[:, 1]
    tau = oracle.wss_tau(f, feq, (0.0, 1.0, 0.0), C)
    np.testing.assert_allclose(tau, [C * d / 9.0, 0.0, 0.0], atol=1e-18)
    # a normal stress component (f_neq ~ c_y^2) has no tangential part
    f2 = feq + d * w * (c[:, 1] ** 2 - 1.0 / 3.0)
    np.testing.assert_allclose(oracle.wss_tau(f2, feq, (0.0, 1.0, 0.0), C), 0.0, atol=1e-18)
    # rotate the normal: tau stays perpendicular to it
    n = np.array([1.0, 2.0, -0.5])
    n /= np.linalg.norm(n)
    t = oracle.wss_tau(f, feq, n, C)
    assert abs(t @ n) < 1e-18


def test_inlet_waveform_spot_values():
    p = 0.9
    assert oracle.inlet_velocity(0.0, p) == 0.1423
    x = 0.05
    eq1 = (-8.579e+08 * x**8 + 4.059e+08 * x**7 - 7.817e+07 * x**6 + 7.994e+06 * x**5 - 4.835e+05 * x**4
           + 1.782e+04 * x**3 - 321.6 * x**2 + 3.215 * x + 0.1423)
    assert abs(oracle.inlet_velocity(x, p) - eq1) < 1e-12
    x = 0.5
    eq3 = -5.123 * x**4 + 14.15 * x**3 - 14.21 * x**2 + 5.865 * x - 0.5696
    assert abs(oracle.inlet_velocity(x, p) - eq3) < 1e-12
    assert oracle.inlet_velocity(0.85, p) == 0.0                    # last 10% of the cycle
    assert abs(oracle.inlet_velocity(x + p, p) - eq3) < 1e-9         # periodic in t


def test_oracle_runs_the_tube_without_nan(tube_geometry):
    ref = oracle.Reference.from_geometry(tube_geometry, tau=0.8, dt=1e-5)
    ref.step(np.full(50, 0.05), nthreads=4)
    f = ref.f()
    assert np.isfinite(f).all()
    _, kind, _ = tube_geometry.nodes()
    rho = f.sum(1)
    assert 0.9 < rho.min() and rho.max() < 1.5
    # the inlet nodes carry the imposed speed after Zou-He
    c, _, _ = lbm.lattice(19)
    inl = kind == lbm.KIND_INLET
    uz = (f[inl] @ c[:, 2]) / rho[inl]
    assert np.median(uz) == pytest.approx(0.05, rel=1e-9)


@pytest.mark.parametrize("modes", [(0, 0, 0), (1, 1, 1)], ids=["reference rules", "bounce-back + NEE"])
def test_oracle_threads_do_not_change_results(tube_geometry, modes):
    # Every per-node loop of the oracle step writes only its own node, so the
    # threaded step must equal the serial one bit for bit.
    wall, inlet, outlet = modes
    out = []
    for nthreads in (1, 8):
        ref = oracle.Reference.from_geometry(tube_geometry, tau=0.8, dt=1e-5, wall_mode=wall, inlet_mode=inlet,
                                             outlet_mode=outlet)
        ref.step(np.full(30, 0.05), nthreads=nthreads)
        out.append((ref.f(), ref.wss()))
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
