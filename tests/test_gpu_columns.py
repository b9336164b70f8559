"""Column-offset kernel family (half-way bounce-back walls) on the GPU.

These boundary modes have no reference counterpart; they are checked against
the oracle's extension modes (same algorithm, push form, FP64) and against
closed forms: periodic Poiseuille flow under a body force, mass conservation.
The FP64 GPU path forms FMAs in the equilibrium, so agreement with the oracle
is to rounding (1e-12 on populations of order w_q), not bitwise.
"""
import numpy as np
import pytest

import lbm
import oracle
import workloads

pytestmark = pytest.mark.gpu

TAU, DT, U = 0.8, 1e-5, 0.04


@pytest.mark.parametrize("inlet,outlet", [(lbm.INLET_NEE_VELOCITY, lbm.OUTLET_NEE_PRESSURE),
                                          (lbm.INLET_REFERENCE_ZOUHE, lbm.OUTLET_REFERENCE)])
def test_bounce_back_family_matches_oracle(tube_geometry, inlet, outlet):
    g = tube_geometry
    sol = lbm.Solver(g, tau=TAU, dt=DT, cte_vel=U, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=inlet,
                     outlet_mode=outlet, parity_checks=0)
    ref = oracle.Reference.from_geometry(g, tau=TAU, dt=DT, wall_mode=1, inlet_mode=inlet, outlet_mode=outlet)
    prev = 0
    for n in (1, 50, 300):
        sol.step(n - prev)
        ref.step(np.full(n - prev, U), nthreads=8)
        prev = n
        d = np.max(np.abs(sol.populations() - ref.f()))
        print(f"step {n}: max |f - f_oracle| = {d:.3e}", flush=True)
        assert d < 1e-12
    w, acc, mag = sol.wss()
    rw, racc, rmag = ref.wss()
    assert np.max(np.abs(w - rw)) <= 1e-9 * np.max(np.abs(rw))
    assert np.max(np.abs(mag - rmag)) <= 1e-9 * np.max(np.abs(rmag))


def test_set_get_round_trip_and_init(tube_geometry):
    rng = np.random.default_rng(5)
    _, _, w = lbm.lattice(19)
    f0 = w[None, :] * (1 + 0.05 * rng.uniform(-1, 1, (tube_geometry.n, 19)))
    sol = lbm.Solver(tube_geometry, tau=TAU, dt=DT, cte_vel=U, wall_mode=lbm.WALL_HALFWAY_BB,
                     inlet_mode=lbm.INLET_NEE_VELOCITY, outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0)
    sol.set_populations(f0)
    assert np.array_equal(sol.populations(), f0)
    ref = oracle.Reference.from_geometry(tube_geometry, tau=TAU, dt=DT, wall_mode=1, inlet_mode=1, outlet_mode=1)
    ref.set_f(f0)
    sol.step(3)
    ref.step(np.full(3, U))
    assert np.max(np.abs(sol.populations() - ref.f())) < 1e-13


def _wide_vs_32bit(g, prec, steps):
    rng = np.random.default_rng(9)
    _, _, w = lbm.lattice(19)
    f0 = w[None, :] * (1 + 0.05 * rng.uniform(-1, 1, (g.n, 19)))
    out = []
    for wide in (False, True):
        sol = lbm.Solver(g, precision=prec, tau=TAU, dt=DT, cte_vel=U, wall_mode=lbm.WALL_HALFWAY_BB,
                         inlet_mode=lbm.INLET_NEE_VELOCITY, outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0)
        sol.set_wide_offsets(wide)
        sol.set_populations(f0)
        sol.step(steps)
        out.append((sol.post_collision(), sol.wss()[2]))
        sol.close()
    return out, w


@pytest.mark.parametrize("prec", [lbm.PREC_F32, lbm.PREC_F64])
def test_wide_offsets_match_32bit_offsets(tube_geometry, prec):
    # States of 4 GiB or more (the ~300M-node share) take 64-bit element
    # offsets; on a small case that path must give the same bits.
    out, _ = _wide_vs_32bit(tube_geometry, prec, 40)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_wide_offsets_fp16_within_one_storage_ulp(tube_geometry):
    # FP16 storage with FP32 arithmetic: the two instantiations fuse the
    # half-to-float conversions and multiply-adds differently, and a result
    # within an FP32 rounding of an FP16 tie stores one FP16 ulp apart
    # (measured: 13 of 475,551 values after one step). One step, every stored
    # value within one FP16 ulp, at most 0.1% of them different.
    out, w = _wide_vs_32bit(tube_geometry, lbm.PREC_F16, 1)
    a, b = out[0][0], out[1][0]
    g = np.abs(a - w[None, :]).astype(np.float16)
    ulp = np.spacing(g).astype(np.float64)
    d = np.abs(a - b)
    assert np.all(d <= ulp * 1.0001)
    assert np.count_nonzero(d) <= 1e-3 * d.size


def _poiseuille(prec, steps, R=8.0, L=12, g=2e-6, tau=0.8):
    lab, origin, (ax, ay) = workloads.tube_mask(R, L)
    geom = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1), normals=False)
    sol = lbm.Solver(geom, tau=tau, dt=DT, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NONE,
                     outlet_mode=lbm.OUTLET_NONE, precision=prec, parity_checks=0, force=(0.0, 0.0, g),
                     wss_every=0)
    s0 = sol.reduce()
    sol.step(steps)
    return geom, sol, s0, (ax, ay), R, g, tau


def test_periodic_poiseuille_closed_form():
    geom, sol, s0, (ax, ay), R, g, tau = _poiseuille(lbm.PREC_F64, 6000)
    ijk, _, _ = geom.nodes()
    rho, u = sol.macroscopic()
    # Guo forcing: the physical velocity is (j + F/2) / rho
    uz = u[:, 2] + 0.5 * g / rho
    r2 = (ijk[:, 0] - ax) ** 2 + (ijk[:, 1] - ay) ** 2
    nu = (tau - 0.5) / 3.0
    # Fit away from the staircase wall (which bends the profile within ~2
    # nodes of it): u(r) = g / (4 nu) (R^2 - r^2), curvature fixed by g, nu.
    core = r2 < (R - 2.5) ** 2
    A = np.stack([np.ones_like(r2), r2], 1)[core]
    coef, *_ = np.linalg.lstsq(A, uz[core], rcond=None)
    assert coef[1] == pytest.approx(-g / (4 * nu), rel=0.02)
    fit = A @ coef
    assert np.sqrt(np.mean((fit - uz[core]) ** 2)) < 0.01 * uz.max()
    # flow is axial and the density uniform
    assert np.max(np.abs(u[:, :2])) < 1e-3 * uz.max()
    s1 = sol.reduce()
    assert abs(s1.mass - s0.mass) < 1e-10 * s0.mass


def test_periodic_poiseuille_matches_oracle_early():
    lab, origin, _ = workloads.tube_mask(5.0, 6)
    geom = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1), normals=False)
    g = 1e-5
    sol = lbm.Solver(geom, tau=0.9, dt=DT, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NONE,
                     outlet_mode=lbm.OUTLET_NONE, parity_checks=0, force=(0.0, 0.0, g), wss_every=0)
    ijk, kind, oid = geom.nodes()
    ref = oracle.Reference(19, tuple(geom.info.grid), ijk, kind, oid, np.zeros((0, 3)), (0, 0, 0),
                           np.zeros(0, np.int32), np.zeros((0, 3)), 0.9, 1.0, wall_mode=1, inlet_mode=2,
                           outlet_mode=2, periodic=(0, 0, 1), force=(0, 0, g))
    sol.step(200)
    ref.step(np.zeros(200))
    assert np.max(np.abs(sol.populations() - ref.f())) < 1e-13


@pytest.mark.parametrize("prec,tol", [(lbm.PREC_F32, 1e-4), (lbm.PREC_F16, 5e-2)])
def test_reduced_precision_poiseuille(prec, tol):
    _, s64, _, _, _, _, _ = _poiseuille(lbm.PREC_F64, 1500)
    _, sp, _, _, _, _, _ = _poiseuille(prec, 1500)
    _, u64 = s64.macroscopic()
    _, up = sp.macroscopic()
    err = np.linalg.norm(up - u64) / np.linalg.norm(u64)
    print(f"precision {prec}: velocity rel L2 vs FP64 = {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("prec,tol", [(lbm.PREC_F64, 1e-9), (lbm.PREC_F32, 2e-5)])
def test_brinkman_damping_momentum_budget_and_darcy_limit(prec, tol):
    # Fully periodic box, uniform porosity alpha, body force F along z. The
    # state stays uniform, and the Guo scheme with Brinkman damping changes
    # the momentum per step by F_tot = F - alpha m, with the physical
    # momentum m = (j + F/2) / (1 + alpha/2). So j_(n+1) = j_n + F - alpha m_n,
    # which relaxes to the Darcy limit m = F / alpha (u = F / (alpha rho)).
    # (The collision's own momentum change, omega (m - j) = omega F_tot / 2,
    # and the forcing term (1 - omega/2) F_tot add up to F_tot for any tau.)
    # Tolerances relative to F: FP64 rounding of the host-side momentum sums
    # (~1e-16 on populations of order w_q); FP32 arithmetic on increments of
    # order F.
    alpha, g = 2.0 ** -7, 1e-5  # alpha exactly representable in the FP32 porosity input
    geom = lbm.Geometry.from_labels(np.ones((6, 6, 6), np.uint8), (0.0, 0.0, 0.0), 1.0, periodic=(1, 1, 1),
                                    normals=False)
    sol = lbm.Solver(geom, tau=0.8, dt=DT, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NONE,
                     outlet_mode=lbm.OUTLET_NONE, precision=prec, parity_checks=0, force=(0.0, 0.0, g), wss_every=0)
    sol.set_porosity(np.full(geom.n, alpha))

    def jz():
        rho, u = sol.macroscopic()
        j = rho * u[:, 2]
        assert np.ptp(j) <= tol * g  # uniform state
        return float(j.mean())

    j = jz()
    for n in range(5):
        sol.step(1)
        m = (j + 0.5 * g) / (1.0 + 0.5 * alpha)
        expect = j + g - alpha * m
        j = jz()
        print(f"step {n + 1}: j = {j:.6e}, budget {expect:.6e}", flush=True)
        assert abs(j - expect) <= tol * g
    sol.step(4000)  # (1 - alpha / (1 + alpha / 2))^4000 ~ 3e-14
    m = (jz() + 0.5 * g) / (1.0 + 0.5 * alpha)
    assert m == pytest.approx(g / alpha, rel=100 * tol)
