"""Known answers for the oracle's stale-link, wall-copy and outlet rules.

Tiny node sets whose outcome follows by hand from the reference rules, so a
plausible slip (a reversed q loop, the copy written the wrong way round, the
upstream node taken on the wrong side, the rules applied in another order)
fails an equality:

* Streaming (Simulation.cpp:238-263, Node.cpp:57-61) only writes a
  population that arrives over an existing link; every other population of
  the target keeps its previous value.
* The reference wall rule (Simulation.cpp:746-763) runs q = 0 .. Q-1 in
  order and sets f[opp q] = f[q] wherever the link q is missing. On a pair
  (q, opp q) with both links missing, the lower index comes first, so both
  members end up with the lower member's value.
* The outlet rule (Simulation.cpp:650-741) copies every population from the
  node one step against the outlet normal, after streaming and before the
  wall rule.
* A wall node collides with u = 0 (Simulation.cpp:92-95), so its f_eq is
  w_q rho.
"""
import numpy as np
import pytest

import oracle

TAU = 0.8


def _pairs(q):
    _, opp, _ = oracle.lattice(q)
    return sorted({(min(a, int(b)), max(a, int(b))) for a, b in enumerate(opp) if a != b})


def _rest(q):
    c, _, _ = oracle.lattice(q)
    return int(np.flatnonzero(np.all(c == 0, axis=1))[0])


def _solver(q, ijk, kind, outlet_id, outlet_normals=np.zeros((0, 3))):
    return oracle.Reference(q, tuple(int(v) + 1 for v in np.max(ijk, axis=0)), np.asarray(ijk), np.asarray(kind),
                            np.asarray(outlet_id), outlet_normals, (0.0, 0.0, -1.0), np.zeros(0, np.int32),
                            np.zeros((0, 3)), TAU, 1.0, wall_mode=0, inlet_mode=0, outlet_mode=0)


def _random_f(n, q, seed):
    _, _, w = oracle.lattice(q)
    rng = np.random.default_rng(seed)
    return w[None, :] * (1 + 0.3 * rng.uniform(-1, 1, (n, q)))


@pytest.mark.parametrize("q", [19, 27])
def test_lone_wall_node(q):
    # One wall node with no neighbour: every link but the rest one is missing.
    ref = _solver(q, [[0, 0, 0]], [1], [-1])
    _, _, w = oracle.lattice(q)
    r = _rest(q)
    f = _random_f(1, q, 3)[0]
    ref.set_f(f[None, :])
    for step in range(3):
        ref.step([0.0])
        rho = sum(f)  # same order as the reference sum, q = 0 .. Q-1
        # collision with u = 0: f* = f + (w rho - f) / tau
        assert np.allclose(ref.fs()[0], f + (1.0 / TAU) * (w * rho - f), rtol=0, atol=1e-15)
        expect = f.copy()
        expect[r] = ref.fs()[0][r]  # the rest population streams onto the node itself
        for lo, hi in _pairs(q):
            expect[lo] = f[lo]  # nothing arrives: stale value, then the lower index wins the pair copy
            expect[hi] = f[lo]
        got = ref.f()[0]
        assert np.array_equal(got, expect), f"step {step + 1}"
        f = got


def test_outlet_takes_every_population_of_the_upstream_node():
    # A column along z: fluid, fluid, outlet (normal +z). The outlet's upstream
    # node is the one below it.
    ijk = [[0, 0, 0], [0, 0, 1], [0, 0, 2]]
    ref = _solver(19, ijk, [0, 0, 3], [-1, -1, 0], np.array([[0.0, 0.0, 1.0]]))
    ref.set_f(_random_f(3, 19, 4))
    for step in range(4):
        ref.step([0.0])
        f = ref.f()
        assert np.array_equal(f[2], f[1]), f"step {step + 1}"
        assert not np.array_equal(f[1], f[0])


def test_outlet_copy_precedes_the_wall_rule():
    # fluid, wall, outlet along z: the outlet copies the wall node's streamed
    # populations, then the wall rule rewrites the wall node's missing links.
    ijk = [[0, 0, 0], [0, 0, 1], [0, 0, 2]]
    ref = _solver(19, ijk, [0, 1, 3], [-1, -1, 0], np.array([[0.0, 0.0, 1.0]]))
    ref.set_f(_random_f(3, 19, 5))
    c, _, _ = oracle.lattice(19)
    present = {qq for qq in range(19) if c[qq][0] == 0 and c[qq][1] == 0}  # +z, -z, rest
    for step in range(3):
        ref.step([0.0])
        f = ref.f()
        streamed = f[2]  # the outlet holds the wall node's populations before its wall rule
        expect = streamed.copy()
        for lo, hi in _pairs(19):
            if lo in present:
                continue
            expect[hi] = streamed[lo]
        assert np.array_equal(f[1], expect), f"step {step + 1}"
        if step == 0:
            # on the first step the missing-link pairs still hold distinct
            # stale values, so the wall rule applied first would show here
            assert all(streamed[lo] != streamed[hi] for lo, hi in _pairs(19) if lo not in present)


def test_a_zero_post_collision_value_is_not_streamed():
    # Wall node A at x = 0, fluid node B at x = 1: B's population q = 0
    # (c = +x) arrives from A. With all of A's populations zero, A collides
    # to f* = 0 exactly (u = 0, rho = 0), and the reference pushes only
    # values other than 0.0 (Simulation.cpp:245-251): B keeps its old f_0.
    # With A nonzero, B's f_0 is A's f*_0.
    c, _, _ = oracle.lattice(19)
    assert tuple(c[0]) == (1, 0, 0)
    ijk = [[0, 0, 0], [1, 0, 0]]
    for a_zero in (True, False):
        ref = _solver(19, ijk, [1, 0], [-1, -1])
        f0 = _random_f(2, 19, 6)
        if a_zero:
            f0[0] = 0.0
        ref.set_f(f0)
        ref.step([0.0])
        fs_a = ref.fs()[0]
        got = ref.f()[1][0]
        if a_zero:
            assert np.all(fs_a == 0.0) and got == f0[1][0]
        else:
            assert fs_a[0] != 0.0 and got == fs_a[0]


@pytest.mark.parametrize("q", [19, 27])
def test_streaming_shifts_each_population_by_its_velocity(q):
    # Fully periodic 4 x 5 x 6 box of fluid nodes (reference order: i outer,
    # k inner): every link exists and push streaming is a pure shift,
    # f_q(x + c_q) = f*_q(x).
    g = (4, 5, 6)
    ijk = np.array([(i, j, k) for i in range(g[0]) for j in range(g[1]) for k in range(g[2])], np.int32)
    n = len(ijk)
    ref = oracle.Reference(q, g, ijk, np.zeros(n, np.uint8), np.full(n, -1, np.int32), np.zeros((0, 3)),
                           (0.0, 0.0, -1.0), np.zeros(0, np.int32), np.zeros((0, 3)), TAU, 1.0, periodic=(1, 1, 1))
    ref.set_f(_random_f(n, q, 7))
    ref.step([0.0])
    c, _, _ = oracle.lattice(q)
    fs, f = ref.fs(), ref.f()
    for qq in range(q):
        t = (ijk + c[qq]) % np.array(g)
        dst = (t[:, 0] * g[1] + t[:, 1]) * g[2] + t[:, 2]
        assert np.array_equal(f[dst, qq], fs[:, qq]), qq


def _moments(f, c):
    rho = f.sum()
    return rho, (f[:, None] * c).sum(0) / rho


def _feq(rho, u, c, w):
    cu = c @ u
    return w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


@pytest.mark.parametrize("q", [19, 27])
def test_non_equilibrium_extrapolation_identities(q):
    # A z column, periodic in x and y: inlet (normal -z), two fluid nodes,
    # outlet (normal +z). The extrapolated boundary nodes take the
    # non-equilibrium part of the node behind them (which has every link, so
    # its streamed populations are what the rule reads):
    #   inlet:  density of the node behind, velocity U along +z;
    #   outlet: the set density, velocity of the node behind.
    U, rho_out = 0.03, 1.002
    ijk = np.array([[0, 0, k] for k in range(4)], np.int32)
    ref = oracle.Reference(q, (1, 1, 4), ijk, np.array([2, 0, 0, 3], np.uint8), np.array([-1, -1, -1, 0], np.int32),
                           np.array([[0.0, 0.0, 1.0]]), (0.0, 0.0, -1.0), np.zeros(0, np.int32), np.zeros((0, 3)),
                           TAU, 1.0, wall_mode=1, inlet_mode=1, outlet_mode=1, outlet_rho=rho_out, periodic=(1, 1, 0))
    c, _, w = oracle.lattice(q)
    c = c.astype(np.float64)
    ref.set_f(_random_f(4, q, 8))
    for step in range(3):
        ref.step([U])
        f = ref.f()
        r1, u1 = _moments(f[1], c)
        r0, u0 = _moments(f[0], c)
        assert abs(r0 - r1) < 1e-15 and np.allclose(u0, [0.0, 0.0, U], rtol=0, atol=1e-15), step
        assert np.allclose(f[0] - _feq(r1, np.array([0.0, 0.0, U]), c, w), f[1] - _feq(r1, u1, c, w), rtol=0, atol=1e-15)
        r2, u2 = _moments(f[2], c)
        r3, u3 = _moments(f[3], c)
        assert abs(r3 - rho_out) < 1e-15 and np.allclose(u3, u2, rtol=0, atol=1e-15), step
        assert np.allclose(f[3] - _feq(rho_out, u2, c, w), f[2] - _feq(r2, u2, c, w), rtol=0, atol=1e-15)


def test_a_zero_rest_population_keeps_its_value():
    # A lone plain fluid node with f_rest = 3 and every other population 0,
    # at tau = 2/3 (1/tau = 1.5 exactly): rho = 3, u = 0, f_eq,rest =
    # w_rest * 3 = 1 and f*_rest = 3 + 1.5 (1 - 3) = 0.0 exactly, so the
    # rest population, whose link points at the node itself, is not pushed
    # and keeps 3 (Simulation.cpp:245-251); with no neighbours every other
    # population keeps its 0 as well. A push of zeros would leave f_rest = 0.
    q = 19
    r = _rest(q)
    ref = oracle.Reference(q, (1, 1, 1), np.zeros((1, 3), np.int32), np.zeros(1, np.uint8),
                           np.full(1, -1, np.int32), np.zeros((0, 3)), (0.0, 0.0, -1.0), np.zeros(0, np.int32),
                           np.zeros((0, 3)), 2.0 / 3.0, 1.0, wall_mode=0, inlet_mode=0, outlet_mode=0)
    f = np.zeros((1, q))
    f[0, r] = 3.0
    ref.set_f(f)
    ref.collide()
    fs = ref.fs()[0]
    assert fs[r] == 0.0 and np.all(fs[np.arange(q) != r] != 0.0)
    ref.step([0.0])
    assert np.array_equal(ref.f(), f)
