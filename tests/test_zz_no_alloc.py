"""lbm_reduce_noalloc: the reduction over the static device workspace,
against lbm_reduce over the solver's own workspace, bit for bit, on two
calls in a row (the second on a smaller solver), for both kernel families."""
import ctypes as C

import numpy as np
import pytest

import lbm
import workloads

pytestmark = pytest.mark.gpu


def _both(sol):
    a, b = lbm.Stats(), lbm.Stats()
    lbm._check(lbm.lib().lbm_reduce(sol._h, C.byref(a)))
    lbm._check(lbm.lib().lbm_reduce_noalloc(sol._h, C.byref(b)))
    return a, b


def _fields(s):
    return [s.mass, *s.momentum, s.wss_l2, s.wss_max, s.iteration, s.time]


@pytest.mark.parametrize("family", ["reference", "columns"])
def test_noalloc_reduce_equals_reduce(tmp_path, family):
    kw = {} if family == "reference" else dict(wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                                               outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0,
                                               precision=lbm.PREC_F32)
    for i, (radius, length) in enumerate(((9.0, 60.0), (4.0, 12.0))):  # the second call on a smaller input
        cfg = workloads.axis_tube(str(tmp_path / str(i)), axis="z", radius=radius, length=length)
        geom = lbm.Geometry.from_config(cfg)
        sol = lbm.Solver(geom, tau=0.8, dt=1e-5, cte_vel=0.03, **kw)
        sol.step(7)
        a, b = _both(sol)
        print(f"{family} {geom.n} nodes: mass {a.mass!r}, wss_l2 {a.wss_l2!r}", flush=True)
        assert _fields(a) == _fields(b)
        assert a.mass > 0 and a.wss_l2 > 0
        sol.close()
