"""config.txt parser (LBM.h:81-188) on the reference's own config.txt.

The expected values are read off the file by hand and from the parser's
rules (a block ends at the first line that is empty before trimming,
LBM.h:163; the blank-only line before `equations:` is then skipped by the
outer loop, so the three equations are recorded).
"""
import math
import os

import pytest

import lbm
from conftest import GOLDEN

REF_CFG = os.path.join(GOLDEN, "reference_config.txt")


def test_reference_config_values():
    c = lbm.parse_config(REF_CFG)
    assert c.stl_file == b"JT.stl"
    assert c.flow_type == b"Constant"
    assert c.dt == 1e-5 and c.rho == 1060 and c.nu_blood == 3.5e-6 and c.tau == 0.68
    assert c.dx == 0.025 and c.pressure == 1 and c.cte_vel == 0.1 and c.reynolds_number == 100
    assert c.num_out == 10 and c.max_iter == 30001 and c.iter_to_write == 10000
    assert c.systole == 0.15 and c.diastole == 0.9 and c.num_cycle == 5
    assert c.n_inlets == 1 and c.n_outlets == 4
    assert c.n_equations == 3
    i0 = c.inlets[0]
    assert i0.plane == b"z" and i0.coordinate == -178.5 and list(i0.normal) == [0, 0, -1]
    assert list(i0.range_x) == [-math.inf, math.inf]
    o = c.outlets
    assert o[0].plane == b"x" and o[0].coordinate == 15.0 and list(o[0].normal) == [1, 0, 0]
    assert o[1].coordinate == -0.800001 and list(o[1].normal) == [-1, 0, 0]
    assert o[2].plane == b"y" and o[2].coordinate == -188.009
    assert list(o[2].range_x) == [1.0, 3.5] and list(o[2].range_z) == [-162.0, -159.0]
    assert list(o[3].range_x) == [7.0, 9.0] and list(o[3].range_z) == [-157, -155]
    assert list(o[3].range_y) == [-math.inf, math.inf]


def test_block_ends_only_at_truly_empty_line(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("parameters:\n  MaxIter: 7\n\ninlets:\n\n    plane: z\n    coordinate: 1.5\n"
                 "    normal: [0, 0, -1]\n   \noutlets:\n    plane: x\n    coordinate: 2\n"
                 "    normal: [1, 0, 0]\n\nequations:\n  E1: [1, 2,3]\n")
    c = lbm.parse_config(str(p))
    # the blank-only line does not end the inlet block, so `outlets:` and the
    # outlet keys are read into it; the last normal wins.
    assert c.max_iter == 7 and c.n_inlets == 1 and c.n_outlets == 0
    assert c.inlets[0].plane == b"x" and c.inlets[0].coordinate == 2.0 and list(c.inlets[0].normal) == [1, 0, 0]
    assert c.n_equations == 1


def test_range_parsing_and_numcycle(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text('parameters:\n  NumCycle: 2.9\n  FlowType: "Pulsatile"\n  stl_file: "a b.stl"\n\n'
                 "outlets:\n    plane: y\n    coordinate: -3\n    range_x: [ 1 ,2,]\n    normal: [0, 1, 0] # c\n")
    c = lbm.parse_config(str(p))
    assert c.num_cycle == 2 and c.flow_type == b"Pulsatile" and c.stl_file == b"a b.stl"
    assert list(c.outlets[0].range_x) == [1.0, 2.0] and c.outlets[0].normal_len == 3


def test_missing_file_and_bad_number(tmp_path):
    with pytest.raises(lbm.LbmError) as e:
        lbm.parse_config(str(tmp_path / "nope.txt"))
    assert e.value.code == lbm.ERR_IO
    p = tmp_path / "bad.txt"
    p.write_text("parameters:\n  MaxIter: abc\n")
    with pytest.raises(lbm.LbmError) as e:
        lbm.parse_config(str(p))
    assert e.value.code == lbm.ERR_ARG
