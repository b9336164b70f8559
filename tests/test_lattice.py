"""Lattice tables (LatticeModel.cpp:92-191, :249-374) pinned by closed forms."""
import numpy as np
import pytest

import lbm
import oracle


@pytest.mark.parametrize("q", [19, 27])
def test_closed_forms(q):
    c, opp, w = lbm.lattice(q)
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(w @ c, 0.0, atol=1e-16)
    np.testing.assert_allclose((w[:, None, None] * c[:, :, None] * c[:, None, :]).sum(0), np.eye(3) / 3.0,
                               atol=1e-15)
    # fourth-order isotropy: sum w c_a^4 = 1/3, sum w c_a^2 c_b^2 = 1/9
    assert abs((w * c[:, 0] ** 4).sum() - 1.0 / 3.0) < 1e-15
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1.0 / 9.0) < 1e-15
    assert (opp[opp] == np.arange(q)).all()
    assert (c[opp] == -c).all()
    assert len({tuple(v) for v in c}) == q


def test_reference_index_order():
    c19, opp19, w19 = lbm.lattice(19)
    assert tuple(c19[18]) == (0, 0, 0) and w19[18] == 12.0 / 36.0   # rest last
    assert tuple(c19[6]) == (1, 1, 0) and tuple(c19[17]) == (0, -1, -1)
    c27, opp27, w27 = lbm.lattice(27)
    assert tuple(c27[0]) == (0, 0, 0) and w27[0] == 8.0 / 27.0      # rest first
    assert tuple(c27[19]) == (1, 1, 1) and tuple(c27[26]) == (-1, 1, 1)
    assert list(opp19) == [1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18]


@pytest.mark.parametrize("q", [19, 27])
def test_product_and_oracle_tables_agree(q):
    a = lbm.lattice(q)
    b = oracle.lattice(q)
    for x, y in zip(a, b):
        assert (x == y).all()
