"""Tetrahedral quadrature used by the reference element matrices (oracle copy).

Restates ``petgem/hvfem.py:1055-1127`` (the symmetric rules of order 2, 4
and 6 -- the only ones used for p = 1..3, since the reference integrates with
order 2p, ``hvfem.py:251``) and the mapping of ``hvfem.py:1603-1608`` from
the [-1,1]^3 reference simplex onto the unit tetrahedron.

The 15-digit literals are reference data and are kept verbatim: the element
matrices are compared with the reference to 1e-12, and a rule with its own
(more accurate) digits would differ from it at ~1e-15, which is harmless,
but keeping the same literals makes the comparison exact up to summation
order.

Pinned by: weights sum to the unit-tet volume 1/6; each rule integrates
monomials x^a y^b z^c up to its order exactly (closed form
a! b! c! / (a+b+c+3)!), see ``tests/test_oracle_basis.py``.
"""
import numpy as np

# rows: (t0, t1, t2, w) on the simplex {t_i >= -1, t0+t1+t2 <= -1}
_RULE_ORDER2 = [
    [-0.723606797749979, -0.723606797749979, -0.723606797749979, 0.333333333333333],
    [0.170820393249937, -0.723606797749979, -0.723606797749979, 0.333333333333333],
    [-0.723606797749979, 0.170820393249937, -0.723606797749979, 0.333333333333333],
    [-0.723606797749979, -0.723606797749979, 0.170820393249937, 0.333333333333333],
]

_RULE_ORDER4 = [
    [-0.500000000000000, -0.500000000000000, -0.500000000000000, -0.105244444444444],
    [-0.857142857142857, -0.857142857142857, -0.857142857142857, 0.060977777777778],
    [-0.857142857142857, -0.857142857142857, 0.571428571428571, 0.060977777777778],
    [-0.857142857142857, 0.571428571428571, -0.857142857142857, 0.060977777777778],
    [0.571428571428571, -0.857142857142857, -0.857142857142857, 0.060977777777778],
    [-0.201192847666402, -0.201192847666402, -0.798807152333598, 0.199111111111111],
    [-0.201192847666402, -0.798807152333598, -0.201192847666402, 0.199111111111111],
    [-0.798807152333598, -0.201192847666402, -0.201192847666402, 0.199111111111111],
    [-0.201192847666402, -0.798807152333598, -0.798807152333598, 0.199111111111111],
    [-0.798807152333598, -0.201192847666402, -0.798807152333598, 0.199111111111111],
    [-0.798807152333598, -0.798807152333598, -0.201192847666402, 0.199111111111111],
]

_A6, _B6, _W6a = -0.570794257481696, -0.287617227554912, 0.053230333677557
_C6, _D6, _W6b = -0.918652082930777, 0.755956248792332, 0.013436281407094
_E6, _F6, _W6c = -0.355324219715449, -0.934027340853653, 0.073809575391540
_G6, _H6, _I6, _W6d = -0.872677996249965, -0.460655337083368, 0.206011329583298, 0.064285714285714

_RULE_ORDER6 = [
    [_A6, _A6, _A6, _W6a], [_B6, _A6, _A6, _W6a], [_A6, _B6, _A6, _W6a], [_A6, _A6, _B6, _W6a],
    [_C6, _C6, _C6, _W6b], [_D6, _C6, _C6, _W6b], [_C6, _D6, _C6, _W6b], [_C6, _C6, _D6, _W6b],
    [_E6, _E6, _E6, _W6c], [_F6, _E6, _E6, _W6c], [_E6, _F6, _E6, _W6c], [_E6, _E6, _F6, _W6c],
    [_G6, _G6, _H6, _W6d], [_G6, _H6, _G6, _W6d], [_G6, _G6, _I6, _W6d], [_G6, _I6, _G6, _W6d],
    [_G6, _H6, _I6, _W6d], [_G6, _I6, _H6, _W6d], [_H6, _G6, _G6, _W6d], [_H6, _G6, _I6, _W6d],
    [_H6, _I6, _G6, _W6d], [_I6, _G6, _H6, _W6d], [_I6, _G6, _G6, _W6d], [_I6, _H6, _G6, _W6d],
]

_RULES = {2: _RULE_ORDER2, 4: _RULE_ORDER4, 6: _RULE_ORDER6}


def gauss_points(order):
    """Points (Q,3) on the unit tetrahedron and weights (Q,) for a rule of ``order``.

    Follows ``hvfem.py:1603-1608``: xi = (1+t1)/2, eta = -(1+t0+t1+t2)/2,
    zeta = (1+t0)/2, w = t3/8.
    """
    if order not in _RULES:
        raise ValueError(f"quadrature order {order} not supported (p <= 3 uses 2, 4, 6)")
    t = np.asarray(_RULES[order], dtype=np.float64)
    w = t[:, 3] / 8.0
    pts = np.empty((t.shape[0], 3), dtype=np.float64)
    pts[:, 0] = (1.0 + t[:, 1]) / 2.0
    pts[:, 1] = -(1.0 + np.sum(t[:, 0:3], axis=1)) / 2.0
    pts[:, 2] = (1.0 + t[:, 0]) / 2.0
    return pts, w


def rule_for_order(p):
    """Quadrature used for Nedelec order ``p`` (order 2p, ``hvfem.py:251``)."""
    return gauss_points(2 * p)
