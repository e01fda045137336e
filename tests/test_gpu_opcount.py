"""Static operation count (GPU): the counting type reproduces hand counts of small functions."""
import pytest

import cfdx

pytestmark = pytest.mark.gpu


def test_op_count_hand_counts():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    c = cfdx.Ctx(0).op_count(cfdx.default_params())
    # prim_to_q: ke = 0.5 rho (u u + v v + w w): 5 mul 2 add; rho u, rho v, rho w: 3 mul; p 2.5 + ke: 1 mul 1 add
    assert (c["cons"]["add"], c["cons"]["mul"], c["cons"]["total"]) == (3, 9, 12)
    # node flux of the sweeps: E + p from the conserved energy 1 add; V 3 mul 2 add; rho V 1; f0 1 mul;
    # f1..f3 3 x (3 mul 1 add); f4 2 mul
    assert (c["node_flux"]["add"], c["node_flux"]["mul"], c["node_flux"]["total"]) == (6, 16, 22)
    # one face: 5 fields x 6 points x 5 projection terms alone are 150 operations; two interface
    # states, Roe average, LLF and the 5 midpoint corrections take the rest
    assert 1000 < c["face"]["total"] < 2500 and c["face"]["sqrt"] >= 4
    assert c["viscous"]["cbrt"] == 1 and c["viscous"]["total"] > 200
