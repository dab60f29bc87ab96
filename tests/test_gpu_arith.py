"""The FP64 arithmetic the stage kernels rely on for bit parity.

The log-mean of the entropy-conserving flux amplifies the rounding of the
ratio a/b (and of beta = rho/(2p)) by 1/|a/b - 1|, so the kernels' division
div_rn (reciprocal refinement plus one Markstein correction, without the
range scaling of the library division) must return the IEEE-rounded
quotient.  The check generates operand pairs on the device (random
exponents in [-60, 60], half of them nearly equal) and compares bits."""
import pytest

from nemo_amd import _lib, solver


@pytest.mark.gpu
def test_division_and_reciprocal_match_ieee_bits():
    cfg = solver.make_config(N=3, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP,
                             timedisc=_lib.TD_LSRK54)
    n = 1 << 26
    print(f"arithmetic check, {n} pairs", flush=True)
    with solver.Solver(cfg) as s:
        div_bad, rcp_bad = s.check_arithmetic(n, seed=12345)
    print(f"  div_rn mismatches {div_bad}, rcp_nr mismatches {rcp_bad}")
    assert div_bad == 0
    # the reciprocal is used where about 1 ulp is enough; it is correctly
    # rounded nearly always
    assert rcp_bad <= n // 100000
