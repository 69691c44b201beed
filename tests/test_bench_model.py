"""The bytes bench.py's sweep_bytes_per_step model charges each kernel of a
step (per cell), against the algorithmic bytes of the step, B_alg = 64*S
(read the state twice more in stages 2 and 3 for the RK base, write it once
per stage: 16*S + 24*S + 24*S): no design may move fewer bytes than B_alg,
a 2D case whose x and y directions run as one fused kernel moves exactly
B_alg, and fusing a 3D case's x and y sweeps removes the rhs round trip
between them and the second pass over the stencil input (3*8*S bytes per
stage: rhs written and read, the input read once more)."""
import pytest

import bench


class _P:
    def __init__(self, D, nf, precision):
        self.num_dims, self.num_fluids, self.precision = D, nf, precision
        self.sys_size = 2 * nf + D + 1
        self.viscous, self.weno_Re_flux = False, False


@pytest.mark.parametrize("precision", [0, 1, 3])
@pytest.mark.parametrize("D", [2, 3])
def test_design_bytes_against_the_algorithmic_ones(D, precision):
    p = _P(D, 2, precision)
    S = p.sys_size
    sep = sum(bench.sweep_bytes_per_step(p).values())
    fused = sum(bench.sweep_bytes_per_step(p, fused=True).values())
    assert sep >= 64 * S and fused >= 64 * S
    if D == 2:
        assert fused == 64 * S
    else:
        assert sep - fused == 3 * 3 * 8 * S


def test_sweep_order():
    assert bench.sweep_order(_P(3, 2, 0)) == ["sweep_x", "sweep_y", "sweep_z"]
    assert bench.sweep_order(_P(3, 2, 3)) == ["sweep_z", "sweep_y", "sweep_x"]
    assert bench.sweep_order(_P(2, 2, 1)) == ["sweep_y", "sweep_x"]
    assert bench.sweep_order(_P(3, 2, 2)) == ["sweep_z", "sweep_y", "sweep_x"]
