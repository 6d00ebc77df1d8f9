"""pic_gauss_residual_noalloc (no device allocation, a file-scope scratch of
2^21 doubles per device) against pic_gauss_residual through the C ABI: two
calls in a row, the second on a smaller grid, equal to FP64 round-off (both
accumulate the charge density with FP64 atomics); grids above the scratch
are rejected."""
import numpy as np
import pytest

from picongpu_amd import abi, setups

pytestmark = pytest.mark.gpu


def residuals(st, steps):
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    sim.step(steps)
    ref = sim.gauss_residual()
    got = sim.gauss_residual_noalloc()
    again = sim.gauss_residual_noalloc()
    sim.close()
    return ref, got, again


def test_noalloc_residual_matches_allocating_one():
    big = setups.thermal(grid=(64, 64, 64), ppc=4)
    small = setups.kelvin_helmholtz(grid=(32, 32, 16))
    for st, steps in ((big, 3), (small, 2)):
        print(f"case noalloc gauss residual {st.name} {st.grid}, {steps} steps", flush=True)
        ref, got, again = residuals(st, steps)
        scale = np.abs(ref).max()
        assert scale > 0
        assert np.abs(got - ref).max() <= 1e-12 * scale
        assert np.abs(again - ref).max() <= 1e-12 * scale


def test_noalloc_rejects_grids_beyond_its_scratch():
    st = setups.thermal(grid=(128, 128, 132), ppc=1)  # 2,162,688 cells > 2^21
    print(f"case noalloc rejection {st.grid}", flush=True)
    sim = abi.Simulation(st.config())
    with pytest.raises(abi.PicError, match="more than its scratch holds"):
        sim.gauss_residual_noalloc()
    sim.close()
