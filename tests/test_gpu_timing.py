"""Per-kernel timing (vrn_timing_enable / vrn_timing_report): mode 1 brackets
every launch of the fused step, mode 2 only the directional sweeps, and a
mode outside 0..2 is refused."""
import pytest

from helpers import progress, state3

pytestmark = pytest.mark.gpu
G = 4.0 / 3.0


def test_timing_modes():
    progress("test_timing_modes")
    from verona_amd import DeviceGrid
    from verona_amd._lib import VeronaHIPError
    nx, ny, nz = 40, 24, 20
    bc = ["outflow", "outflow", "periodic", "periodic", "outflow", "outflow"]
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        dev.upload("P", state3(nx, ny, nz, seed=5))
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision="fp64fast",
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=1 / nx, dy=1 / ny, dz=1 / nz, bc=bc)
        reps = {}
        for mode in (1, 2):
            dev.timing(mode)
            for _ in range(2):
                dev.step()
            reps[mode] = dev.timing_report()
            dev.timing(False)
        with pytest.raises(VeronaHIPError, match="timing mode"):
            dev.timing(3)
    progress(f"  mode 1 {sorted(reps[1])}, mode 2 {sorted(reps[2])}")
    assert {"xysweep", "zsweep_c2p", "ghost", "rescue"} <= set(reps[1])
    assert set(reps[2]) == {"xysweep", "zsweep_c2p"}
    for k in ("xysweep", "zsweep_c2p"):
        assert reps[2][k][0] == 6 and reps[2][k][1] > 0.0
