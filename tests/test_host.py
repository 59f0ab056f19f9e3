"""Host-side logic of the package (CPU only)."""
import math

import pytest

from verona_amd import Boundary, Cart, EosTypes, host_array
from verona_amd.bc import splitmix64, uniform_stream


def test_splitmix64_known_values():
    # published first outputs of SplitMix64 for seed 0 and seed 1234567
    g = splitmix64(0)
    assert next(g) == 0xE220A8397B1DCDAF
    g = splitmix64(1234567)
    assert next(g) == 6457827717110365317
    u = list(next(uniform_stream(42)) for _ in range(1))
    assert 0.0 <= u[0] < 1.0


def test_boundary_kinds():
    b = Boundary(outflow=("ylo", "yhi", "zlo", "zhi", "xhi"))
    assert b.kinds((False, False, False)) == ["none", "outflow", "outflow", "outflow", "outflow", "outflow"]
    assert b.kinds((True, False, False))[:2] == ["periodic", "periodic"]
    bn = Boundary(outflow=("xlo", "xhi"), nozzle=dict(radius=1, rho=1, u=1, umag=1, eps=0))
    assert bn.kinds((False,) * 3)[4] == "nozzle"
    with pytest.raises(ValueError):
        Boundary(outflow=("top",))
    assert Cart((True, False, True)).periodic == (True, False, True)


def test_host_array_layout():
    P = host_array(3, 4, 5, 6)
    assert P.shape == (5, 10, 11, 12) and P.flags["F_CONTIGUOUS"]
    # Verona byte order: variable fastest, then x, y, z
    P[2, 7, 1, 3] = 1.0
    flat = P.ravel(order="K")
    assert flat[2 + 5 * (7 + 10 * (1 + 11 * 3))] == 1.0
    assert host_array(2, 4, 5).shape == (4, 10, 11)


def test_eos():
    e = EosTypes.Polytrope(4.0 / 3.0)
    assert EosTypes.Pressure(3.0, e) == pytest.approx(1.0)
    c = EosTypes.SoundSpeed(1.0, 3.0, e)
    assert c == pytest.approx(math.sqrt((4 / 3) * (1 / 3) * 3 / (1 + 4)))
    assert c < 1


def test_parvector_mirror():
    from verona_amd import Verona3D, Verona2D
    P = Verona3D.ParVector3D(8, 6, 4)
    assert (P.size_X, P.size_Y, P.size_Z) == (14, 12, 10)
    Q = Verona2D.ParVector2D(8, 6)
    assert Q.arr.shape == (4, 14, 12)
