"""Pins the oracle's MAC / centred-Vec3 advection, outflow extrapolation and
3D wall boundary conditions (the checks live in tests/oracle_pins.py).  CPU
only."""
import pytest

from tests.oracle_pins import all_pins

PINS = all_pins()


@pytest.mark.parametrize("name,pin", PINS, ids=[p[0] for p in PINS])
def test_pin(oracle, name, pin):
    pin(oracle)
