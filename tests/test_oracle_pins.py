"""The oracle against the closed forms of tests/oracle_pins.py: the sheared
linear RHS, the Phillips dispersion relation, the ARK43 / ARK43_QGLeith and
BS3(2) error norms and step control, and AB3's local error constant."""
import pytest

from oracle import qgn_oracle as orc
from oracle_pins import PINS


@pytest.mark.parametrize("name", sorted(PINS))
def test_oracle_pin(name):
    PINS[name](orc)
