"""Each pin of tests/oracle_pins.py is only as good as the mistakes it
catches.  This test plants one plausible mistake at a time in a copy of the
oracle's source (in memory, on the CPU), runs every pin on that copy, and
requires at least one pin to fail for every mistake, while all pass on the
unchanged oracle."""
import os
import sys
import types

import pytest

import oracle_pins

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "qgn_oracle.py")

# (name, [(text in the oracle, replacement), ...])
MUTATIONS = [
    ("beta - Qy", [("(p.beta + self.qyBar)", "(p.beta - self.qyBar)")]),
    ("uBar in reversed layer order", [("self.uBar[:, None, None] * (1j * self.kx)",
                                       "self.uBar[::-1, None, None] * (1j * self.kx)")]),
    ("ARK error norm over (nx ny)^2", [("err1 = dt * np.max(np.abs(grid)) / float(p.nx * p.ny)",
                                        "err1 = dt * np.max(np.abs(grid)) / float(p.nx * p.ny) ** 2")]),
    ("QG-Leith with ARK43's reject and safety factors", [
        ("reject_factor = 0.5 if self.leith else 0.75", "reject_factor = 0.75"),
        ("safety = 0.2 if self.leith else 0.75", "safety = 0.75")]),
    ("AB3 third step as AB2", [("q = q_hat + (dt / 12.0) * (23.0 * r2 - 16.0 * self.r1 + 5.0 * self.r0)",
                                "q = q_hat + dt * (1.5 * r2 - 0.5 * self.r1)")]),
    ("ARK43 with QG-Leith's safety factor", [("safety = 0.2 if self.leith else 0.75", "safety = 0.2")]),
    ("ARK error without its first term", [("spec = (BE[0] * (N[1] + L[1]) + BE[2]", "spec = (BE[2]")]),
    ("implicit factor 1/2 instead of 1/4", [("Mq = 1.0 / (1.0 + 0.25 * dt * lam)", "Mq = 1.0 / (1.0 + 0.5 * dt * lam)")]),
    ("BS3 error norm unnormalised", [("err_phys = m.c2r(err_spec) / float(p.nx * p.ny)",
                                      "err_phys = m.c2r(err_spec)")]),
    ("BS3 reject factor 0.5", [("dt = 0.75 * dt\n                self.rejects += 1\n                tries += 1\n"
                                "                if tries > self.max_rejects:\n"
                                "                    raise RuntimeError(\"BSRK3",
                                "dt = 0.5 * dt\n                self.rejects += 1\n                tries += 1\n"
                                "                if tries > self.max_rejects:\n"
                                "                    raise RuntimeError(\"BSRK3")]),
    ("BS3 safety 0.2", [("0.75 * p.TOL, 0.1, 0.4 / 3.0", "0.2 * p.TOL, 0.1, 0.4 / 3.0")]),
    ("AB3 coefficient 16 -> 15", [("- 16.0 * self.r1", "- 15.0 * self.r1")]),
    ("beta and Qy acting on psi instead of v", [("* self.v_hat          # :270-273", "* self.psi_hat          # :270-273")]),
]


def _mutant(edits):
    src = open(ORACLE).read()
    for old, new in edits:
        assert src.count(old) == 1, f"mutation anchor not found exactly once: {old!r}"
        src = src.replace(old, new)
    mod = types.ModuleType("qgn_oracle_mutant")
    mod.__file__ = ORACLE
    sys.modules[mod.__name__] = mod      # dataclasses looks the module up while the class is built
    try:
        exec(compile(src, ORACLE, "exec"), mod.__dict__)
    finally:
        del sys.modules[mod.__name__]
    return mod


@pytest.mark.parametrize("name,edits", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pins_catch_mutation(name, edits):
    mod = _mutant(edits)
    caught = []
    for pin_name, pin in oracle_pins.PINS.items():
        try:
            pin(mod)
        except (AssertionError, RuntimeError, FloatingPointError):
            caught.append(pin_name)
    assert caught, f"no pin fails when the oracle has the mistake: {name}"


def test_unmutated_copy_passes_every_pin():
    mod = _mutant([])
    for pin in oracle_pins.PINS.values():
        pin(mod)
