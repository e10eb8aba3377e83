"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
the CPU oracle) are reproduced bit for bit by the oracle: a change to the oracle's arithmetic,
order or RNG shows here. The GPU path is held to the same fixtures in test_gpu_golden.py."""

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def make():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("name", ["rng", "deposit", "solve", "contact", "confinement", "step"])
def test_oracle_reproduces_the_golden_fixture(oracle, make, name):
    want = load(name)
    got = make.CASES[name]()
    assert sorted(got) == sorted(want)
    for key in want:
        np.testing.assert_array_equal(np.asarray(got[key]), want[key], err_msg=f"{name}.{key}")


def test_fixtures_hold_the_interesting_cases():
    step = load("step")
    assert step["daughters"].sum() > 0 and step["count"] > 250  # divisions happened
    assert load("solve")["converged"] == 1
    conf = load("confinement")
    assert (conf["pushed"] != conf["velocity"]).any(axis=1).sum() > 20  # some cells pushed
    assert (conf["pushed"] == conf["velocity"]).all(axis=1).sum() > 20  # some untouched
