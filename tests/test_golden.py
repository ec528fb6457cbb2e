"""The oracle against the committed golden fixtures (tests/golden/*.npz).

Every output is re-derived from the inputs stored in the fixture and must be
bit-identical: the oracle is deterministic (fixed chunking, fixed summation
order), so any change to it shows up here until the fixtures are rewritten
with tests/golden/make_golden.py."""
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as M
from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLD))
import make_golden as G  # noqa: E402


def load(name):
    return np.load(GOLD / f"{name}.npz", allow_pickle=False)


def _opt(a):
    return None if a.size == 0 else a


@pytest.mark.parametrize("name", list(G.PUSH_CASES))
def test_push_fixture(name):
    f = load(name)
    inp = G.push_inputs(name)
    st = tuple(f[k] for k in ("x", "vx", "vy", "vz", "w"))
    for a, b in zip(st, inp["state"]):
        np.testing.assert_array_equal(a, b)  # the inputs are reproducible too
    inp["state"] = st
    inp["E"], inp["R"], inp["G"] = _opt(f["E"]), _opt(f["R"]), _opt(f["G"])
    out, mom, cnt = G.push_expected(inp)
    np.testing.assert_array_equal(np.stack(out), f["out"])
    np.testing.assert_array_equal(mom, f["moments"])
    np.testing.assert_array_equal(cnt, f["counters"])
    assert f["counters"][4] > 0  # every fixture exercises ionization


@pytest.mark.parametrize("nb", [3, 5])
def test_resample_fixture(nb):
    f = load(f"resample_{nb}bin_f64")
    inp = G.resample_inputs(nb, np.float64)
    inp["state"] = tuple(f[k] for k in ("x", "vx", "vy", "vz", "w"))
    counts, offsets, perm, out, cnt = G.resample_expected(inp)
    np.testing.assert_array_equal(counts, f["counts"])
    np.testing.assert_array_equal(offsets, f["offsets"])
    np.testing.assert_array_equal(perm, f["perm"])
    np.testing.assert_array_equal(np.stack(out), f["out"])
    np.testing.assert_array_equal(cnt, f["counters"])
    assert cnt[8] > 0 and cnt[9] > 0  # one split and one merge at least
    ip, ci = G.resample_inplace_expected(inp)
    np.testing.assert_array_equal(np.stack(ip), f["out_inplace"])
    np.testing.assert_array_equal(ci, f["counters"])


@pytest.mark.parametrize("scheme", [0, 1])
def test_field_fixture(scheme):
    f = load(f"field_scheme{scheme}")
    inp = G.field_inputs()
    for k_in, k in (("mom", "mom"), ("ni_in", "ni"), ("E_in", "E"), ("chol", "chol")):
        np.testing.assert_array_equal(f[k_in], inp[k])
    got = G.field_expected(inp, scheme)
    for k, a in zip(("ni", "ns", "Te", "g2", "V", "E"), got):
        np.testing.assert_array_equal(a, f[k])


def test_mlp_fixture():
    f = load("mlp")
    inp = G.mlp_inputs()
    np.testing.assert_array_equal(inp["master"], f["master"])
    bits, y, loss, grad = G.mlp_expected(inp)
    np.testing.assert_array_equal(bits, f["bits"])
    np.testing.assert_array_equal(y, f["y"])
    assert float(loss) == float(f["loss"])
    np.testing.assert_array_equal(grad, f["grad"])
    _ = M
