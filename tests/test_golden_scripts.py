"""The committed 50-digit fixtures are what their scripts in tests/golden/
compute: each element's entries (all three kinds) and the bar and beam
gradients are recomputed here and compared digit for digit.  The plate
gradients and the tilted plates, which take the plate script about 25 s, are
left to a run of the script itself."""

import importlib.util
import json
import os

import mpmath as mp
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _script(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(GOLDEN, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _fixture(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _entries(K, n):
    return [[i, j, mp.nstr(K[i, j], 30)] for i in range(n) for j in range(n) if K[i, j] != 0]


def test_plate_entries_reproduce():
    mod = _script("make_plate_golden")
    fx = _fixture("plate_mitc4_golden.json")["cases"]
    cases = mod.CASES + mod._random_cases(20251203, 3)
    assert len(cases) == len(fx)
    for (xl, E, nu, t, kf), c in zip(cases, fx):
        assert c["props"] == [E, nu, t, kf]
        assert _entries(mod.plate_k(xl, E, nu, t, kf), 24) == c["entries"]


def test_beam_fixture_reproduces():
    mod = _script("make_beam_golden")
    fx = _fixture("beam_golden.json")
    assert fx["params"] == list(mod.PARAMS)
    for (L, props), c in zip(mod.CASES, fx["cases"]):
        assert _entries(mod.beam_k(L, props), 12) == c["entries"]
        w = {(i, j): mp.mpf(g) for (i, j, _), g in zip(c["entries"], c["G"])}
        assert [mp.nstr(v, 30) for v in mod.gradients(L, props, w)] == c["grad_G"]


@pytest.mark.parametrize("idx", range(3))
def test_truss_fixture_reproduces(idx):
    mod = _script("make_truss_golden")
    x1, x2, props = mod.CASES[idx]
    c = _fixture("truss_golden.json")["cases"][idx]
    assert _entries(mod.truss_k(x1, x2, *props), 12) == c["entries"]
    u = c["u"]
    w = {(i, j): -mp.mpf(u[i]) * mp.mpf(u[j]) for i, j, _ in c["entries"]}
    gp, gx = mod.gradients(x1, x2, props, w)
    assert [None if v is None else mp.nstr(v, 30) for v in gp] == c["grad_u"]["params"]
    assert [[mp.nstr(v, 30) for v in r] for r in gx] == c["grad_u"]["xyz"]
