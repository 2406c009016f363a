"""Pins the sequential-simulation oracle (oracle/sgsim.py) on the CPU:
against the compiled reference's realisations and variances for the same
paths and draws (tests/golden/sgsim_vectors.npz, table and analytic mode),
against the ordinary-kriging oracle where the simulation reduces to it, and
by changes that must move the result (the co-location rule, the simple
kriging mean term, the draw)."""
import os

import numpy as np
import pytest

import problems
from oracle import okrige, sgsim
from runners import rel_err

VEC = np.load(os.path.join(os.path.dirname(__file__), "golden", "sgsim_vectors.npz"))
CASES = problems.sgsim_cases()
IDS = [c["name"] for c in CASES]


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_oracle_matches_reference_vectors(spec, mode):
    e, v = sgsim.simulate(spec, table_mode=(mode == "table"))
    er, vr = VEC[f"{spec['name']}/{mode}/sim"], VEC[f"{spec['name']}/{mode}/variance"]
    assert e.shape == er.shape and not np.isnan(er).any()
    assert np.max(rel_err(e, er)) < 1e-10
    assert np.max(rel_err(v, vr)) < 1e-10


def test_first_block_is_ordinary_kriging_plus_draw():
    """the first block on the path sees only observations: its realisation
    is the ordinary kriging estimate plus sqrt(variance) times its draw"""
    spec = CASES[0]
    e, v = sgsim.simulate(spec, table_mode=False)
    b = spec["path"][0] - 1
    ok = dict(spec)
    ok["grid"] = spec["grid"][b:b + 1]
    o = okrige.solve(ok, table_mode=False)
    assert rel_err(v[b:b + 1], o["variance"])[0] < 1e-12
    expect = o["estimate"][0] + np.sqrt(o["variance"][0]) * spec["sample"][0]
    assert np.max(rel_err(e[b], expect)) < 1e-12


def test_zero_draws_give_equal_realisations():
    """with every draw zero each realisation is the same sequence of
    conditional means"""
    spec = dict(CASES[1])
    spec["sample"] = np.zeros_like(spec["sample"])
    e, _ = sgsim.simulate(spec)
    assert np.array_equal(e[:, 0], e[:, 1]) and np.array_equal(e[:, 0], e[:, 2])


@pytest.mark.parametrize("change", ["colocation", "sk_mean", "draw"])
def test_terms_matter(change):
    spec = dict(CASES[4] if change == "colocation" else CASES[3])
    key = spec["name"]
    src = open(sgsim.__file__).read()
    if change == "colocation":
        src = src.replace("if not dup:", "if True:")
    elif change == "sk_mean":
        src = src.replace("avg += (1.0 - wo.sum()) * skm", "pass")
    else:
        src = src.replace("L * smp[p, s]", "0.5 * L * smp[p, s]")
    ns = {}
    exec(compile(src, "oracle/sgsim.py (changed)", "exec"), ns)
    e, _ = ns["simulate"](spec, table_mode=True)
    assert np.max(rel_err(e, VEC[f"{key}/table/sim"])) > 1e-4
