"""Graph-captured passes and the per-bucket schedule autotuner.

A captured pass consumes every count on the device (no host synchronisation
inside) and is replayed over the resident inputs: each replay, with eager
work and synchronisations between replays, gives bitwise the results of the
host-orchestrated pass it was captured after.  Autotuned schedules change
only which solve instance and grid run a bucket: results stay within the
FP64 bar of the oracle."""
import numpy as np
import pytest

import problems
from runners import make_gpu, max_errors, run_oracle

pytestmark = pytest.mark.gpu


def _cases():
    from krigekit_amd import synth
    return {
        "ok3d_octant": lambda: synth.ok3d(scale=1.0 / 16),       # octant search, factor reuse
        "lattice_ties": lambda: problems.lattice(nmax=8),         # flagged queries: exact search
        "near_duplicates": problems.near_duplicates,              # FP32+refine re-solves
        "ladder": problems.ladder,                                # Bunch-Kaufman fallback, NaN
        "gq_blocks": lambda: problems.gq_blocks(),                # block support
        "plain_knn": lambda: problems.random_problem(ndim=3, nobs=3000, nq=5000, nmax=24, seed=41),
    }


@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
@pytest.mark.parametrize("case", list(_cases()))
@pytest.mark.parametrize("reuse", [1, 0], ids=["grouped", "bucketed"])
def test_graph_replays_match_the_pass(case, precision, reuse):
    from krigekit_amd.kriging import device_synchronize
    spec = _cases()[case]()
    print("case graph", case, precision, reuse, flush=True)
    k = make_gpu(spec, precision=precision)
    k.set_factor_reuse(reuse, 4)
    k.solve()
    e0, v0 = k.get_results(copy=True)
    k.graph_capture()
    for rep in range(3):
        k.graph_launch()
        k.fetch()
        e, v = k.get_results(copy=True)
        assert np.array_equal(e, e0, equal_nan=True) and np.array_equal(v, v0, equal_nan=True), (case, rep)
        # eager work between the replays
        k.solve_resident()
        device_synchronize()
    k.close()


def test_graph_is_dropped_when_inputs_change():
    spec = problems.random_problem(ndim=2, nobs=800, nq=1000, nmax=12, seed=43)
    k = make_gpu(spec)
    k.solve()
    k.graph_capture()
    k.graph_launch()
    k.set_grid(spec["grid"][::-1].copy())
    with pytest.raises(RuntimeError):
        k.graph_launch()
    k.close()


def test_autotune_table_round_trip(tmp_path):
    """the tuner writes one schedule per non-empty bucket; results with the
    tuned table stay within 1e-10 of the oracle, and a captured pass uses it"""
    from krigekit_amd.kriging import set_tuning
    from krigekit_amd import synth
    spec = synth.ok3d(scale=1.0 / 64)
    path = str(tmp_path / "tuning.txt")
    try:
        k = make_gpu(spec)
        k.solve()
        k.autotune(path, reps=2)
        rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
        assert rows and all(len(r) == 7 and int(r[4]) >= (8, 12, 16, 24, 32, 48, 64)[int(r[3]) - 1]
                            for r in rows)
        k.solve()
        g = dict(estimate=k.get_results(copy=True)[0], variance=k.get_results(copy=True)[1])
        o = run_oracle(spec)
        ee, ev = max_errors(g, o)
        assert ee < 1e-10 and ev < 1e-10, (ee, ev)
        k.graph_capture()
        k.graph_launch()
        k.fetch()
        e, v = k.get_results(copy=True)
        assert np.array_equal(e, g["estimate"], equal_nan=True)
        k.close()
        set_tuning(path)
        with pytest.raises(RuntimeError):
            set_tuning(str(tmp_path / "missing.txt"))
    finally:
        # back to the table the library loaded at start
        import os
        import krigekit_amd.kriging as kr
        default = os.path.join(os.path.dirname(kr.LIB_PATH), "tuning.txt")
        set_tuning(default if os.path.exists(default) else None)
