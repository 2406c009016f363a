"""The C-ABI library: loads, exports every entry point include/krigekit_amd.h
declares, and keeps the reference's error conventions on host-side paths
(no GPU work is started by anything here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "krigekit_amd.h")
LIB = os.path.join(ROOT, "krigekit_amd", "libkrigekit_amd.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (python -m krigekit_amd.build)")


def declared_symbols():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(krige_\w+|kk_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 45
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_reference_import_symbols_present():
    """the symbols krigekit's wrapper binds at import (src/krigekit/kriging.py:142-286)"""
    lib = ctypes.CDLL(LIB)
    need = """krige_create krige_ind_create krige_ind_set_ncat krige_destroy krige_initialize
    krige_set_obs krige_set_obs_drift krige_update_obs_value krige_reset_vgm krige_set_vgm
    krige_set_vgm_block krige_set_grid krige_set_grid_block krige_set_grid_cv krige_set_grid_drift
    krige_set_sim krige_set_search krige_set_grad krige_solve krige_get_nblocks krige_get_nsim
    krige_get_block_coord krige_get_estimate krige_get_estimate_all krige_get_variance
    krige_get_variance_all krige_get_factor_info krige_get_factor_matrices krige_get_factor_system
    krige_get_last_error krige_solver_stats krige_to_str krige_get_max_threads krige_get_num_threads""".split()
    assert [s for s in need if not hasattr(lib, s)] == []


@pytest.fixture()
def K():
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=2)
    yield k
    k.close()


def test_getters_before_solve_raise(K):
    with pytest.raises(RuntimeError, match="solve\\(\\) has not been called"):
        K.get_results()


def test_duplicate_coordinates_rejected(K):
    coord = np.array([[0.0, 0.0], [1.0, 1.0], [0.0, 0.0]])
    with pytest.raises(RuntimeError, match="Duplicate coordinate found! Station 1 and Station 3"):
        K.set_obs(1, coord, np.zeros(3))


def test_unknown_variogram_type(K):
    K.set_obs(1, np.array([[0.0, 0.0], [1.0, 1.0]]), np.zeros(2))
    with pytest.raises(RuntimeError, match="unknown vtype"):
        K.set_vgm(1, 1, "xyz")


def test_error_message_format(K):
    with pytest.raises(RuntimeError) as e:
        K.set_search(1)
    msg = str(e.value)
    assert msg.startswith("KRIGING ERROR: Location: t_kriging%set_search")
    assert "Description: set_obs() needs to be called before set_search()." in msg


def test_bad_handle():
    from krigekit_amd import kriging as kk
    lib = kk._load_lib()
    n = ctypes.c_int(7)
    assert lib.krige_get_nblocks(ctypes.c_int64(12345), ctypes.byref(n)) == 1
    assert n.value == 0
    assert "Invalid kriging object handle" in kk._last_error()
    assert lib.krige_get_nblocks(ctypes.c_int64(0), ctypes.byref(n)) == 1
    assert "Null kriging object handle" in kk._last_error()


def test_out_of_scope_entry_points_report(K):
    from krigekit_amd import kriging as kk
    lib = kk._load_lib()
    lib.krige_set_nscore.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    assert lib.krige_set_nscore(ctypes.c_int64(K._handle), 1, 0.0, 1.0, 1, 1, 0.0, 1.0, 0, None) == 1
    assert "not provided by this backend" in kk._last_error()


def test_set_sim_without_simulations_is_a_no_op(K):
    """set_sim_base returns at once when nsim = 0 (kriging_base.F90:1235)"""
    from krigekit_amd import kriging as kk
    lib = kk._load_lib()
    assert lib.krige_set_sim(ctypes.c_int64(K._handle), 0, None, 0, 0, None) == 0


def test_error_slot_cleared_by_next_call(K):
    from krigekit_amd import kriging as kk
    with pytest.raises(RuntimeError):
        K.get_results()
    n = ctypes.c_int(0)
    assert kk._load_lib().krige_get_nblocks(ctypes.c_int64(K._handle), ctypes.byref(n)) == 0
    assert kk._last_error() == ""


def test_sector_search_needs_two_dims():
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=1)
    k.set_obs(1, np.array([[0.0], [1.0], [2.0]]), np.zeros(3), nmax=1)
    with pytest.raises(RuntimeError, match="sector search needs ndim >= 2"):
        k.set_search(1, sector_search=True)
    k.close()


def test_destroy_zeroes_handle():
    from krigekit_amd import kriging as kk
    lib = kk._load_lib()
    h = ctypes.c_int64(0)
    assert lib.krige_create(ctypes.byref(h)) == 0 and h.value > 0
    assert lib.krige_destroy(ctypes.byref(h)) == 0 and h.value == 0


# ---- weight store: host-side behaviour (no device work) ---------------------
def _store_obj(nobs=20, nb=6, nmax=4, **kw):
    from krigekit_amd.kriging import Kriging
    rng = np.random.default_rng(3)
    k = Kriging(ndim=2, **kw)
    k.set_obs(1, rng.uniform(0, 10, (nobs, 2)), rng.normal(size=nobs), nmax=nmax)
    k.set_grid(rng.uniform(0, 10, (nb, 2)))
    return k


def _store(nb=6, nm=4, nobs=20, seed=4):
    rng = np.random.default_rng(seed)
    nnear = rng.integers(1, nm + 1, (nb, 1)).astype(np.int32)
    inear = np.zeros((nb, 1, nm), dtype=np.int32)
    weight = np.zeros((nb, 1, nm))
    for b in range(nb):
        n = nnear[b, 0]
        inear[b, 0, :n] = np.sort(rng.choice(nobs, n, replace=False)) + 1
        weight[b, 0, :n] = rng.dirichlet(np.ones(n))
    return dict(nnear=nnear, inear=inear, weight=weight, variance=rng.uniform(0, 1, nb))


def test_get_weights_without_store_raises():
    """kriging_capi_common.F90:503-519"""
    k = _store_obj()
    with pytest.raises(RuntimeError, match="Weight store not allocated"):
        k.get_weights()
    k.close()


def test_set_weights_round_trip_and_free():
    """set_weights keeps the arrays as given (kriging_base.F90:2930-2977),
    get_weights returns them; free_weight_store releases the store"""
    k = _store_obj()
    w = _store()
    k.set_weights(w)
    assert k.use_old_weight
    g = k.get_weights()
    for key in ("nnear", "inear", "weight", "variance"):
        np.testing.assert_array_equal(g[key], w[key])
    k.free_weight_store()
    with pytest.raises(RuntimeError, match="Weight store not allocated"):
        k.get_weights()
    k.close()


def test_set_weights_without_variance_gives_zero():
    k = _store_obj()
    w = _store()
    w.pop("variance")
    k.set_weights(w)
    assert np.all(k.get_weights()["variance"] == 0.0)
    k.close()


def test_set_weights_shape_checks():
    """the reference's messages (kriging_base.F90:2950-2968)"""
    k = _store_obj()
    w = _store(nb=5)
    with pytest.raises(RuntimeError, match="nnear shape mismatch"):
        k.set_weights(w)
    w = _store(nm=3)
    with pytest.raises(RuntimeError, match="inear shape mismatch"):
        k.set_weights(w)
    w = _store()
    w["nnear"][2, 0] = 9
    with pytest.raises(RuntimeError, match="nnear out of range"):
        k.set_weights(w)
    k.close()


def test_set_weights_with_store_weight_raises():
    k = _store_obj(store_weight=True)
    with pytest.raises(ValueError, match="store_weight=True"):
        k.set_weights(_store())
    k.close()


def test_set_weights_needs_grid():
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=2)
    k.set_obs(1, np.random.default_rng(1).uniform(0, 1, (5, 2)), np.zeros(5), nmax=4)
    with pytest.raises(RuntimeError, match="call set_grid\\(\\) before alloc_weight_store"):
        k.set_weights(_store())
    k.close()


def test_replay_checks_before_device_work(tmp_path):
    """use_old_weight errors raised before any device work: no store, an
    empty store, a factor file of other dimensions or cut short
    (kriging_base.F90:3074-3106, 2896-2917)"""
    k = _store_obj(use_old_weight=True)
    with pytest.raises(RuntimeError, match="requires set_weights\\(\\) to be called first"):
        k.solve()
    w = _store()
    w["nnear"][:] = 0
    k.set_weights(w)
    with pytest.raises(RuntimeError, match="in-memory weight store is empty"):
        k.solve()
    k.close()
    fac = tmp_path / "w.fac"
    fac.write_text("           6           1           5\n")
    k = _store_obj(use_old_weight=True, weight_file=str(fac))
    with pytest.raises(RuntimeError, match="weight_file dimensions do not match"):
        k.solve()
    k.close()
    fac.write_text("           6           1           4\n1  2\n  3  4\n  0.5\n  0.25  0.75\n2  1\n")
    k = _store_obj(use_old_weight=True, weight_file=str(fac))
    with pytest.raises(RuntimeError, match="Unexpected end-of-file or read error in factor file"):
        k.solve()
    k.close()
