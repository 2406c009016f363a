"""The allocation-free entry point kk_solve_resident_noalloc, called through
the C ABI, against kk_solve_resident's results on the same handle: two calls
in a row, the second on a smaller problem, bitwise equal; sizes above the
fixed scratch are rejected."""
import ctypes

import numpy as np
import pytest

import problems
from krigekit_amd import synth
from krigekit_amd.kriging import _load_lib
from runners import make_gpu

pytestmark = pytest.mark.gpu


def _hip():
    """the HIP runtime the library itself is linked against (the path it was
    loaded from, read from the process's memory map)"""
    _load_lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    lib_paths = sorted(p for p in paths if "/torch/" not in p)
    assert lib_paths, "libamdhip64 not loaded"
    hip = ctypes.CDLL(lib_paths[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), 8 * max(n, 1)) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(k, est, var):
    lib = _load_lib()
    f = lib.kk_solve_resident_noalloc
    f.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f(k._handle, est.p, var.p)


def _last_error():
    buf = ctypes.create_string_buffer(4096)
    _load_lib().krige_get_last_error(buf, 4096)
    return buf.value.decode()


@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
def test_noalloc_matches_resident_solve(precision):
    hip = _hip()
    # a 3-D octant problem (131,072 nodes), then a smaller one that reaches the
    # symmetric-indefinite fallback, its jittered retry and failures (ladder)
    # and, at FP32+refine, the FP64 re-solve of near-duplicate systems
    for spec in (synth.ok3d(scale=1.0 / 32), problems.ladder(), problems.near_duplicates()):
        print("case noalloc", spec["name"], precision, len(spec["grid"]), flush=True)
        k = make_gpu(spec, precision=precision)
        k.solve()
        e_ref, v_ref = k.get_results(copy=True)
        nb = len(e_ref)
        est, var = DevBuf(hip, nb), DevBuf(hip, nb)
        assert _noalloc(k, est, var) == 0, _last_error()
        e, v = est.get(), var.get()
        assert np.array_equal(e, e_ref, equal_nan=True) and np.array_equal(v, v_ref, equal_nan=True)
        # the handle's own results are untouched
        e2, v2 = k.get_results(copy=True)
        assert np.array_equal(e2, e_ref, equal_nan=True)
        est.free()
        var.free()
        k.close()


def test_noalloc_rejects_what_it_cannot_hold():
    hip = _hip()
    est, var = DevBuf(hip, 1), DevBuf(hip, 1)
    # no finished solve on the handle yet
    k = make_gpu(problems.ok2d_small())
    assert _noalloc(k, est, var) != 0
    assert "finished single-variable solve" in _last_error()
    k.close()
    # 2,097,152 nodes of the 3-D octant workload need more than the fixed scratch
    spec = synth.ok3d(scale=0.5)
    k = make_gpu(spec)
    k.solve()
    assert _noalloc(k, est, var) != 0
    assert "MiB of scratch" in _last_error()
    k.close()
    est.free()
    var.free()
