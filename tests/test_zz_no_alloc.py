"""help_gpu_pcg_apply_noalloc (the allocation-free entry point) against
help_gpu_pcg_apply on the same handle: the same kernels on the same data, so
the results must be bit-identical.  Two calls in a row on different handles,
the second with a smaller input, then the first handle again (the shared
scratch keeps the last call's contents, which the kernels overwrite)."""
import ctypes as C
import os

import numpy as np
import pytest

from help_amd import LinearEquation, instances
from help_amd import _lib as L

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the backend library is linked against (already mapped
    into this process once the library is loaded)."""
    L.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    path = sorted(paths)[0] if paths else "libamdhip64.so"
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class DevVec:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, 1) == 0   # host to device

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0   # device to host
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _handle(inst, K, seed, fp32=1):
    rng = np.random.default_rng(seed)
    d = np.exp(rng.uniform(-2, 2, inst.n) * np.log(10.0))
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=K, pcg_fp32_precond=fp32)
    be.construct(inst)
    be.renew(d)
    return be


def _check(hip, be, seed):
    r = np.random.default_rng(seed).standard_normal(be.m)
    ref = be.pcg_apply(r)
    dr, dz = DevVec(hip, be.m), DevVec(hip, be.m)
    try:
        dr.put(r)
        dz.put(np.full(be.m, np.nan))
        st = L.lib().help_gpu_pcg_apply_noalloc(be.h, dr.p, dz.p)
        assert st == L.HELP_OK, L.lib().help_gpu_last_error(be.h)
        z = dz.get()
    finally:
        dr.free()
        dz.free()
    print(f"noalloc m={be.m}: max |diff| {np.abs(z - ref).max():.3e}", flush=True)
    assert np.array_equal(z, ref)


def test_noalloc_apply_matches_pcg_apply():
    hip = _hip()
    big = _handle(instances.zipf(3000, 9000, 6, 1.1, seed=51), 40, 1)
    small = _handle(instances.random_lp(300, 600, 4, seed=52), 20, 2)
    fp64 = _handle(instances.random_lp(400, 800, 4, seed=53), 30, 3, fp32=0)
    try:
        _check(hip, big, 11)
        _check(hip, small, 12)     # second call, smaller input
        _check(hip, big, 13)       # the first handle again after the smaller one
        _check(hip, fp64, 14)      # FP64 preconditioner factor
        # host pointers are rejected, not dereferenced
        r = np.zeros(small.m)
        z = np.zeros(small.m)
        st = L.lib().help_gpu_pcg_apply_noalloc(small.h, r.ctypes.data, z.ctypes.data)
        assert st == L.HELP_ERR_ARG
    finally:
        for be in (big, small, fp64):
            be.destruct()
