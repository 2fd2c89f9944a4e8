"""
tdse_tdma_batched_noalloc (the allocation-free Thomas entry point) against
tdse_tdma_batched on the same systems, through the C ABI. Device buffers for
the caller's side are made with the HIP runtime the library itself links
(libamdhip64.so.7), so the test process holds one HIP runtime only.
"""

import ctypes

import numpy as np
import pytest

from ionization_amd import _lib
from oracle import mesh_cpu

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
MAX_POINTS = 1 << 21
MAX_SYSTEMS = 1 << 16


@pytest.fixture(scope="module")
def hip():
    _lib.load()  # loads /opt/rocm's libamdhip64.so.7 as the library's dependency
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    for f in (rt.hipMalloc, rt.hipMemcpy, rt.hipFree):
        f.restype = ctypes.c_int
    return rt


@pytest.fixture(scope="module")
def ctx():
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.tdse_ctx_create(0, ctypes.c_size_t(256 << 20), ctypes.byref(h)))
    yield h
    lib.tdse_ctx_destroy(h)


class DeviceArrays:
    """Caller-owned device copies of complex128 host arrays."""

    def __init__(self, hip, arrays):
        self.hip = hip
        self.ptrs = []
        for a in arrays:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
            self.ptrs.append(p)

    def download(self, i, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptrs[i], out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def ragged_systems(rng, n_systems, max_len):
    lengths = rng.integers(1, max_len, n_systems).astype(np.int64)
    lengths[:3] = [1, 2, 3]
    # a gap between systems: entries outside every system are left alone
    offsets = np.concatenate([[0], np.cumsum(lengths + 2)[:-1]]).astype(np.int64)
    total = int(offsets[-1] + lengths[-1] + 5)
    c = lambda: rng.standard_normal(total) + 1j * rng.standard_normal(total)
    sub, sup, d = c(), c(), c()
    diag = c() + 4.0
    return [np.ascontiguousarray(a) for a in (sub, diag, sup, d)], offsets, lengths, total


def call_noalloc(ctx, hip, arrs, offsets, lengths, total, x0):
    lib = _lib.load()
    dev = DeviceArrays(hip, arrs + [x0])
    try:
        rc = lib.tdse_tdma_batched_noalloc(ctx, *dev.ptrs, offsets.ctypes.data_as(_lib.c_int64_p),
                                           lengths.ctypes.data_as(_lib.c_int64_p), len(offsets), total)
        _lib.check(rc, ctx)
        return dev.download(4, x0)
    finally:
        dev.free()


def call_batched(ctx, arrs, offsets, lengths, x0):
    lib = _lib.load()
    x = x0.copy()
    _lib.check(lib.tdse_tdma_batched(ctx, *[_lib.dptr(a) for a in arrs], _lib.dptr(x),
                                     offsets.ctypes.data_as(_lib.c_int64_p),
                                     lengths.ctypes.data_as(_lib.c_int64_p), len(offsets)), ctx)
    return x


@pytest.mark.parametrize("seed", [0])
def test_noalloc_equals_batched_on_two_calls_second_smaller(ctx, hip, seed):
    rng = np.random.default_rng(seed)
    for n_systems, max_len in [(300, 3000), (37, 700)]:
        print(f"noalloc tdma: {n_systems} systems, lengths < {max_len}", flush=True)
        arrs, offsets, lengths, total = ragged_systems(rng, n_systems, max_len)
        x0 = np.full(total, 7.0 + 3.0j)
        got = call_noalloc(ctx, hip, arrs, offsets, lengths, total, x0)
        want = call_batched(ctx, arrs, offsets, lengths, x0)
        covered = np.zeros(total, dtype=bool)
        for o, n in zip(offsets, lengths):
            covered[o:o + n] = True
        # same kernel, same inputs: identical unknowns; entries outside every system untouched
        assert np.array_equal(got[covered], want[covered])
        assert np.array_equal(got[~covered], x0[~covered])
        # and both solve the systems (the oracle's Thomas recurrence, cy.pyx:31-48)
        for i in [0, 1, 2, 3, n_systems // 2, n_systems - 1]:
            o, n = offsets[i], lengths[i]
            if n < 2:
                ref = arrs[3][o:o + 1] / arrs[1][o:o + 1]
            else:
                ref = mesh_cpu.thomas(arrs[0][o:o + n], arrs[1][o:o + n], arrs[2][o:o + n], arrs[3][o:o + n])
            assert np.linalg.norm(got[o:o + n] - ref) <= 1e-12 * np.linalg.norm(ref)


def test_noalloc_rejects_sizes_beyond_its_scratch(ctx, hip):
    lib = _lib.load()
    arrs = [np.ones(4, dtype=np.complex128) for _ in range(5)]
    dev = DeviceArrays(hip, arrs)
    try:
        def call(offsets, lengths, total):
            off = np.asarray(offsets, dtype=np.int64)
            ln = np.asarray(lengths, dtype=np.int64)
            return lib.tdse_tdma_batched_noalloc(ctx, *dev.ptrs, off.ctypes.data_as(_lib.c_int64_p),
                                                 ln.ctypes.data_as(_lib.c_int64_p), len(off), total)

        assert call([0], [4], MAX_POINTS + 1) == -1
        assert b"scratch maximum" in lib.tdse_last_error(ctx)
        many = np.zeros(MAX_SYSTEMS + 1, dtype=np.int64)
        assert call(many, many + 1, 4) == -1
        assert b"scratch maximum" in lib.tdse_last_error(ctx)
        assert call([2], [3], 4) == -1  # system past the end of the arrays
        assert call([0], [0], 4) == -1  # empty system
        assert call([0], [4], 4) == 0
    finally:
        dev.free()
