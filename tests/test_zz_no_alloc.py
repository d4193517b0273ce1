"""mpm_sort_keys_noalloc -- the radix sort of mpm_probe_sort on device
memory, with its temporaries in a fixed device array instead of hipMalloc --
against mpm_probe_sort, bit for bit, on two calls in a row (the second
smaller, so a stale tail of the first call's temporaries would show), an
in-place call and the size limit.  Device buffers come from the HIP runtime
the library itself is linked against (loaded through ctypes): a second HIP
runtime in the process (PyTorch's own) would not see the device."""
import ctypes as C

import numpy as np
import pytest

from mpm_cuda_amd import capi

pytestmark = pytest.mark.gpu


def _hip():
    capi.lib()                                   # loads the library and with it its HIP runtime
    for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.skip("HIP runtime not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, n):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(n, 1) * 4) == 0
        self.n = n

    def put(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.n, np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _device_sort(hip, keys, bits, in_place=False):
    n = keys.size
    d = DevBuf(hip, n)
    srt = d if in_place else DevBuf(hip, n)
    perm = DevBuf(hip, n)
    try:
        d.put(keys)
        capi.sort_keys_noalloc(d.p.value or 0, n, bits, srt.p.value or 0, perm.p.value or 0, 0)
        assert hip.hipDeviceSynchronize() == 0
        return srt.get(), perm.get()
    finally:
        for b in {id(x): x for x in (d, srt, perm)}.values():
            b.free()


def test_noalloc_sort_matches_probe_sort(gpu):
    hip = _hip()
    rng = np.random.default_rng(44)
    # 24-bit keys: three 8-bit passes, the result comes back from the
    # scratch buffers; then 17-bit keys (two 9-bit passes) on fewer keys
    for n, bits in ((capi.SORT_NOALLOC_MAX_KEYS, 24), (70_001, 17), (1, 5)):
        keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[0]                # many equal keys: stability matters
        want_s, want_p = capi.probe_sort(keys, bits)
        got_s, got_p = _device_sort(hip, keys, bits)
        print(n, bits, "sorted keys equal:", np.array_equal(got_s, want_s), "permutation equal:",
              np.array_equal(got_p, want_p))
        assert np.array_equal(got_s, want_s) and np.array_equal(got_p, want_p), (n, bits)
    keys = rng.integers(0, 2 ** 20, size=333_333, dtype=np.uint64).astype(np.uint32)
    want_s, want_p = capi.probe_sort(keys, 20)
    got_s, got_p = _device_sort(hip, keys, 20, in_place=True)
    assert np.array_equal(got_s, want_s) and np.array_equal(got_p, want_p)


def test_noalloc_sort_refuses_what_it_cannot_hold(gpu):
    hip = _hip()
    d = DevBuf(hip, 16)
    try:
        p = d.p.value
        with pytest.raises(capi.MpmError, match="MPM_SORT_NOALLOC_MAX_KEYS"):
            capi.sort_keys_noalloc(p, capi.SORT_NOALLOC_MAX_KEYS + 1, 20, p, p)
        with pytest.raises(capi.MpmError):
            capi.sort_keys_noalloc(p, 16, 0, p, p)
        with pytest.raises(capi.MpmError):
            capi.sort_keys_noalloc(0, 16, 8, p, p)
        capi.sort_keys_noalloc(0, 0, 8, 0, 0)            # nothing to sort
    finally:
        d.free()
