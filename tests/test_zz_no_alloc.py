"""The allocation-free forward pass (nhqg_mlp_forward_noalloc) against the
allocating one (nhqg_mlp_forward), through the C ABI.

Both run the same three layer kernels in the same order, so the results
are bit-identical.  Two calls in a row, the second with a smaller batch,
show that the shared scratch array carries nothing from one call into the
next; sizes beyond the stated maxima are rejected before any launch.  The
caller's device buffers come from the HIP runtime the library itself
links (libamdhip64.so.7, already loaded with libnhqg.so)."""

import ctypes as C

import numpy as np
import pytest

from nhqg_amd import _native

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """A caller-owned device allocation (hipMalloc / hipFree)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(a.nbytes), 1) == 0          # host -> device

    def download(self, shape):
        out = np.empty(shape, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(out.nbytes), 2) == 0        # device -> host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def hip():
    _native.load()                       # loads libnhqg.so and its HIP runtime
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _mlp(width=128, seed=0):
    m = _native.MLP(16, width, 16, max_batch=4096)
    rng = np.random.default_rng(seed)
    m.set_params((0.2 * rng.standard_normal(m.n_params)).astype(np.float32))
    return m


def test_noalloc_forward_matches_forward_on_two_calls(hip):
    m = _mlp()
    rng = np.random.default_rng(1)
    xs = [rng.standard_normal((b, 16)).astype(np.float32) for b in (3000, 700)]
    want = [m.forward(x) for x in xs]
    bufs = []
    for x in xs:
        xd, yd = DeviceBuffer(hip, x.nbytes), DeviceBuffer(hip, x.nbytes)
        xd.upload(x)
        yd.upload(np.full(x.shape, np.nan, np.float32))
        bufs.append((xd, yd))
    assert hip.hipDeviceSynchronize() == 0
    print("noalloc forward: batches 3000 then 700", flush=True)
    for (xd, yd), x in zip(bufs, xs):
        m.forward_noalloc(xd.ptr.value, x.shape[0], yd.ptr.value, 0)
    assert hip.hipDeviceSynchronize() == 0
    for (xd, yd), x, w in zip(bufs, xs, want):
        assert np.array_equal(yd.download(x.shape), w)
        xd.free()
        yd.free()


def test_noalloc_forward_rejects_sizes_beyond_the_maximum(hip):
    x, y = DeviceBuffer(hip, 4097 * 16 * 4), DeviceBuffer(hip, 4097 * 16 * 4)
    print("noalloc forward: size checks", flush=True)
    with pytest.raises(_native.NativeError, match="batch must be"):
        _mlp().forward_noalloc(x.ptr.value, 4097, y.ptr.value, 0)
    with pytest.raises(_native.NativeError, match="width must be"):
        _mlp(width=256).forward_noalloc(x.ptr.value, 8, y.ptr.value, 0)
    x.free()
    y.free()
