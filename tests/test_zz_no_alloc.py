"""pic_deposit_noalloc (device arrays, the caller's stream, no allocation)
against pic_deposit, bit for bit, through the C ABI: two calls in a row, the
second with fewer particles and a smaller grid, for every shape; sizes beyond
its fixed scratch are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import picgolf
    return picgolf.load()


@pytest.fixture(scope="module")
def hip(lib):
    """The HIP runtime libpicgolf itself loaded (device memory for the test)."""
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line)
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def dev_copy(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0       # host to device
    return p


def noalloc(L, hip, x, y, w, N, shape):
    xd = dev_copy(hip, x)
    yd = dev_copy(hip, y) if y is not None else None
    rd = dev_copy(hip, np.full(N, np.nan))
    try:
        rc = L.pic_deposit_noalloc(xd, yd, x.size, float(w), N, shape, rd, None)
        assert rc == 0, L.pic_noalloc_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        rho = np.zeros(N)
        assert hip.hipMemcpy(rho.ctypes.data, rd, rho.nbytes, 2) == 0    # device to host
    finally:
        for p in (xd, yd, rd):
            if p is not None:
                hip.hipFree(p)
    return rho


def reference(x, y, w, N, shape):
    import picgolf
    from picgolf.core import Config
    scheme = "leapfrog" if shape in ("ngp",) else "midpoint"
    cfg = Config(N=N, P=x.size, W=100.0, dt=1 / (6 * N), shape="gauss7" if shape == "area" else shape, scheme=scheme)
    with picgolf.Context(cfg) as c:
        return c.deposit(x, y, w=w, shape=shape)


SHAPES = {"ngp": 0, "area": 1, "gauss6": 6, "gauss7": 7}


@pytest.mark.parametrize("shape", ["ngp", "area", "gauss6", "gauss7"])
def test_two_calls_match_pic_deposit(lib, hip, shape):
    rng = np.random.default_rng(SHAPES[shape] + 3)
    for N, P in ((4096, 300001), (64, 5000)):          # the second call is smaller
        x = rng.uniform(-0.05, 1.05, P) if shape != "ngp" else rng.random(P)
        y = rng.random(P) if shape not in ("ngp",) else None
        w = 400.0 / P * N
        print(f"noalloc deposit shape={shape} N={N} P={P}", flush=True)
        got = noalloc(lib, hip, x, y, w, N, SHAPES[shape])
        ref = reference(x, y, w, N, shape)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def test_rejects_sizes_beyond_its_scratch(lib, hip):
    x = dev_copy(hip, np.zeros(16))
    rho = dev_copy(hip, np.zeros(8192))
    try:
        rc = lib.pic_deposit_noalloc(x, None, 16, 1.0, 8192, 7, rho, None)
        assert rc == -1
        assert b"4096" in lib.pic_noalloc_last_error()
    finally:
        hip.hipFree(x)
        hip.hipFree(rho)
