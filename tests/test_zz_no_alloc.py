"""gfsph_neighbour_pairs_noalloc (the allocation-free neighbour set) through
the C ABI against gfsph_neighbour_pairs on the same inputs: the same pair
sets bit for bit, on two calls in a row (the second smaller, with
out-of-domain particles and the other kernel family), the count beyond a
short output capacity, and the size limit.  The test allocates the device
inputs and outputs with the HIP runtime the library uses; the entry point
itself allocates nothing."""
import ctypes

import numpy as np
import pytest

import gfsph
from gfsph import cases as C

pytestmark = pytest.mark.gpu


def _sorted(ij):
    ij = np.asarray(ij, dtype=np.int32).reshape(-1, 2)
    return ij[np.lexsort((ij[:, 1], ij[:, 0]))]


def _hip():
    """The HIP runtime libgfsph.so is linked against (found among the
    process's mapped libraries once the library is loaded)."""
    gfsph.load_library()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if "libamdhip64.so" in path:
            hip = ctypes.CDLL(path)
            hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            hip.hipFree.argtypes = [ctypes.c_void_p]
            hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            return hip
    raise RuntimeError("no HIP runtime mapped")


class _Dev:
    """Device buffers of the test (the entry point under test allocates none)."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.hip = _hip()
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(x, hsml, flag, cap, skf=1):
    lib = gfsph.load_library()
    d = _Dev()
    try:
        n = x.shape[0]
        dx = d.put(np.asarray(x, dtype=np.float64))
        dh = d.put(np.asarray(hsml, dtype=np.float64))
        df = None if flag is None else d.put(np.asarray(flag, dtype=np.int32))
        dn = d.put(np.full(1, 7, dtype=np.uint64))
        ij0 = np.full((max(cap, 1), 2), -1, dtype=np.int32)
        dij = d.put(ij0)
        rc = lib.gfsph_neighbour_pairs_noalloc(skf, n, dx, dh, df, cap, dn, dij, None)
        assert rc == 0, lib.gfsph_last_error(None)
        # the blocking copies below follow the null stream's work
        k = int(d.get(dn, np.zeros(1, np.uint64))[0])
        ij = d.get(dij, ij0)
    finally:
        d.close()
    return k, ij[: min(k, cap)]


def test_noalloc_pairs_equal_the_allocating_entry_point():
    # call 1: the 200,000-particle release of the fractal workload
    case = C.cfg_fractal_frictional(integrator=2, nsteps=1)
    P = case.particles
    want = gfsph.neighbour_pairs(P.x, P.hsml, None)
    k, ij = _noalloc(P.x, P.hsml, None, cap=want.shape[0] + 100)
    print(f"\nnoalloc call 1: n={P.n}, {k} pairs")
    assert k == want.shape[0]
    np.testing.assert_array_equal(_sorted(ij), want)
    # call 2, smaller: a random cloud with varied smoothing lengths, some
    # particles out of the domain, Gauss kernel (support 3 h)
    rng = np.random.Generator(np.random.Philox(17))
    n = 3000
    x = rng.uniform(0.0, 60.0, (n, 2))
    hs = rng.uniform(0.5, 1.5, n)
    flag = (rng.uniform(size=n) < 0.05).astype(np.int32)
    want2 = gfsph.neighbour_pairs(x, hs, flag, skf=2)
    k2, ij2 = _noalloc(x, hs, flag, cap=want2.shape[0], skf=2)
    print(f"noalloc call 2: n={n}, {k2} pairs")
    assert k2 == want2.shape[0]
    np.testing.assert_array_equal(_sorted(ij2), want2)


def test_noalloc_capacity_and_limit():
    rng = np.random.Generator(np.random.Philox(18))
    n = 2000
    x = rng.uniform(0.0, 40.0, (n, 2))
    hs = np.full(n, 1.2)
    want = gfsph.neighbour_pairs(x, hs)
    cap = want.shape[0] // 3
    k, ij = _noalloc(x, hs, None, cap=cap)
    assert k == want.shape[0] and ij.shape[0] == cap
    got = {tuple(p) for p in ij.tolist()}
    assert len(got) == cap and got <= {tuple(p) for p in want.tolist()}
    # beyond the maximum: refused on the host, before any device work
    lib = gfsph.load_library()
    rc = lib.gfsph_neighbour_pairs_noalloc(1, 262145, ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0,
                                           ctypes.c_void_p(16), None, None)
    assert rc == -5
    assert b"262144" in lib.gfsph_last_error(None)
