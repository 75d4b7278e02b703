"""dmx_get_entries_noalloc (the entry gather with its device staging in a
fixed per-device array) against dmx_get_entries, through the C ABI: equal bit
for bit on two calls in a row, the second with fewer entries, in both of rho's
stored orientations; the size limit and argument errors come back as
DMX_EINVAL."""
import ctypes as C

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def gather(lib, h, fn, idx):
    import dmsim_amd
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    re = np.empty(idx.size)
    im = np.empty(idx.size)
    P = C.POINTER
    rc = getattr(lib, fn)(h, idx.size, idx.ctypes.data_as(P(C.c_uint64)),
                          re.ctypes.data_as(P(C.c_double)), im.ctypes.data_as(P(C.c_double)))
    assert rc == dmsim_amd.DMX_OK, (fn, rc, lib.dmx_last_error())
    return re, im


@pytest.mark.parametrize("method", [1, 2])  # plain (leaves rho transposed), fused
def test_noalloc_gather_matches_get_entries(method):
    import dmsim_amd
    lib = dmsim_amd.load()
    n = 10
    s = dmsim_amd.Sim(n)
    s.set_method(method)
    s.append_table(W.random_mixed(n, 300, 4040 + method))
    s.upload()
    s.run()
    rng = np.random.default_rng(method)
    num = 1 << (2 * n)
    for count in (dmsim_amd.NOALLOC_MAX_ENTRIES, 777):  # second call smaller
        idx = rng.integers(0, num, count, dtype=np.uint64)
        want = gather(lib, s._h, "dmx_get_entries", idx)
        got = gather(lib, s._h, "dmx_get_entries_noalloc", idx)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), count
    # against the downloaded matrix as well (get_dm's convention)
    re, im = s.get_dm()
    idx = rng.integers(0, num, 1000, dtype=np.uint64)
    got = s.get_entries_noalloc(idx)
    assert np.array_equal(got.real, re[idx]) and np.array_equal(got.imag, im[idx])
    s.close()


def test_noalloc_gather_limits():
    import dmsim_amd
    lib = dmsim_amd.load()
    s = dmsim_amd.Sim(4)
    s.append(W.OP["H"], [0])
    s.upload()
    s.run()
    P = C.POINTER
    too_many = np.zeros(dmsim_amd.NOALLOC_MAX_ENTRIES + 1, dtype=np.uint64)
    out = np.empty(too_many.size)
    rc = lib.dmx_get_entries_noalloc(s._h, too_many.size, too_many.ctypes.data_as(P(C.c_uint64)),
                                     out.ctypes.data_as(P(C.c_double)), None)
    assert rc == dmsim_amd.DMX_EINVAL and b"65536" in lib.dmx_last_error()
    bad = np.array([256], dtype=np.uint64)  # 4^4 = 256 entries: out of range
    rc = lib.dmx_get_entries_noalloc(s._h, 1, bad.ctypes.data_as(P(C.c_uint64)),
                                     out.ctypes.data_as(P(C.c_double)), None)
    assert rc == dmsim_amd.DMX_EINVAL
    assert lib.dmx_get_entries_noalloc(s._h, 0, None, None, None) == dmsim_amd.DMX_OK
    assert s.get_entries_noalloc([0, 3, 17]).shape == (3,)
    s.close()
