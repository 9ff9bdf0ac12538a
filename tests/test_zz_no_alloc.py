"""akmc_itl_rates_noalloc (the dumbbell-rate step with caller-supplied device
buffers and the library's static working array) against akmc_itl_rates,
bit for bit, on two calls in a row (the second smaller), through the C ABI.
The caller's device buffers and stream come from the HIP runtime libakmc
itself links (/opt/rocm/lib/libamdhip64.so.7), called through ctypes: a
second runtime in the process (torch bundles its own) would not share
device pointers with it."""
import ctypes

import numpy as np
import pytest

import akmc
from akmc import workloads

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        akmc.lib()  # the runtime libakmc loaded
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        vp = ctypes.c_void_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
        assert _hip.hipSetDevice(0) == 0
    return _hip


class DevBuf:
    """A device buffer the caller owns (hipMalloc), filled from a host array."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.nbytes = max(a.nbytes, 1)
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), self.nbytes) == 0
        if a.nbytes:
            assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self, like):
        out = np.empty_like(like)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = None

    @property
    def addr(self):
        return self.p.value


def _noalloc(box, t, sites, orient, nu, T):
    n = len(sites)
    bufs = [DevBuf(np.asarray(t, np.uint16)), DevBuf(np.asarray(sites, np.int64)), DevBuf(np.asarray(orient, np.int8)),
            DevBuf(np.full(max(n, 1), 0xEE, np.uint8)), DevBuf(np.full((max(n, 1), 8), -1.0))]
    st = ctypes.c_void_p()
    assert hip().hipStreamCreate(ctypes.byref(st)) == 0
    try:
        akmc.itl_rates_noalloc(box, bufs[0].addr, n, bufs[1].addr, bufs[2].addr, nu, T, bufs[3].addr, bufs[4].addr,
                               device=0, stream=st.value)
        assert hip().hipStreamSynchronize(st) == 0
        av = bufs[3].get(np.zeros(max(n, 1), np.uint8))[:n]
        r = bufs[4].get(np.zeros((max(n, 1), 8)))[:n]
    finally:
        hip().hipStreamDestroy(st)
        for b in bufs:
            b.free()
    return av, r


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def test_noalloc_matches_itl_rates_on_two_calls():
    # A dense case first, then a smaller lattice with fewer dumbbells: the
    # second call must not see anything the first left in the working array.
    cases = [((16, 16, 16), 12, 0.02, 573.0), ((6, 4, 9), 4, 0.10, 563.0)]
    sizes = []
    for box, seed, f, T in cases:
        t, sites, orient = workloads.typed_lattice(box, seed, f, solutes=(0.05, 0.05, 0.05), f_v=0.02)
        print("no-alloc case: box", box, "dumbbells", len(sites), flush=True)
        av_ref, r_ref = akmc.itl_rates(box, t, sites, orient, 6e12, T)
        av, r = _noalloc(box, t, sites, orient, 6e12, T)
        assert np.array_equal(av, av_ref)
        assert _same(r, r_ref)
        assert (r_ref > 0).sum() > 0
        sizes.append(len(sites))
    assert sizes[1] < sizes[0]


def test_noalloc_rejects_bad_input_and_recovers():
    box = (8, 8, 8)
    t, sites, orient = workloads.typed_lattice(box, 21, 0.01, solutes=(0.05, 0.05, 0.05), f_v=0.02)
    assert len(sites) >= 4
    # Sites out of order: an error status, no output written.
    bad = sites.copy()
    bad[[0, 1]] = bad[[1, 0]]
    print("no-alloc case: unordered sites", flush=True)
    with pytest.raises(akmc.AkmcError) as ei:
        _noalloc(box, t, bad, orient, 6e12, 573.0)
    assert ei.value.code == akmc.AKMC_ERR_INVALID
    # A site index past the box.
    bad = sites.copy()
    bad[-1] = 2 * 8 * 8 * 8 + 5
    print("no-alloc case: site out of the box", flush=True)
    with pytest.raises(akmc.AkmcError):
        _noalloc(box, t, bad, orient, 6e12, 573.0)
    # More dumbbells than the working array holds: refused before any launch.
    d = DevBuf(np.zeros(8, np.int64))
    with pytest.raises(akmc.AkmcError) as ei:
        akmc.itl_rates_noalloc(box, d.addr, akmc.ITL_NOALLOC_MAX + 1, d.addr, d.addr, 6e12, 573.0, d.addr, d.addr)
    assert ei.value.code == -3
    d.free()
    # A valid call afterwards gives the reference results again.
    print("no-alloc case: valid call after the errors", flush=True)
    av_ref, r_ref = akmc.itl_rates(box, t, sites, orient, 6e12, 573.0)
    av, r = _noalloc(box, t, sites, orient, 6e12, 573.0)
    assert np.array_equal(av, av_ref) and _same(r, r_ref)
