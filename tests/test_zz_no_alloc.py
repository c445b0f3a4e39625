"""md_macro_noalloc (no allocation; partials in the library's fixed device
array) against md_macro of a context holding the same state, through the C
ABI, on two calls in a row, the second with fewer particles (needs an
MI355X).  Device buffers come from the HIP runtime the library itself links."""
import ctypes as C

import numpy as np
import pytest

from tests.common import AR_MASS, engine_for, gas_state, liquid_state

pytestmark = pytest.mark.gpu


def hip_runtime():
    import moldyn_amd as M
    M.load_library()  # loads libamdhip64 as the library's dependency
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, host):
        self.hip = hip
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def macro_noalloc(hip, lib, v, pot, vir, mass, box):
    dv, dp, dw = DevBuf(hip, v), DevBuf(hip, pot), DevBuf(hip, vir)
    out = DevBuf(hip, np.full(11, np.nan))
    b = np.ascontiguousarray(box, dtype=np.float64)
    dpp = C.POINTER(C.c_double)
    rc = lib.md_macro_noalloc(C.cast(dv.ptr, dpp), C.cast(dp.ptr, dpp), C.cast(dw.ptr, dpp), len(v), mass,
                              b.ctypes.data_as(dpp), C.cast(out.ptr, dpp), None)
    assert rc == 0, lib.md_global_error()
    assert hip.hipDeviceSynchronize() == 0
    res = out.read()
    for buf in (dv, dp, dw, out):
        buf.free()
    return res


def test_macro_noalloc_matches_md_macro_twice():
    from moldyn_amd import _lib as L
    lib = L.load()
    hip = hip_runtime()
    # a drifting centre of mass, so that the thermal sums would show cancellation
    big = liquid_state(8, temperature=140.0, seed=4)  # 2,048 atoms
    big.vel = big.vel + np.array([3.0, -1.0, 2.0])
    small = gas_state(700, 6.0, 8)
    for s in (big, small):
        e = engine_for(s, "fp64_list")
        e.update_force()
        ref = e.macro()
        d = e.download(("v", "pot", "virial", "box"))
        e.close()
        got = macro_noalloc(hip, lib, d["v"], d["pot"], d["virial"], AR_MASS, d["box"])
        want = np.concatenate([[ref["kinetic"], ref["potential"], ref["thermal"], ref["temperature"],
                                ref["pressure"]], ref["momentum"], ref["com_velocity"]])
        print(s.n, got, want)
        # FP64 sums in another fixed order: reduction-order tolerance
        scale = np.abs(want[:5])
        assert np.all(np.abs(got[:5] - want[:5]) <= 1e-12 * scale), (got, want)
        pscale = np.abs(d["v"]).sum() * AR_MASS
        assert np.all(np.abs(got[5:8] - want[5:8]) <= 1e-13 * pscale)
        assert np.all(np.abs(got[8:] - want[8:]) <= 1e-13 * np.abs(d["v"]).sum(0).max())


def test_macro_noalloc_rejects_bad_sizes():
    from moldyn_amd import _lib as L
    lib = L.load()
    dpp = C.POINTER(C.c_double)
    box = np.ones(3)
    one = np.zeros(3)
    assert lib.md_macro_noalloc(one.ctypes.data_as(dpp), None, None, 0, AR_MASS, box.ctypes.data_as(dpp),
                                one.ctypes.data_as(dpp), None) == L.MD_ERR_ARG
    assert lib.md_macro_noalloc(one.ctypes.data_as(dpp), None, None, (1 << 40) + 1, AR_MASS,
                                box.ctypes.data_as(dpp), one.ctypes.data_as(dpp), None) == L.MD_ERR_ARG
    assert b"[1, 2^40]" in lib.md_global_error()
