"""qvmhip_sample_noalloc (static scratch, no allocation) against
qvmhip_sample_ex on the same amplitudes and draws: identical indices and
near-boundary counts, on two calls in a row on one GPU, the second with a
smaller state and fewer draws (the first call's scratch contents must not
leak into the second).  Device buffers for the caller's side come from the
HIP runtime the engine links (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest

from qvm_amd import _native as N

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def noalloc_sample(h, amps, n, dtype, u, rule, eps):
    """One call of the entry point on caller-owned device buffers."""
    da = DevBuf(h, amps.nbytes)
    du = DevBuf(h, u.nbytes)
    do = DevBuf(h, 8 * u.size)
    dn = DevBuf(h, 8)
    da.put(amps)
    du.put(u)
    dn.put(np.array([12345], dtype=np.uint64))  # stale count: the call clears it
    code = N.F64 if dtype == "f64" else N.F32
    r = N.RULE_GE if rule == "ge" else N.RULE_GT
    N.check(N.lib().qvmhip_sample_noalloc(da.p, n, code, du.p, u.size, r, eps, do.p, dn.p, None))
    assert h.hipDeviceSynchronize() == 0
    out = do.get(np.uint64, u.size)
    near = int(dn.get(np.uint64, 1)[0])
    for b in (da, du, do, dn):
        b.free()
    return out, near


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_matches_sample_ex_two_calls(dtype):
    from qvm_amd.state import HipPureState
    assert N.device_count() >= 1
    h = hip()
    rng = np.random.default_rng(91)
    results = []
    # the second call is smaller in both the state and the draws
    for n, nshots, rule in ((22, 40000, "ge"), (11, 3000, "gt")):
        psi = rand_state(n, rng)
        st = HipPureState(n, dtype)
        st.write_amplitudes(psi)
        stored = st.amplitudes()  # the state's own values (FP32 rounded)
        cdf = np.cumsum(np.abs(stored) ** 2)
        u = np.concatenate([rng.random(nshots), cdf[rng.integers(0, cdf.size, 64)]])
        want, want_near = st.sample_ex(u, rule, 1e-15)
        amps = stored if dtype == "f64" else stored.astype(np.complex64)
        got, near = noalloc_sample(h, amps, n, dtype, u, rule, 1e-15)
        print(f"noalloc n={n} {dtype} rule={rule} draws={u.size} near={near}", flush=True)
        np.testing.assert_array_equal(got, want)
        assert near == int(want_near.sum())
        results.append((got, want))
        st.close()
    assert results[0][0].size != results[1][0].size


def test_noalloc_rejects_states_beyond_its_scratch():
    N.lib()
    rc = N.lib().qvmhip_sample_noalloc(C.c_void_p(16), 27, N.F64, C.c_void_p(16), 1, N.RULE_GE, 0.0,
                                       C.c_void_p(16), None, None)
    assert rc == N.EINVAL
    assert "26" in N.last_error()
