"""amrgpu_cfl_noalloc (the entry point that allocates nothing) against
amrgpu_cfl on the same meshes: per-instance CFL minima bit for bit, on two
calls in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

import amrhip
from amrhip import dpt

pytestmark = pytest.mark.gpu


def forest(seed, n, lmin, lmax):
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(2, 1, (0, 1), 0.0, 0.0, 0.5, n, lmin, lmax)
    for _ in range(lmax - lmin):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    s = f.soln()
    s[..., 0] = rng.uniform(0.2, 2.0, s.shape[:-1])
    s[..., 1] = rng.uniform(-1.0, 1.0, s.shape[:-1])
    s[..., 2] = rng.uniform(-1.0, 1.0, s.shape[:-1])
    s[..., 3] = rng.uniform(2.0, 5.0, s.shape[:-1])
    return f


class DeviceBuffers:
    """device copies of host arrays through the HIP runtime the library
    itself links (test plumbing: the entry point under test allocates
    nothing, so its inputs and outputs come from here)"""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # H2D
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0   # D2H
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def check(n, forests, h_base, cfl):
    L = amrhip.lib()
    insts = [amrhip.make_inst(h, (0, 0, 0, 0), c, 1.0) for h, c in zip(h_base, cfl)]
    ci, fi, geom, soln, cinst = amrhip.batch_external(forests)
    g = amrhip.Gpu()
    g.setup(n, 3, insts, ci, fi, geom, soln, cinst)
    ref = g.cfl()
    U, nq, nn, st = C.c_void_p(), C.c_int(), C.c_int(), C.c_void_p()
    assert L.amrgpu_device_state(g.h, C.byref(U), None, C.byref(nq), C.byref(nn), C.byref(st)) == 0
    dev = DeviceBuffers()
    level = dev.upload(np.concatenate([f.levels() for f in forests]).astype(np.int32))
    inst = dev.upload(np.asarray(cinst, np.int32))
    init = np.full(len(forests), -1.0)
    dt = dev.upload(init)
    hb = np.ascontiguousarray(h_base, np.float64)
    cf = np.ascontiguousarray(cfl, np.float64)
    print(f"noalloc call: n={n}, {nq.value} quads, {len(forests)} instances", flush=True)
    rc = L.amrgpu_cfl_noalloc(n, nq.value, len(forests), dpt(hb), dpt(cf), level, inst, U, dt, None)
    assert rc == 0, amrhip.last_error()
    got = dev.download(dt, init)
    dev.free()
    g.close()
    assert np.array_equal(got, ref), (got, ref)
    return got


def test_noalloc_cfl_equals_amrgpu_cfl_on_two_calls():
    big = [forest(1, 8, 1, 3), forest(2, 8, 2, 4), forest(3, 8, 1, 2)]
    first = check(8, big, [0.5, 0.5, 0.25], [0.3, 0.2, 0.3])
    small = [forest(4, 8, 1, 2), forest(5, 8, 1, 1)]
    second = check(8, small, [0.5, 1.0], [0.3, 0.3])
    assert len(second) == 2 and np.all(np.isfinite(first)) and np.all(second > 0)


def test_noalloc_cfl_16x16():
    check(16, [forest(6, 16, 1, 2), forest(7, 16, 1, 3)], [0.5, 0.5], [0.3, 0.3])


def test_noalloc_cfl_rejects_too_many_instances():
    """the size check comes before any device work"""
    L = amrhip.lib()
    h = np.ones(65)
    one = C.c_void_p(1)     # never dereferenced: the call is refused first
    assert L.amrgpu_cfl_noalloc(8, 4, 65, dpt(h), dpt(h), one, one, one, one, None) == -1
    assert "65 instances" in amrhip.last_error()
