"""sc_mp_update_noalloc: the material-point update on caller-owned device
planes with the library's fixed scratch array (no device allocation).  Its
results must be those of sc_mp_update on the same inputs, bitwise (the same
kernels run), over two calls in a row on one GPU with the second smaller, so
that the second call runs on the scratch the first one left behind."""
import ctypes as C

import numpy as np
import pytest

from conftest import PRODUCT_SO
from solver_cpp_amd.capi import Library, MpDevicePlanes, ptr
from solver_cpp_amd.geometry import uniform_orientations

IP = C.POINTER(C.c_int)
MAX_POINTS = 32768  # SC_MP_NOALLOC_MAX_POINTS


@pytest.fixture(scope="module")
def lib():
    return Library(PRODUCT_SO)


def _inputs(lib, n, strain, seed):
    rng = np.random.default_rng(seed)
    q = uniform_orientations(n, seed + 1)
    Fp0 = np.zeros((n, 9))
    for c in range(n):
        lib.sc_qu2om(ptr(q[c]), ptr(Fp0[c]))
    F = np.tile(np.eye(3).reshape(-1), (n, 1)) + strain * rng.uniform(-1, 1, (n, 9))
    m, p, _ = lib.defaults()
    return dict(F=np.ascontiguousarray(F), F0=np.tile(np.eye(3).reshape(-1), (n, 1)), Fp0=Fp0,
                Lp0=np.zeros((n, 9)), S0=np.zeros((n, 9)), xi0=np.full((n, 12), m.xi_0_sl),
                gamma0=np.zeros((n, 12)), Lp=np.zeros((n, 9))), m, p


OUT9 = ("P", "Fp", "S", "Fe")
OUT12 = ("xi", "gamma")


def _host_update(lib, x, m, p, dt):
    """sc_mp_update (host arrays, the entry point the new one is a variant of)."""
    n = x["F"].shape[0]
    out = {k: np.zeros((n, 9)) for k in OUT9}
    out.update({k: np.zeros((n, 12)) for k in OUT12})
    Lp = x["Lp"].copy()
    d = np.zeros((n, 81))
    br = np.zeros(n, dtype=np.int32)
    it = np.zeros(2 * n, dtype=np.int32)
    lib.check(lib.sc_mp_update(n, dt, C.byref(m), C.byref(p), *(ptr(x[k]) for k in
                                                                 ("F", "F0", "Fp0", "Lp0", "S0", "xi0", "gamma0")),
                               ptr(Lp), *(ptr(out[k]) for k in OUT9), *(ptr(out[k]) for k in OUT12), ptr(d),
                               ptr(br, IP), ptr(it, IP)), "sc_mp_update")
    out.update(Lp=Lp, dPdF=d, broken=br, iterations=it.reshape(n, 2))
    return out


class _Hip:
    """Device memory through the HIP runtime the library itself links (one
    runtime per process: a second copy, such as torch's, cannot share it)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.h.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        self._ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _device_update(lib, x, m, p, dt):
    """sc_mp_update_noalloc on device planes (structure of arrays) the test owns."""
    hip = _Hip()
    n = x["F"].shape[0]
    host = {k: np.ascontiguousarray(v.T) for k, v in x.items()}
    for k in OUT9:
        host[k] = np.zeros((9, n))
    for k in OUT12:
        host[k] = np.zeros((12, n))
    host["dPdF"] = np.zeros((81, n))
    host["broken"] = np.full(n, -1, dtype=np.int32)
    host["iterations"] = np.full((2, n), -1, dtype=np.int32)
    try:
        dev = {k: hip.put(v) for k, v in host.items()}
        pl = MpDevicePlanes(**{k: dev[k].value for k in (f for f, _ in MpDevicePlanes._fields_)})
        lib.check(lib.sc_mp_update_noalloc(n, dt, C.byref(m), C.byref(p), C.byref(pl), None), "sc_mp_update_noalloc")
        hip.sync()
        res = {k: hip.get(dev[k], host[k]) for k in host}
    finally:
        hip.free()
    out = {k: res[k].T for k in OUT9 + OUT12 + ("Lp", "dPdF")}
    out.update(broken=res["broken"], iterations=res["iterations"].T)
    return out


@pytest.mark.gpu
def test_noalloc_matches_mp_update_twice(lib):
    dt = 0.5
    for n, strain, seed in ((4096, 2e-2, 31), (1000, 1e-3, 32)):
        print(f"[noalloc] n={n} strain={strain} seed={seed}", flush=True)
        x, m, p = _inputs(lib, n, strain, seed)
        a = _host_update(lib, x, m, p, dt)
        b = _device_update(lib, x, m, p, dt)
        np.testing.assert_array_equal(b["broken"], a["broken"])
        np.testing.assert_array_equal(b["iterations"], a["iterations"])
        for k in OUT9 + OUT12 + ("Lp", "dPdF"):
            np.testing.assert_array_equal(b[k], a[k], err_msg=f"{k} (n={n})")
        if strain > 1e-2:  # the first call did cut back (the scratch planes were used)
            assert (a["iterations"][:, 1] > 12).any()


def test_noalloc_refuses_more_points_than_its_scratch(lib):
    """(CPU) A size beyond the fixed scratch is refused before any device work."""
    m, p, _ = lib.defaults()
    pl = MpDevicePlanes()
    rc = lib.sc_mp_update_noalloc(MAX_POINTS + 1, 0.5, C.byref(m), C.byref(p), C.byref(pl), None)
    assert rc != 0
    assert b"exceed" in lib.sc_last_error()
