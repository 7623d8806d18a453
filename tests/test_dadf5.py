"""DADF5 result files: the HDF5 writer beneath sc_rve_result_* (CPU) and the
result file of a solved RVE (GPU).

Pins: the stored Fletcher-32 checksums and the object layout of the
reference's own result file 2x2x2_tensionX_material_numerics.hdf5 (kept as
data in reference_goldens.json); an independent reader (the published HDF5
library when it is installed, and tests/golden/hdf5_read.py, which verifies
every checksum); the orientations O that file stores next to its F_e.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

sys.path.insert(0, GOLDEN)
from hdf5_read import H5File  # noqa: E402

HDF5_LIB = "/opt/conda/lib/libhdf5.so.103"
MECH = "/increment_1/phase/Aluminum/mechanical/"


def _write(lib, path, compression, items):
    """items: (kind, name, payload[, extra]) written through sc_dadf5_*."""
    h = C.c_void_p()
    lib.check(lib.sc_dadf5_create(os.fsencode(path), int(compression), C.byref(h)), "sc_dadf5_create")
    try:
        for it in items:
            kind, name = it[0], it[1].encode()
            if kind == "f64":
                a = np.ascontiguousarray(it[2], dtype=np.float64)
                dims = (C.c_int64 * a.ndim)(*a.shape)
                rc = lib.sc_dadf5_dataset_f64(h, name, a.ndim, dims, a.ctypes.data_as(C.POINTER(C.c_double)))
            elif kind == "str":
                rc = lib.sc_dadf5_dataset_str(h, name, it[2].encode())
            elif kind == "attr_str":
                rc = lib.sc_dadf5_attr_str(h, name, it[2].encode(), it[3].encode())
            else:
                rc = lib.sc_dadf5_attr_f64(h, name, it[2].encode(), float(it[3]))
            lib.check(rc, f"sc_dadf5 {kind} {it[1]}")
    finally:
        lib.check(lib.sc_dadf5_close(h), "sc_dadf5_close")


def _stored_checksum(f: H5File, path: str) -> int:
    import struct
    lay = [d for t, d in f.messages(f.resolve(path)) if t == 0x08][0]
    assert lay[1] == 2, path  # chunked
    chunks = list(f._chunks(struct.unpack_from("<Q", lay, 3)[0], lay[2] - 1))
    assert len(chunks) == 1
    return struct.unpack_from("<I", chunks[0][1], len(chunks[0][1]) - 4)[0]


class Hdf5Lib:
    """The installed HDF5 library through ctypes (read side only)."""

    def __init__(self):
        self.h = C.CDLL(HDF5_LIB)
        self.h.H5open()
        for fn in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Dget_type", "H5Aopen", "H5Aget_type"):
            getattr(self.h, fn).restype = C.c_int64
        self.h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        self.h.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        self.h.H5Aopen.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        self.h.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        self.h.H5Aread.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
        for fn in ("H5Dget_space", "H5Dget_type", "H5Aget_type", "H5Sget_simple_extent_ndims", "H5Tget_size",
                   "H5Dclose", "H5Aclose", "H5Sclose", "H5Tclose", "H5Fclose"):
            getattr(self.h, fn).argtypes = [C.c_int64]
        self.h.H5Tget_size.restype = C.c_size_t
        self.h.H5Sget_simple_extent_dims.argtypes = [C.c_int64, C.c_void_p, C.c_void_p]
        self.double = C.c_int64.in_dll(self.h, "H5T_NATIVE_DOUBLE_g").value

    def open(self, path):
        f = self.h.H5Fopen(os.fsencode(path), 0, 0)
        assert f >= 0, "H5Fopen failed"
        return f

    def read_f64(self, f, name):
        d = self.h.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        s = self.h.H5Dget_space(d)
        n = self.h.H5Sget_simple_extent_ndims(s)
        dims = (C.c_uint64 * max(n, 1))()
        self.h.H5Sget_simple_extent_dims(s, dims, None)
        out = np.empty(tuple(dims[:n]), dtype=np.float64)
        rc = self.h.H5Dread(d, self.double, 0, 0, 0, out.ctypes.data)
        self.h.H5Sclose(s)
        self.h.H5Dclose(d)
        assert rc >= 0, f"H5Dread {name} failed (checksum or format)"
        return out

    def read_raw(self, f, name):
        """Bytes of a dataset in its file type (strings, records)."""
        d = self.h.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        t = self.h.H5Dget_type(d)
        s = self.h.H5Dget_space(d)
        n = self.h.H5Sget_simple_extent_ndims(s)
        dims = (C.c_uint64 * max(n, 1))()
        self.h.H5Sget_simple_extent_dims(s, dims, None)
        size = self.h.H5Tget_size(t) * int(np.prod(dims[:n]))
        buf = C.create_string_buffer(size)
        rc = self.h.H5Dread(d, t, 0, 0, 0, buf)
        for fn, x in (("H5Tclose", t), ("H5Sclose", s), ("H5Dclose", d)):
            getattr(self.h, fn)(x)
        assert rc >= 0, name
        return buf.raw

    def read_attr(self, f, obj_path, name):
        d = self.h.H5Dopen2(f, obj_path.encode(), 0)
        assert d >= 0, obj_path
        a = self.h.H5Aopen(d, name.encode(), 0)
        assert a >= 0, name
        t = self.h.H5Aget_type(a)
        buf = C.create_string_buffer(self.h.H5Tget_size(t) * 64)
        rc = self.h.H5Aread(a, t, buf)
        for fn, x in (("H5Tclose", t), ("H5Aclose", a), ("H5Dclose", d)):
            getattr(self.h, fn)(x)
        assert rc >= 0
        return buf.raw


@pytest.fixture(scope="module")
def hdf5():
    if not os.path.exists(HDF5_LIB):
        pytest.skip("no HDF5 library installed")
    return Hdf5Lib()


def test_writer_stores_reference_chunks_bitwise(product, goldens, tmp_path):
    """The reference file's values written back: same shapes, same bytes per
    chunk, hence the same Fletcher-32 checksums the reference stored."""
    g, st = goldens["dadf5_2x2x2_increment_1"], goldens["dadf5_2x2x2_storage"]
    data = {"F": g["F"], "F_e": g["Fe"], "F_p": g["Fp"], "L_p": g["Lp"], "P": g["P"], "O": g["O"],
            "xi_sl": g["xi"]}
    items = [("f64", MECH + k, np.array(v)) for k, v in data.items()]
    items += [("f64", "/increment_1/geometry/u_n", np.array(g["u_n"])),
              ("str", "/setup/material.yaml", st["material_yaml"]),
              ("str", "/setup/numerics.yaml", st["numerics_yaml"]),
              ("attr_f64", "/increment_1", "t/s", g["t"])]
    path = str(tmp_path / "w.hdf5")
    _write(product, path, True, items)
    f = H5File(path)
    for k, v in data.items():
        np.testing.assert_array_equal(f.read(MECH + k), np.array(v))
    assert f.attrs("/increment_1")["t/s"] == g["t"]
    assert f.read("/setup/numerics.yaml")[0].decode() == st["numerics_yaml"]
    for p, want in st["fletcher32"].items():
        if p.endswith("u_p"):
            continue  # the reference's u_p holds uninitialised memory (not kept)
        assert _stored_checksum(f, p) == want, p


def test_writer_read_by_hdf5_library(product, hdf5, tmp_path):
    """Large compressed datasets (two B-tree levels of chunks, a partial last
    chunk), small ones, strings and attributes read back by libhdf5 with its
    checksum verification on."""
    rng = np.random.default_rng(7)
    big = np.cumsum(rng.standard_normal((200_001, 3, 3)), axis=0)  # 13.7 chunks of 14563 rows
    huge = rng.standard_normal((1_000_003, 9))                      # 69 chunks -> 2-level chunk tree
    small = rng.standard_normal((5, 12))
    text = "material:\n  - phase: Aluminum\n" * 10000               # > 256 KiB: shuffled + deflated
    for compression in (True, False):
        path = str(tmp_path / f"big{int(compression)}.hdf5")
        _write(product, path, compression,
               [("f64", "a/b/big", big), ("f64", "a/huge", huge), ("f64", "small", small),
                ("str", "setup/text", text), ("attr_str", "a/b/big", "unit", "Pa"),
                ("attr_f64", "a", "t/s", 2.5)])
        f = hdf5.open(path)
        np.testing.assert_array_equal(hdf5.read_f64(f, "/a/b/big"), big)
        np.testing.assert_array_equal(hdf5.read_f64(f, "/a/huge"), huge)
        np.testing.assert_array_equal(hdf5.read_f64(f, "/small"), small)
        assert hdf5.read_raw(f, "/setup/text").rstrip(b"\0").decode() == text
        assert hdf5.read_attr(f, "/a/b/big", "unit").rstrip(b"\0") == b"Pa"
        hdf5.h.H5Fclose(f)
        r = H5File(path)
        np.testing.assert_array_equal(r.read("/a/huge"), huge)
        np.testing.assert_array_equal(r.read("/a/b/big"), big)
    # compression shrinks the smooth dataset
    assert os.path.getsize(tmp_path / "big1.hdf5") < os.path.getsize(tmp_path / "big0.hdf5")


def test_orientation_of_reference_elastic_deformation(product, goldens):
    """O = om2qu(transpose(rotational part of F_e)) (phase.f90:590) against
    the O the reference stored next to its F_e.  The rotations are tiny
    (1e-22, 1e-31), so each component is compared relative to itself: a
    sign or transposition mistake flips them."""
    g = goldens["dadf5_2x2x2_increment_1"]
    for Fe_rows, O_ref in zip(g["Fe"], g["O"]):
        Fe = np.asfortranarray(np.array(Fe_rows))  # [i][j] = Fe_ij -> column-major buffer
        R = np.zeros(9)
        product.sc_rotational_part(Fe.ravel(order="F").copy().ctypes.data_as(C.POINTER(C.c_double)),
                                   R.ctypes.data_as(C.POINTER(C.c_double)))
        om = R.reshape(3, 3, order="F").T.ravel(order="F").copy()
        q = np.zeros(4)
        product.sc_om2qu(om.ctypes.data_as(C.POINTER(C.c_double)), q.ctypes.data_as(C.POINTER(C.c_double)))
        O_ref = np.array(O_ref)
        assert abs(q[0] - O_ref[0]) < 1e-15
        for k in (1, 2, 3):
            assert abs(q[k] - O_ref[k]) <= 1e-6 * abs(O_ref[k]), (k, q, O_ref)


def test_om2qu_inverts_qu2om(product):
    """om2qu(qu2om(q)) = q for unit quaternions with w > 0 (every branch of
    the matrix-to-quaternion conversion: trace dominant and each diagonal)."""
    rng = np.random.default_rng(3)
    qs = [rng.standard_normal(4) for _ in range(200)]
    qs += [np.array([0.05, 1, 0.1, 0.2]), np.array([0.05, 0.1, 1, 0.2]), np.array([0.05, 0.1, 0.2, 1])]
    dp = C.POINTER(C.c_double)
    for q in qs:
        q = q / np.linalg.norm(q)
        if q[0] < 0:
            q = -q
        om, back = np.zeros(9), np.zeros(4)
        product.sc_qu2om(q.ctypes.data_as(dp), om.ctypes.data_as(dp))
        product.sc_om2qu(om.ctypes.data_as(dp), back.ctypes.data_as(dp))
        np.testing.assert_allclose(back, q, atol=1e-14)


def _objects(f: H5File, path="/"):
    """(path, shape or None, attribute names) of every object below path."""
    import struct
    out = []
    oh = f.resolve(path)
    msgs = f.messages(oh)
    attrs = sorted(f.attrs(path))
    if any(t == 0x08 for t, _ in msgs):
        dims = f._dataspace([d for t, d in msgs if t == 0x01][0])
        out.append((path, tuple(dims), attrs))
    else:
        out.append((path, None, attrs))
        for k in sorted(f.children(oh)):
            out += _objects(f, path.rstrip("/") + "/" + k)
    return out


@pytest.mark.gpu
def test_rve_result_file(product, goldens, hdf5, tmp_path):
    """A solved RVE's job file: increments at the load step's f_out cadence,
    the reference file's object layout (paths, shapes, attribute names), and
    the values of the solver state it was written from."""
    from solver_cpp_amd.geometry import polycrystal
    from solver_cpp_amd.rve import Rve, tension_x

    st = goldens["dadf5_2x2x2_storage"]
    cells = (6, 5, 4)
    N = int(np.prod(cells))
    mat, ori, size = polycrystal(cells, 3, seed=11)
    print(f"rve result file: cells {cells}", flush=True)
    r = Rve(product, cells, size, mat, ori)
    r.init()
    path = str(tmp_path / "job.hdf5")
    r.result_open(path, st["material_yaml"], st["numerics_yaml"])
    load = tension_x(t=2.0, N=4)
    load.f_out = 2
    r.loop_over_loadcase([load])
    r.result_write(99, 2.0)
    r.result_close()

    f = H5File(path)
    assert f.keys("/") == ["cell_to", "increment_1", "increment_3", "increment_99", "setup"]
    assert f.attrs("/increment_3")["t/s"] == pytest.approx(1.5, abs=1e-15)
    # layout of the reference's file, with this grid's sizes
    nodes = (cells[0] + 1) * (cells[1] + 1) * cells[2]
    want = []
    for o in st["objects"]:
        p = o["path"]
        shape = None
        if "shape" in o:
            shape = tuple({8: N, 18: nodes}.get(d, d) if k == 0 else d for k, d in enumerate(o["shape"]))
            if p.startswith("/setup"):
                shape = (1,)
        want.append((p.replace("increment_1", "increment_99"), shape, o["attrs"]))
    got = [o for o in _objects(f) if not o[0].startswith(("/increment_1", "/increment_3"))]
    assert got == [w for w in want if not w[0].startswith("/increment_1")]
    # values: the state after the last increment
    m = "/increment_99/phase/Aluminum/mechanical/"

    def rows(name):  # get_field (3,3,N) column-major -> [c][i][j]
        return r.field(name).reshape(N, 3, 3).transpose(0, 2, 1)

    for name, field in (("P", "P"), ("F_p", "Fp"), ("L_p", "Lp"), ("F_e", "Fe")):
        np.testing.assert_array_equal(f.read(m + name), rows(field), err_msg=name)
    np.testing.assert_allclose(f.read(m + "F"), rows("F"), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(f.read(m + "xi_sl"), r.field("xi").reshape(N, 12))
    O = f.read(m + "O")
    dp = C.POINTER(C.c_double)
    for c in range(0, N, 7):
        R, q = np.zeros(9), np.zeros(4)
        Fe = f.read(m + "F_e")[c].ravel(order="F").copy()
        product.sc_rotational_part(Fe.ctypes.data_as(dp), R.ctypes.data_as(dp))
        om = R.reshape(3, 3, order="F").T.ravel(order="F").copy()
        product.sc_om2qu(om.ctypes.data_as(dp), q.ctypes.data_as(dp))
        np.testing.assert_array_equal(O[c], q)
    x_p = r.field("x_p").reshape(N, 3)
    step = np.array(size) / np.array(cells)
    i, j, k = np.meshgrid(*[np.arange(n) for n in cells], indexing="ij")
    x0 = np.stack([(i.T.ravel() + 0.5) * step[0], (j.T.ravel() + 0.5) * step[1], (k.T.ravel() + 0.5) * step[2]], 1)
    np.testing.assert_allclose(f.read("/increment_99/geometry/u_p"), x_p - x0, rtol=0, atol=1e-18)
    labels = f.attrs(m + "xi_sl")["systems"]
    assert [s.decode() for s in labels][:2] == ["[ 0  1 -1]( 1  1  1)", "[-1  0  1]( 1  1  1)"]
    # the HDF5 library reads it, checksums verified
    h = hdf5.open(path)
    np.testing.assert_array_equal(hdf5.read_f64(h, m + "P"), rows("P"))
    rec = hdf5.read_raw(h, "/cell_to/phase")
    assert rec[:16] == b"Aluminum" + (0).to_bytes(8, "little") and len(rec) == 16 * N
    hdf5.h.H5Fclose(h)
