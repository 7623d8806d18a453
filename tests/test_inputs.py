"""Job inputs read by the HIP library's host code: the reference's 16^3 VTI
fixture, the grid example's material.yaml, numerics.yaml and tensionX.yaml.
CPU only (the readers run on the host)."""
import base64
import ctypes as C
import os
import re
import zlib

import numpy as np
import pytest

from conftest import GOLDEN, PRODUCT_SO
from solver_cpp_amd.capi import GridNumerics, Library, LoadStep, Material, PhaseNumerics, ptr

IP = C.POINTER(C.c_int)
VTI = os.path.join(GOLDEN, "grid16_20grains.vti")


@pytest.fixture(scope="module")
def lib():
    return Library(PRODUCT_SO, boundary=True)


def _vti_material_by_hand(text):
    """Independent decoding of the 'material' array: UInt32 header
    [nblocks, block size, last size, compressed sizes...], zlib blocks, Int64."""
    body = re.search(r'Name="material"[^>]*>\s*([A-Za-z0-9+/=\s]+?)\s*</DataArray>', text).group(1)
    body = "".join(body.split())
    head = base64.b64decode(body[:8])
    nblock = int(np.frombuffer(head[:4], "<u4")[0])
    hlen = 4 * (3 + nblock)
    hchars = (hlen + 2) // 3 * 4
    h = np.frombuffer(base64.b64decode(body[:hchars]), "<u4")
    data = base64.b64decode(body[hchars:])
    out, off = b"", 0
    for b in range(nblock):
        out += zlib.decompress(data[off:off + int(h[3 + b])])
        off += int(h[3 + b])
    return np.frombuffer(out, "<i8")


def test_vti_cells_size_origin_match_reference_test(lib):
    """test_discretization_grid.cpp VTISetup: cells 16^3, size (1,1,1), origin 0."""
    text = open(VTI, "rb").read()
    cells, size, origin = np.zeros(3, np.int32), np.zeros(3), np.ones(3)
    lib.f_VTI_readCellsSizeOrigin(text, len(text), ptr(cells, IP), ptr(size), ptr(origin))
    assert cells.tolist() == [16, 16, 16]
    assert size.tolist() == [1.0, 1.0, 1.0]
    assert origin.tolist() == [0.0, 0.0, 0.0]


def test_vti_material_dataset(lib):
    text = open(VTI, "rb").read()
    n = np.array([4096], np.int32)
    ds = np.zeros(4096, np.int32)
    label = b"material"
    lib.f_VTI_readDataset_int(text, len(text), label, len(label), ptr(ds, IP), ptr(n, IP))
    assert n[0] == 4096
    ref = _vti_material_by_hand(text.decode())
    np.testing.assert_array_equal(ds, ref + 1)  # returned 1-based (VTI.f90:494)
    assert ref.min() == 0 and ref.max() == 19 and len(np.unique(ref)) == 20
    cells, size, origin = np.zeros(3, np.int32), np.zeros(3), np.zeros(3)
    mat = np.zeros(4096, np.int32)
    assert lib.sc_read_vti(VTI.encode(), ptr(cells, IP), ptr(size), ptr(origin), ptr(mat, IP), 4096) == 0
    np.testing.assert_array_equal(mat, ref)


def test_vti_errors_are_reported(lib):
    bad = b'<VTKFile type="PolyData"></VTKFile>'
    cells, size, origin = np.zeros(3, np.int32), np.zeros(3), np.zeros(3)
    lib.f_VTI_readCellsSizeOrigin(bad, len(bad), ptr(cells, IP), ptr(size), ptr(origin))
    assert cells.tolist() == [-1, -1, -1]
    assert b"ImageData" in lib.sc_last_error()


def test_material_yaml(lib):
    m = Material()
    n = C.c_int()
    ori = np.zeros((32, 4))
    assert lib.sc_read_material(os.path.join(GOLDEN, "material.yaml").encode(), C.byref(m), ptr(ori), 32,
                                C.byref(n)) == 0, lib.sc_last_error()
    assert n.value == 20
    assert (m.C_11, m.C_12, m.C_44) == (106.75e9, 60.41e9, 28.34e9)
    assert (m.xi_0_sl, m.xi_inf_sl, m.h_0_sl_sl) == (31e6, 63e6, 75e6)
    assert (m.n_sl, m.a_sl, m.dot_gamma_0_sl) == (20.0, 2.25, 1e-3)
    assert list(m.h_sl_sl) == [1, 1, 1.4, 1.4, 1.4, 1.4, 1.4]
    assert (m.atol_xi, m.atol_gamma) == (1.0, 1e-6)
    np.testing.assert_array_equal(ori[0], [1, 0, 0, 0])
    np.testing.assert_array_equal(ori[1], [0.7936696712125002, -0.28765777461664166, -0.3436487135089419,
                                           0.4113964260949434])
    np.testing.assert_allclose(np.linalg.norm(ori[:20], axis=1), 1.0, rtol=1e-14)


def test_numerics_yaml(lib):
    g, p = GridNumerics(), PhaseNumerics()
    lib.sc_default_grid_numerics(C.byref(g))
    lib.sc_default_phase_numerics(C.byref(p))
    assert lib.sc_read_numerics(os.path.join(GOLDEN, "numerics.yaml").encode(), C.byref(g), C.byref(p)) == 0
    assert (g.itmin, g.itmax) == (4, 40)
    assert g.eps_rel_div == 5e-4 and g.divergence_correction == 2  # untouched defaults


def test_load_yaml(lib):
    steps = (LoadStep * 4)()
    n = C.c_int()
    assert lib.sc_read_load(os.path.join(GOLDEN, "tensionX.yaml").encode(), steps, 4, C.byref(n)) == 0, \
        lib.sc_last_error()
    assert n.value == 1
    s = steps[0]
    assert s.deformation_type == 1 and s.stress_type == 0  # dot_F, P
    assert (s.t, s.N, s.r, s.f_out, s.estimate_rate) == (5.0, 10, 1.0, 4, 1)
    np.testing.assert_array_equal(np.array(s.deformation), [1e-3, 0, 0, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(np.array(s.deformation_mask), [0, 0, 0, 0, 1, 0, 0, 0, 1])
    np.testing.assert_array_equal(np.array(s.stress_mask), [1, 1, 1, 1, 0, 1, 1, 1, 0])
    assert list(s.rot_q) == [1, 0, 0, 0]


def test_yaml_forms(lib, tmp_path):
    """Block sequences of flow rows, multi-line flow rows, comments, R."""
    f = tmp_path / "load.yaml"
    f.write_text("""---
solver:
  mechanical: spectral_basic   # comment
loadstep:
  - boundary_conditions:
      mechanical:
        F: [[1.01, 0, 0],
            [0, x, 0],
            [0, 0, x]]
        dot_P:
          - [x, x, x]
          - [x, 5.0e+6, x]
          - [x, x, 0]
        R: [0.5, 0.5, 0.5, 0.5]
    discretization:
      t: 20
      N: 40
      r: 2
    estimate_rate: false
""")
    steps = (LoadStep * 2)()
    n = C.c_int()
    assert lib.sc_read_load(str(f).encode(), steps, 2, C.byref(n)) == 0, lib.sc_last_error()
    s = steps[0]
    assert s.deformation_type == 0 and s.stress_type == 1
    assert s.deformation[0] == 1.01 and s.stress[4] == 5e6 and s.stress_mask[4] == 0
    assert list(s.rot_q) == [0.5] * 4
    assert (s.t, s.N, s.r, s.estimate_rate) == (20.0, 40, 2.0, 0)
