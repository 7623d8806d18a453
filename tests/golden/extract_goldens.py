#!/usr/bin/env python3
"""Write the reference's golden vectors as small JSON/NPZ fixtures.

Reads (as text) the reference's GoogleTest sources under
``src/grid/test/`` and the original-DAMASK stdout trace
``output_damask_f.sh``; writes ``tests/golden/reference_goldens.json`` and
``tests/golden/fortran_trace.npz``.  Only data (inputs and expected outputs)
is written; no reference source text is kept.

Usage: python tests/golden/extract_goldens.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TESTDIR = os.path.join(REF, "src", "grid", "test")

_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def _parse_braces(text: str, start: int):
    """Parse a nested brace list starting at text[start] == '{'.
    Returns (nested python list, index after the closing brace)."""
    assert text[start] == "{"
    stack = [[]]
    i = start + 1
    while stack:
        ch = text[i]
        if ch == "{":
            stack.append([])
            i += 1
        elif ch == "}":
            top = stack.pop()
            if stack:
                stack[-1].append(top)
            else:
                return top, i + 1
            i += 1
        elif text.startswith("c(", i):
            j = text.index(")", i)
            re_, im_ = [float(x) for x in text[i + 2:j].split(",")]
            stack[-1].append([re_, im_])
            i = j + 1
        else:
            m = re.match(_NUM, text[i:])
            if m and (i == 0 or not (text[i - 1].isalnum() or text[i - 1] == "_")):
                stack[-1].append(float(m.group(0)))
                i += len(m.group(0))
            else:
                i += 1
    raise ValueError("unbalanced")


def set_values(src: str, name: str, occurrence: int = 0):
    """Nested list of `name.setValues({...})` (first index outermost)."""
    pat = re.compile(re.escape(name) + r"\.setValues\(\{")
    matches = list(pat.finditer(src))
    m = matches[occurrence]
    lst, _ = _parse_braces(src, m.end() - 1)
    return lst


def in_test(src: str, test: str) -> str:
    i = src.index(test)
    j = src.find("TEST_F(", i + 1)
    return src[i: j if j > 0 else len(src)]


def tensor_colmajor(nested, shape, complex_=False):
    """Nested list with first index outermost -> flat Fortran-order list."""
    a = np.array(nested, dtype=np.float64)
    if complex_:
        a = a[..., 0] + 1j * a[..., 1]
    a = a.reshape(shape)
    flat = a.reshape(-1, order="F")
    if complex_:
        return {"re": flat.real.tolist(), "im": flat.imag.tolist(), "shape": list(shape)}
    return {"data": flat.tolist(), "shape": list(shape)}


def read(name):
    with open(os.path.join(TESTDIR, name)) as f:
        return f.read()


def main():
    sys.path.insert(0, HERE)
    out = {}
    sp = read("test_spectral.cpp")
    t = in_test(sp, "SpectralTestInit")
    out["spectral_init_2x1x1"] = {
        "cells": [2, 1, 1], "size": [2e-5, 1e-5, 1e-5], "derivative": 0,
        "xi1st": tensor_colmajor(set_values(t, "expected_xi1st"), (3, 2, 1, 1), True),
        "xi2nd": tensor_colmajor(set_values(t, "expected_xi2nd"), (3, 2, 1, 1), True),
        "source": "src/grid/test/test_spectral.cpp:15-48"}
    t = in_test(sp, "SpectralTestConstitutiveResponse")
    out["constitutive_response_2x1x1"] = {
        "F": tensor_colmajor(set_values(t, "F"), (3, 3, 2, 1, 1)),
        "P": tensor_colmajor(set_values(t, "mech2_homogenization_P_result"), (3, 3, 2)),
        "dPdF": tensor_colmajor(set_values(t, "mech2_homogenization_dPdF_result"), (3, 3, 3, 3, 2)),
        "P_av": tensor_colmajor(set_values(t, "expected_P_av"), (3, 3)),
        "C_volAvg": tensor_colmajor(set_values(t, "expected_C_vol_avg"), (3, 3, 3, 3)),
        "C_minMaxAvg": tensor_colmajor(set_values(t, "expected_C_minmax_avg"), (3, 3, 3, 3)),
        "source": "src/grid/test/test_spectral.cpp:58-216"}

    mb = read("test_mech_base.cpp")
    t = in_test(mb, "MechBaseTestInitDivergenceCorrection")
    out["scaled_geom_size"] = {"cells": [2, 1, 1], "size": [3, 4, 5], "none": [3, 4, 5],
                               "size_only": [0.75, 1, 1.25], "size_grid": [0.6, 0.8, 1],
                               "source": "src/grid/test/test_mech_base.cpp:55-77"}
    t = in_test(mb, "MechBaseTestUpdateCoords")
    out["update_coords_2x1x1"] = {
        "F": tensor_colmajor(set_values(t, "F_"), (3, 3, 2, 1, 1)),
        "x_n": tensor_colmajor(set_values(t, "expected_x_n"), (3, 3, 2, 2)),
        "x_p": tensor_colmajor(set_values(t, "expected_x_p"), (3, 2, 1, 1)),
        "source": "src/grid/test/test_mech_base.cpp:79-152"}
    t = in_test(mb, "MechBaseTestForwardField")
    out["forward_field_2x1x1"] = {
        "field_last_inc": tensor_colmajor(set_values(t, "field_last_inc"), (3, 3, 2, 1, 1)),
        "rate": tensor_colmajor(set_values(t, "rate"), (3, 3, 2, 1, 1)),
        "aim": tensor_colmajor(set_values(t, "aim"), (3, 3)),
        "delta_t": 1.0,
        "expected": tensor_colmajor(set_values(t, "expected_forwarded_field"), (3, 3, 2, 1, 1)),
        "source": "src/grid/test/test_mech_base.cpp:199-237"}
    t = in_test(mb, "MechBaseTestMaskedCompliance")
    out["masked_compliance"] = {
        "C": tensor_colmajor(set_values(t, "C"), (3, 3, 3, 3)),
        "rot_q": [1.0, 0.0, 0.0, 0.0],
        "stress_mask_rowmajor": [[1, 1, 1], [1, 0, 1], [1, 1, 0]],
        "expected": tensor_colmajor(set_values(t, "expected_masked_compliance"), (3, 3, 3, 3)),
        "source": "src/grid/test/test_mech_base.cpp:239-321"}
    t = in_test(mb, "MechBaseTestDivergenceRMS")
    rms = float(re.search(r"ASSERT_EQ\(rms,\s*(" + _NUM + r")\)", t).group(1))
    out["divergence_rms_4x1x1"] = {
        "cells": [4, 1, 1], "size": [4e-5, 1e-5, 1e-5], "scaled_size": [4, 1, 1], "wgt": 0.25,
        "field": tensor_colmajor(set_values(t, "tensor_field"), (3, 3, 4, 1, 1)),
        "xi1st": tensor_colmajor(set_values(t, "spectral.xi1st"), (3, 3, 1, 1), True),
        "expected_rms": rms,
        "source": "src/grid/test/test_mech_base.cpp:323-387"}
    t = in_test(mb, "MechBaseTestGammaConvolution")
    out["gamma_convolution_2x1x1"] = {
        "field_aim": tensor_colmajor(set_values(t, "field_aim"), (3, 3)),
        "C_ref": tensor_colmajor(set_values(t, "mech_base.C_ref"), (3, 3, 3, 3)),
        "field": tensor_colmajor(set_values(t, "field"), (3, 3, 2, 1, 1)),
        "expected": tensor_colmajor(set_values(t, "expected_gamma_field"), (3, 3, 2, 1, 1)),
        "source": "src/grid/test/test_mech_base.cpp:389-490"}
    t = in_test(mb, "TestCalculateRate")
    out["calculate_rate_2x1x1"] = {
        "field": tensor_colmajor(set_values(t, "field"), (3, 3, 2, 1, 1)),
        "field_0": tensor_colmajor(set_values(t, "field_0"), (3, 3, 2, 1, 1)),
        "avRate": tensor_colmajor(set_values(t, "avRate"), (3, 3)),
        "dt": 100.2, "heterogeneous": 1,
        "expected": tensor_colmajor(set_values(t, "expected_rate"), (3, 3, 2, 1, 1)),
        "source": "src/grid/test/test_mech_base.cpp:492-545"}

    ms = read("test_mech_solver_basic.cpp")
    t = in_test(ms, "TestBasicForward")
    out["basic_forward_2x1x1"] = {
        "dot_F": [[1e-3, 0, 0], [0, "x", 0], [0, 0, "x"]],
        "P": [["x", "x", "x"], ["x", 0, "x"], ["x", "x", 0]],
        "cut_back": 0, "guess": 0, "delta_t": 100.2, "delta_t_old": 1.0, "t_remaining": 1002.0,
        "expected_aim": tensor_colmajor(set_values(t, "expected_aim"), (3, 3)),
        "expected_F": tensor_colmajor(set_values(t, "forwarded_field"), (3, 3, 2, 1, 1)),
        "source": "src/grid/test/test_mech_solver_basic.cpp:160-230"}
    t = in_test(ms, "TestFormResidual")
    err_bc = float(re.search(r"EXPECT_EQ\(mech_basic\.err_BC,\s*(" + _NUM + r")\)", t).group(1))
    out["form_residual_bc"] = {
        "S": tensor_colmajor(set_values(t, "mech_basic.S"), (3, 3, 3, 3)),
        "P_av": tensor_colmajor(set_values(t, "mech_basic.P_av"), (3, 3)),
        "F_aim": [[1.1002000000000001, 0, 0], [0, 1, 0], [0, 0, 1]],
        "stress_mask_rowmajor": [[1, 1, 1], [1, 0, 1], [1, 1, 0]],
        "expected_rotated": tensor_colmajor(set_values(t, "expected_rotated"), (3, 3)),
        "expected_F_aim": tensor_colmajor(set_values(t, "expected_F_aim"), (3, 3)),
        "expected_err_BC": err_bc,
        "source": "src/grid/test/test_mech_solver_basic.cpp:232-376"}

    # ---- DAMASK result file: 2x2x2 single crystal, first tension increment ---
    # (read with the own minimal HDF5 parser hdf5_read.py; numbers only)
    from hdf5_read import H5File
    h5 = H5File(os.path.join(REF, "2x2x2_tensionX_material_numerics.hdf5"))
    base = "/increment_1/phase/Aluminum/mechanical/"
    O = h5.read(base + "O")
    assert np.allclose(O, [1.0, 0.0, 0.0, 0.0], atol=1e-12)  # every cell: material.yaml entry 1
    assert b"N_iter_max: 40" in h5.read("/setup/numerics.yaml")[0]
    out["dadf5_2x2x2_increment_1"] = {
        "cells": [2, 2, 2], "t": float(h5.attrs("/increment_1")["t/s"]),
        "orientation": [1.0, 0.0, 0.0, 0.0], "N_iter_min": 4, "N_iter_max": 40,
        # (N,3,3) row-major per cell as stored: [c][i][j] = X_ij
        **{k: h5.read(base + src).tolist() for k, src in
           (("F", "F"), ("P", "P"), ("Fp", "F_p"), ("Lp", "L_p"), ("Fe", "F_e"), ("xi", "xi_sl"))},
        "O": O.tolist(),
        "u_n": h5.read("/increment_1/geometry/u_n").tolist(),
        "source": "2x2x2_tensionX_material_numerics.hdf5:/increment_1 (DADF5 3.0.0-alpha8)"}

    # ---- the same file's storage: every object, its shape and attribute
    # names, and the stored Fletcher-32 checksum of each single-chunk dataset
    # (values are in the group above; the setup texts are kept here)
    import struct
    objects, checksums = [], {}

    def walk(path):
        oh = h5.resolve(path)
        kids = h5.children(oh)
        attrs = sorted(h5.attrs(path)) if path != "/" else sorted(h5.attrs("/"))
        msgs = h5.messages(oh)
        layout = [d for t, d in msgs if t == 0x08]
        if layout:
            dims = h5._dataspace([d for t, d in msgs if t == 0x01][0])
            objects.append({"path": path, "shape": list(dims), "attrs": attrs,
                            "layout": ["compact", "contiguous", "chunked"][layout[0][1]]})
            if layout[0][1] == 2:
                chunks = list(h5._chunks(struct.unpack_from("<Q", layout[0], 3)[0], layout[0][2] - 1))
                if len(chunks) == 1:
                    checksums[path] = struct.unpack_from("<I", chunks[0][1], len(chunks[0][1]) - 4)[0]
        else:
            objects.append({"path": path, "attrs": attrs, "group": True})
            for k in sorted(kids):
                walk(path.rstrip("/") + "/" + k)

    walk("/")
    out["dadf5_2x2x2_storage"] = {
        "objects": objects, "fletcher32": checksums,
        "material_yaml": h5.read("/setup/material.yaml")[0].decode(),
        "numerics_yaml": h5.read("/setup/numerics.yaml")[0].decode(),
        "source": "2x2x2_tensionX_material_numerics.hdf5 (DADF5 3.0.0-alpha8)"}

    with open(os.path.join(HERE, "reference_goldens.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")

    # ---- original-DAMASK trace (3x3x3, 20 grains, tensionX t=5 N=10) ----------
    with open(os.path.join(REF, "output_damask_f.sh"), errors="replace") as f:
        lines = f.read().splitlines()
    evals_F, evals_r, inc_of_eval, pav = [], [], [], []
    inc = 0
    conv_incs = []
    cur_F = None
    for ln in lines:
        s = ln.strip()
        m = re.match(r"Time\s+\S+s: Increment (\d+)/", s)
        if m:
            inc = int(m.group(1))
            continue
        if s.startswith("FR F1"):
            cur_F = np.array([float(x) for x in s[5:].split()])
            continue
        if s.startswith("FR r2"):
            r = np.array([float(x) for x in s[5:].split()])
            assert cur_F is not None and cur_F.size == 243 and r.size == 243
            evals_F.append(cur_F)
            evals_r.append(r)
            inc_of_eval.append(inc)
            cur_F = None
            continue
        m = re.match(r"increment (\d+) converged", s)
        if m:
            conv_incs.append(int(m.group(1)))
    np.savez_compressed(os.path.join(HERE, "fortran_trace.npz"), F=np.array(evals_F), r=np.array(evals_r),
                        increment=np.array(inc_of_eval, dtype=np.int32),
                        converged_increments=np.array(conv_incs, dtype=np.int32))
    print(f"wrote {len(out)} golden groups, {len(evals_F)} trace evaluations")
    write_job_inputs()


def write_job_inputs():
    """Input data of the reference's grid example and geometry fixture:
    - grid16_20grains.vti: the 16^3 VTK image data held as a string by the
      reference's VTISetup fixture (test_discretization_grid.cpp);
    - material.yaml / numerics.yaml / tensionX.yaml: the values of
      examples/grid/{material,numerics,tensionX}.yaml, re-emitted from the
      parsed data (not the original text)."""
    import yaml

    src = read("test_discretization_grid.cpp")
    a = src.index('R""""(') + len('R""""(')
    b = src.index(')""""', a)
    vti = src[a:b].strip() + "\n"
    with open(os.path.join(HERE, "grid16_20grains.vti"), "w") as f:
        f.write(vti)
    ex = os.path.join(REF, "examples", "grid")

    def load(name):
        with open(os.path.join(ex, name)) as f:
            return yaml.safe_load(f)

    def emit(name, data, header):
        with open(os.path.join(HERE, name), "w") as f:
            f.write(header)
            yaml.safe_dump(data, f, default_flow_style=None, sort_keys=False, width=100)

    emit("material.yaml", load("material.yaml"),
         "# values of the grid example's material.yaml: one cF phenopowerlaw phase, 20 entries\n")
    emit("numerics.yaml", load("numerics.yaml"), "# values of the grid example's numerics.yaml\n")
    emit("tensionX.yaml", load("tensionX.yaml"), "# values of the grid example's tensionX.yaml\n")
    print("wrote grid16_20grains.vti, material.yaml, numerics.yaml, tensionX.yaml")


if __name__ == "__main__":
    main()
