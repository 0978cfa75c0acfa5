"""libomni3d.so on the CPU: it loads, exports every entry point its header
declares, reports the missing GPU as an error (no silent fallback), and its
host-only direction generator produces icosahedral sets."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import omni3d
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "omni3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|const char\s*\*|int64_t)\s+\**\s*(\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for n in ("omni3d_create", "omni3d_reconstruct", "omni3d_es_inner", "omni3d_es_inner_select",
              "omni3d_devide_count_inner", "sor3D", "calRHS", "omni3d_forward",
              "omni3d_backward", "omni3d_run_parameter_file", "initialIntegration",
              "omni3dparallellinesESInner", "omni3dparallellinesESInnerWeighted",
              "omni3dparallellinesESInnerSelect", "omni3dparallellinesESInnerSelectFixedBC",
              "omni3dparallellinesESInnerWeightedMiniCurl", "devidecountInner",
              "omni3d_set_velocity"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(omni3d.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares the same set
    assert set(omni3d.EXPORTS) <= set(declared_functions())


def test_default_params_match_reference_defaults():
    p = omni3d.default_params()
    assert (p.dx, p.dy, p.dz, p.density, p.line_spacing) == (1.0, 1.0, 1.0, 1.0, 1.0)
    assert p.method == omni3d.PLAIN and p.precision == omni3d.FP64
    assert p.max_iterations == 10            # Parameter_Omni3D.dat line 14
    assert p.seed_pn_with_p == 1             # kernel.cu:8813
    assert p.traversal == omni3d.STAIRCASE   # kernel.cu:1764-1840


def test_trilinear_needs_plain_affine():
    for kw in ({"method": omni3d.SELECT}, {"method": omni3d.PLAIN_MASK},
               {"solver": omni3d.NAIVE}, {"traversal": 2}):
        with pytest.raises(omni3d.Omni3DError) as e:
            omni3d.Context((8, 8, 8), **dict({"traversal": omni3d.TRILINEAR}, **kw))
        assert e.value.code == omni3d.ERR_ARG


def test_reference_signatures_checked_before_device_work():
    L = omni3d.lib()
    # initialIntegration / omni3d_initial_integration: grid and pointers first
    assert L.omni3d_initial_integration(0, 4, 4, 1.0, 1.0, 1.0, 1.0, None, None, None, 0.0,
                                        None, None) == omni3d.ERR_ARG
    # the reference-order drop-ins: long dims, then the index pointer (not
    # read); a null field pointer is an argument error
    es = [8, 8, 8, 1, 1.0, None, None, None, None, 1.0, 1.0, 1.0, 1.0, None, None, None]
    assert L.omni3d_es_inner(*es, None, None, None, None) == omni3d.ERR_ARG
    assert L.omni3d_es_inner_select(*es, None, None, None, None, 0.5, None) == omni3d.ERR_ARG
    # the void reference-named forms report through the global error
    L.omni3dparallellinesESInner(*es, None, None, None)
    assert b"null" in L.omni3d_last_global_error()


def test_no_gpu_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(omni3d.Omni3DError) as e:
        omni3d.Context((8, 8, 8))
    assert e.value.code == omni3d.ERR_NODEVICE


def test_bad_arguments_rejected_before_any_device_work():
    with pytest.raises(omni3d.Omni3DError) as e:
        omni3d.Context((8, 8, 1))
    assert e.value.code == omni3d.ERR_ARG
    with pytest.raises(omni3d.Omni3DError) as e:
        omni3d.Context((8, 8, 8), line_spacing=0.0)
    assert e.value.code == omni3d.ERR_ARG
    with pytest.raises(omni3d.Omni3DError) as e:
        omni3d.Context((8, 8, 8), method=omni3d.WEIGHTED_MINICURL + 1)
    assert e.value.code == omni3d.ERR_ARG
    L = omni3d.lib()
    out = C.c_float(0)
    # threshold histogram: grid and pointers checked first
    assert L.omni3d_threshold_hist(0, 4, 4, None, 0.5, C.c_void_p(C.addressof(out)),
                                   omni3d.HOST, None) == omni3d.ERR_ARG
    buf = np.zeros(8, np.float32)
    assert L.omni3d_threshold_hist(2, 2, 2, C.c_void_p(buf.ctypes.data), 0.5, None,
                                   omni3d.HOST, None) == omni3d.ERR_ARG
    assert L.omni3d_threshold_hist(2, 2, 2, C.c_void_p(buf.ctypes.data), 0.5,
                                   C.c_void_p(C.addressof(out)), 7, None) == omni3d.ERR_ARG
    # boundary pairs: a context is required
    assert L.omni3d_boundary_pairs(None, 4, C.c_void_p(buf.ctypes.data), omni3d.F32,
                                   omni3d.HOST, None) == omni3d.ERR_ARG
    # the variant launchers check their pointers before any device work
    es = [8, 8, 8, 1, 1.0, None, None, None, None, 1.0, 1.0, 1.0, 1.0, None, None, None]
    assert L.omni3d_es_inner_select_fixedbc(*es, None, None, None, None, 0.5,
                                            None) == omni3d.ERR_ARG
    assert L.omni3d_es_inner_weighted_minicurl(*es, None, None, None, None,
                                               None) == omni3d.ERR_ARG


@pytest.mark.parametrize("divide", [1, 2, 3, 4, 5])
def test_generated_directions_are_icosahedral(divide):
    k = omni3d.generate_directions(divide).astype(np.float64)
    n = 10 * 4 ** (divide - 1) + 2
    assert k.shape == (n, 3)
    np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, atol=1e-6)
    # exact antipodal pairs
    d = np.abs(k[:, None, :] + k[None, :, :]).max(axis=2)
    assert np.all(d.min(axis=1) == 0.0)
    # pole-aligned: (0, 0, +-1) is a vertex
    assert np.any(np.all(k == [0.0, 0.0, 1.0], axis=1))
    if divide == 1:
        # the 12 icosahedron vertices: dot products 1, -1, +-1/sqrt(5)
        g = np.round(np.abs(k @ k.T), 6)
        assert set(np.unique(g)) == {round(1 / np.sqrt(5), 6), 1.0}
    else:
        # 12 vertices keep 5 nearest neighbours, all others have 6
        dist = np.linalg.norm(k[:, None] - k[None], axis=2)
        np.fill_diagonal(dist, np.inf)
        dmin = dist.min(axis=1, keepdims=True)
        nn = (dist < dmin * 1.3).sum(axis=1)
        assert (nn == 5).sum() == 12 and (nn == 6).sum() == n - 12


def test_fixture_direction_sets(ref_dirs2, ref_dirs3):
    """The reference's GridIcosahedronDivide{2,3}.dat: 42 and 162 unit
    vectors (to the 6 printed digits) in near-antipodal pairs: the shipped
    sets are a slightly perturbed icosahedral grid, not the exact one the
    generator above builds, so parity tests load these files."""
    for k, n in ((ref_dirs2, 42), (ref_dirs3, 162)):
        k = k.astype(np.float64)
        assert k.shape == (n, 3)
        np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, atol=2e-6)
        d = np.abs(k[:, None, :] + k[None, :, :]).max(axis=2)
        assert d.min(axis=1).max() < 0.02


def test_parameter_file_reader():
    """Parameter_Omni3D.dat (kernel.cu:8540-8571): the shipped file's first
    22 lines, then a file with pref and a region of interest on lines 16-22."""
    import os
    from conftest import GOLDEN
    pf = omni3d.read_parameter_file(os.path.join(GOLDEN, "Parameter_Omni3D.dat"))
    assert (pf.nx, pf.ny, pf.nz) == (64, 64, 64)
    assert (pf.dx, pf.dy, pf.dz, pf.density, pf.scale) == (0.00024, 0.00024, 0.00024, 1000.0, 1)
    assert pf.line_spacing == 0.7 and pf.n_angles == 10242 and pf.n_iterations == 10
    assert pf.grid_file == b"GridIcosahedronDivide6.dat"
    assert pf.acc_dir == b".\\AccelerationFilesSample\\"
    assert pf.threshold == 52921780.0 and pf.pref == 0.0
    assert list(pf.roi) == [0, 63, 0, 63, 0, 63]


def test_parameter_file_roi(tmp_path):
    lines = ["12", "10", "8", "0.5", "0.5", "0.5", "1.2", "1", "1", "42",
             "GridIcosahedronDivide2.dat", "acc", "out", "3", "7.5", "101325", "1", "10", "0", "8",
             "2", "6"]
    p = tmp_path / "Parameter_Omni3D.dat"
    p.write_text("\r\n".join(lines) + "\r\n")
    pf = omni3d.read_parameter_file(p)
    assert (pf.nx, pf.ny, pf.nz, pf.n_iterations) == (12, 10, 8, 3)
    assert pf.pref == 101325.0 and pf.threshold == 7.5
    assert list(pf.roi) == [1, 10, 0, 8, 2, 6]
    bad = tmp_path / "bad.dat"
    bad.write_text("\n".join(lines[:16] + ["0", "12"] + lines[18:]) + "\n")
    with pytest.raises(omni3d.Omni3DError):
        omni3d.read_parameter_file(bad)
