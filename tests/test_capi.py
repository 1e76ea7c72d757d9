"""C-ABI boundary tests that need no GPU: exports, model reading, errors."""
import os
import re

import numpy as np
import pytest

from helpers import ROOT, chdir, prepare


@pytest.fixture(scope="module")
def lib():
    from openmc_amd import lib as _lib
    return _lib


def test_library_exports_every_declared_symbol(lib):
    header = open(os.path.join(ROOT, "include", "openmc_amd", "capi.h")).read()
    funcs = re.findall(r"^\s*(?:int|size_t|int64_t|uint64_t|void|const char\*|const int\*)\s+\**(\w+)\(",
                       header, re.M)
    globs = re.findall(r"^extern\s+\w+\s+(\w+)(?:\[\d+\])?;", header, re.M)
    assert len(funcs) > 25 and "openmc_init" in funcs and "openmc_tally_results" in funcs
    for name in funcs + globs:
        assert hasattr(lib._dll, name), name


def test_globals_have_reference_values(lib):
    from ctypes import c_int
    codes = {"OPENMC_E_UNASSIGNED": -1, "OPENMC_E_ALLOCATE": -2, "OPENMC_E_OUT_OF_BOUNDS": -3,
             "OPENMC_E_INVALID_SIZE": -4, "OPENMC_E_INVALID_ARGUMENT": -5, "OPENMC_E_INVALID_TYPE": -6,
             "OPENMC_E_INVALID_ID": -7, "OPENMC_E_GEOMETRY": -8, "OPENMC_E_DATA": -9,
             "OPENMC_E_PHYSICS": -10, "OPENMC_E_WARNING": 1}
    for k, v in codes.items():
        assert c_int.in_dll(lib._dll, k).value == v
    assert lib.settings.event_based in (False, True)
    assert lib.settings.run_mode == "fixed source"


def test_model_reading_cube(lib):
    d, model, nucs = prepare(0, n_particles=2000, n_batches=2)
    summ, dens, _ = lib.check_model(d)
    assert summ["n_nuclides"] == 4 and summ["n_cells"] == 1 and summ["n_surfaces"] == 6
    assert summ["n_levels"] == 1 and summ["n_tally_bins"] == 1000
    assert summ["energy_min"] == 1e-5 and summ["energy_max"] == 2e7
    # Material::normalize_density with "sum" units: N_i = (a_i / S) * S
    a = np.array([p for _, p, _ in model.materials[0].nuclides])
    S = 0.0
    for v in a:
        S += v
    assert np.array_equal(dens, (a / S) * S)


def test_model_reading_lattice(lib):
    d, model, nucs = prepare(1, n_particles=2000, n_batches=2, grid_lo=2000, grid_hi=6000, n_nuclides=12,
                             lattice_n=5, mesh=(10, 10, 20), half_height=20.0,
                             energy_bins=[1e-5, 1.0, 1e3, 1e5, 2e7])
    summ, dens, _ = lib.check_model(d)
    assert summ["n_levels"] == 2 and summ["n_lattices"] == 1 and summ["n_cells"] == 4
    assert summ["n_tally_bins"] == 10 * 10 * 20 * 4
    # atom/b-cm units: N_i = (a_i / sum a) * 0.1
    for k, m in enumerate(model.materials):
        a = np.array([p for _, p, _ in m.nuclides])
        S = 0.0
        for v in a:
            S += v
        assert np.array_equal(dens[4 * k:4 * (k + 1)], (a / S) * 0.1)


def test_errors_map_to_exceptions(lib, tmp_path):
    with pytest.raises(lib.OpenMCError):
        lib.check_model(str(tmp_path))
    (tmp_path / "model.xml").write_text("<model><settings><particles>10</particles></settings>")
    with pytest.raises(lib.InvalidArgumentError, match="XML parse error"):
        lib.check_model(str(tmp_path))
    d, model, nucs = prepare(0, n_particles=2000, n_batches=2)
    txt = open(os.path.join(d, "model.xml")).read().replace('type="x-plane"', 'type="cone"', 1)
    (tmp_path / "model.xml").write_text(txt)
    os.symlink(os.path.join(d, "data"), str(tmp_path / "data"))
    with pytest.raises(lib.InvalidArgumentError, match="cone"):
        lib.check_model(str(tmp_path))


def test_separate_xml_files(lib, tmp_path):
    """settings.xml / materials.xml / geometry.xml / tallies.xml instead of model.xml."""
    import xml.etree.ElementTree as ET
    d, model, nucs = prepare(0, n_particles=2000, n_batches=2)
    root = ET.parse(os.path.join(d, "model.xml")).getroot()
    for child in root:
        ET.ElementTree(child).write(str(tmp_path / (child.tag + ".xml")))
    os.symlink(os.path.join(d, "data"), str(tmp_path / "data"))
    s1, _, _ = lib.check_model(d)
    s2, _, _ = lib.check_model(str(tmp_path))
    assert s1 == s2


def test_init_without_gpu_reports_error(lib):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    d, model, nucs = prepare(0, n_particles=2000, n_batches=2)
    with chdir(d):
        with pytest.raises(lib.OpenMCError, match="GPU"):
            lib.init(output=False)


def _bindings():
    funcs, globs = {}, {}
    for line in open(os.path.join(ROOT, "tests", "golden", "openmc_lib_bindings.txt")):
        kind, name, extra = line.split()
        if kind == "func":
            funcs[name] = int(extra)
        else:
            globs[name] = extra
    return funcs, globs


def _header_decls():
    """name -> parameter count, and extern globals name -> C type, of capi.h."""
    header = open(os.path.join(ROOT, "include", "openmc_amd", "capi.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    funcs = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_0-9]+\s*\**\s+\**(\w+)\(([^;{]*?)\);", header, re.M | re.S):
        params = " ".join(m.group(2).split())
        funcs[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    globs = {m.group(2): m.group(1) for m in
             re.finditer(r"^extern\s+([A-Za-z_0-9]+\s*\**)\s*(\w+)(?:\[\d+\])?;", header, re.M)}
    return funcs, globs


def test_library_exports_every_name_openmc_lib_binds(lib):
    """Every function and global that the reference's openmc/lib/*.py looks up
    when it is imported (names scanned from its text into
    tests/golden/openmc_lib_bindings.txt) is exported, so the unchanged
    openmc.lib can load this library."""
    funcs, globs = _bindings()
    assert len(funcs) >= 190 and len(globs) >= 30
    missing = [n for n in list(funcs) + list(globs) if not hasattr(lib._dll, n)]
    assert not missing, missing


def test_header_signatures_match_reference_argtypes(lib):
    """The parameter count of each declaration in capi.h equals the length of
    the argtypes list the reference binds for it, and each global has the C
    type of the ctype the reference reads it through."""
    funcs, globs = _bindings()
    decl, gdecl = _header_decls()
    missing = sorted(n for n in funcs if n not in decl)
    assert not missing, missing
    wrong = {n: (decl[n], k) for n, k in funcs.items() if k >= 0 and decl[n] != k}
    assert not wrong, wrong
    ctype_of = {"c_bool": "bool", "c_int": "int", "c_int32": "int32_t", "c_int64": "int64_t",
                "c_double": "double", "c_char_p": "char*", "(c_char*256)": "char"}
    for name, ctype in globs.items():
        assert name in gdecl, name
        assert gdecl[name].replace(" ", "") == ctype_of[ctype], (name, gdecl[name], ctype)


def test_model_getters_before_init_return_reference_errors(lib):
    """Without a loaded model the getters fail with the reference's codes."""
    from ctypes import CDLL, byref, c_bool, c_double, c_int, c_int32
    d = CDLL(lib._dll._name)  # a second handle: no errcheck on its functions
    d.nuclides_size.restype = d.tally_filters_size.restype = d.n_materials.restype = __import__("ctypes").c_size_t
    i = c_int32()
    assert d.openmc_material_get_id(0, byref(i)) == -3          # OPENMC_E_OUT_OF_BOUNDS
    assert d.openmc_cell_get_id(0, byref(i)) == -3
    assert d.openmc_get_cell_index(1, byref(i)) == -7           # OPENMC_E_INVALID_ID
    assert d.openmc_get_mesh_index(1, byref(i)) == -7
    assert d.openmc_filter_get_id(0, byref(i)) == -3
    assert d.openmc_legendre_filter_get_order(0, byref(c_int())) == -3
    assert d.openmc_get_keff((c_double * 2)()) == -1            # OPENMC_E_UNASSIGNED
    from ctypes import c_char
    assert b"eigenvalue" in (c_char * 256).in_dll(d, "openmc_err_msg").value
    assert d.openmc_calculate_volumes() == -1
    assert d.cells_size() == 0 and d.n_meshes() == 0 and d.n_materials() == 0
    assert d.nuclides_size() == 0 and d.tally_filters_size() == 0
    n = c_int()
    d.openmc_get_n_batches(byref(n), c_bool(False))
    assert n.value == 0
    assert lib.settings.weight_windows_on is False
    assert lib.settings.cmfd_run is False and lib.settings.restart_run is False
    assert lib.settings.max_lost_particles == 10


def test_zernike_moments_closed_forms(lib):
    """calc_zn / calc_zn_rad against the textbook radial polynomials
    R_0^0 = 1, R_1^1 = r, R_2^0 = 2r^2 - 1, R_2^2 = r^2, R_3^1 = 3r^3 - 2r,
    R_3^3 = r^3, R_4^0 = 6r^4 - 6r^2 + 1, R_4^2 = 4r^4 - 3r^2, R_4^4 = r^4,
    R_6^0 = 20r^6 - 30r^4 + 12r^2 - 1, in the order (0,0), (1,-1), (1,1),
    (2,-2), (2,0), (2,2), ... with sin(|m| phi) for m < 0, cos(m phi) for m > 0."""
    for r, phi in [(0.0, 0.0), (0.3, 0.7), (0.81, -2.2), (1.0, 3.0)]:
        R = {(0, 0): 1.0, (1, 1): r, (2, 0): 2 * r**2 - 1, (2, 2): r**2, (3, 1): 3 * r**3 - 2 * r,
             (3, 3): r**3, (4, 0): 6 * r**4 - 6 * r**2 + 1, (4, 2): 4 * r**4 - 3 * r**2, (4, 4): r**4}
        want = []
        for n in range(5):
            for m in range(-n, n + 1, 2):
                rad = R[(n, abs(m))]
                want.append(rad * (np.sin(-m * phi) if m < 0 else np.cos(m * phi) if m > 0 else 1.0))
        np.testing.assert_allclose(lib.calc_zn(4, r, phi), want, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(lib.calc_zn_rad(6, r),
                                   [1.0, 2 * r**2 - 1, 6 * r**4 - 6 * r**2 + 1, 20 * r**6 - 30 * r**4 + 12 * r**2 - 1],
                                   rtol=1e-12, atol=1e-14)


def test_xml_lone_ampersand_is_a_parse_error(lib, tmp_path):
    """An '&' with no closing ';' is reported, not looped on."""
    d, model, nucs = prepare(0, n_particles=2000, n_batches=2)
    txt = open(os.path.join(d, "model.xml")).read().replace("<material ", '<material name="R&D" ', 1)
    (tmp_path / "model.xml").write_text(txt)
    os.symlink(os.path.join(d, "data"), str(tmp_path / "data"))
    with pytest.raises(lib.InvalidArgumentError, match="closing ';'"):
        lib.check_model(str(tmp_path))
