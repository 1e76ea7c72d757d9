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
