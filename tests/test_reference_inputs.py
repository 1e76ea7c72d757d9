"""The model.xml reader on inputs the reference's own Python API wrote (its
regression-test inputs, copied into tests/golden/reference_inputs/ by
make_reference_inputs.py), with their nuclide names pointed at synthetic
data.  Read without a GPU (omc_load_model_host) and checked through the same
getters openmc.lib uses; features outside the fixed-source neutron path
(photons, S(a,b), particle filters) must be refused with a message, never
dropped.  Reference reader: src/initialize.cpp:360-470 and the from_xml
readers it calls.
"""
import os
import re
import shutil

import numpy as np
import pytest

from openmc_amd import data

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "reference_inputs")

MASS_NEUTRON = 1.00866491595
N_AVOGADRO = 0.602214076


@pytest.fixture(scope="module")
def lib():
    from openmc_amd import lib as _lib
    return _lib


def _nuclide_names(*texts):
    names = []
    for t in texts:
        for n in re.findall(r'<nuclide name="([A-Za-z0-9_]+)"', t):
            if n not in names:
                names.append(n)
    return names


def _library(tmp, names):
    """Synthetic nuclides under the reference's names, and cross_sections.xml."""
    nucs = [data.synthetic_nuclide(n, 9000 + k, n_grid=1500, n_res=3, pts_per_res=8) for k, n in enumerate(names)]
    return data.write_library(nucs, os.path.join(tmp, "data")), {n.name: n for n in nucs}


def _load(lib, d, xs):
    os.environ["OPENMC_CROSS_SECTIONS"] = xs
    try:
        lib.load_model_host(d)
    finally:
        del os.environ["OPENMC_CROSS_SECTIONS"]


FIXED_SETTINGS = """<?xml version='1.0' encoding='utf-8'?>
<settings>
  <run_mode>fixed source</run_mode>
  <particles>100</particles>
  <batches>2</batches>
  <source><space type="box"><parameters>{box}</parameters></space></source>
</settings>
"""


def test_fixed_source_model(lib, tmp_path):
    """fixed_source: a sphere of O16/U238 at 7.5 g/cc, a point source with
    the default Watt spectrum, a flux tally with no filter."""
    txt = open(os.path.join(FIX, "fixed_source_model.xml")).read()
    xs, nucs = _library(str(tmp_path), _nuclide_names(txt))
    shutil.copy(os.path.join(FIX, "fixed_source_model.xml"), tmp_path / "model.xml")
    _load(lib, str(tmp_path), xs)
    try:
        assert lib._dll.cells_size() == 1 and lib._dll.n_materials() == 1
        mats = lib.materials()
        m = mats[1]
        assert m.nuclides == ["O16", "U238"]
        # g/cc: the densities reproduce 7.5 g/cc with the nuclides' awr, in the
        # ratio of the atom fractions (Material::normalize_density)
        awr = {k: v.awr for k, v in nucs.items()}
        rho = sum(N * awr[n] * MASS_NEUTRON / N_AVOGADRO for n, N in zip(m.nuclides, m.densities))
        assert rho == pytest.approx(7.5, rel=1e-13)
        assert m.get_density("g/cm3") == pytest.approx(7.5, rel=1e-13)
        assert m.densities[0] / m.densities[1] == pytest.approx(1.0 / 0.0001, rel=1e-13)
        from ctypes import byref, c_bool, c_double
        dep = c_bool()
        lib._dll.openmc_material_get_depletable(0, byref(dep))
        assert dep.value is True
        t = c_double()
        lib._dll.openmc_material_get_temperature(0, byref(t))
        assert t.value == 294.0
        tal = lib.tallies[1]
        assert tal.filters == [] and tal.scores == ["flux"] and tal.estimator == "tracklength"
        assert lib.settings.particles == 100
        with pytest.raises(lib.OpenMCError, match="without a GPU"):
            lib.simulation_init()
        # an option out of range is an error code, not an exception through the C ABI
        with pytest.raises(lib.InvalidArgumentError, match="learned_err_bound"):
            lib.set_option("learned_err_bound", 65)
    finally:
        lib.finalize()


def test_weightwindows_model(lib, tmp_path):
    """weightwindows: as written it asks for photon transport and a particle
    filter (refused, named); without them the neutron windows, their mesh
    and the mesh x energy tally read back exactly, both window sets in order."""
    txt = open(os.path.join(FIX, "weightwindows_model.xml")).read()
    xs, nucs = _library(str(tmp_path), _nuclide_names(txt))
    (tmp_path / "model.xml").write_text(txt)
    with pytest.raises(lib.OpenMCError, match="photon transport"):
        _load(lib, str(tmp_path), xs)
    # neutrons only: drop photon transport and the particle filter
    t2 = txt.replace("<photon_transport>true</photon_transport>", "")
    t2 = re.sub(r'<filter id="3" type="particle">.*?</filter>', "", t2, flags=re.S)
    t2 = t2.replace("<filters>1 2 3</filters>", "<filters>1 2</filters>")
    (tmp_path / "model.xml").write_text(t2)
    # both neutron window sets are read, in order (a particle takes the first
    # set with a valid window at its position and energy)
    ww1 = re.search(r'(<weight_windows id="1">.*?</weight_windows>)', t2, flags=re.S).group(1)
    ww2 = re.search(r'(<weight_windows id="2">.*?</weight_windows>)', t2, flags=re.S).group(1)
    _load(lib, str(tmp_path), xs)
    try:
        assert lib.settings.weight_windows_on is True
        from ctypes import POINTER, byref, c_double, c_int, c_int32, c_size_t
        assert lib._dll.openmc_weight_windows_size() == 2
        for k, (wid, ww) in enumerate(((1, ww1), (2, ww2))):
            idx = c_int32()
            lib._dll.openmc_get_weight_windows_index(wid, byref(idx))
            assert idx.value == k
            lo, hi = POINTER(c_double)(), POINTER(c_double)()
            n = c_size_t()
            lib._dll.openmc_weight_windows_get_bounds(k, lo, hi, n)
            want_lo = np.array(re.search(r"<lower_ww_bounds>(.*?)</lower_ww_bounds>", ww, re.S).group(1).split(),
                               float)
            want_hi = np.array(re.search(r"<upper_ww_bounds>(.*?)</upper_ww_bounds>", ww, re.S).group(1).split(),
                               float)
            assert n.value == 5 * 6 * 7 * 2 == want_lo.size
            assert np.array_equal(np.ctypeslib.as_array(lo, (n.value,)), want_lo)
            assert np.array_equal(np.ctypeslib.as_array(hi, (n.value,)), want_hi)
        eb, ne = POINTER(c_double)(), c_size_t()
        lib._dll.openmc_weight_windows_get_energy_bounds(0, eb, ne)
        assert list(np.ctypeslib.as_array(eb, (ne.value,))) == [0.0, 0.5, 2.0e7]
        r = c_double()
        lib._dll.openmc_weight_windows_get_survival_ratio(0, byref(r))
        assert r.value == 3.0
        lib._dll.openmc_weight_windows_get_max_lower_bound_ratio(0, byref(r))
        assert r.value == 1.5
        lib._dll.openmc_weight_windows_get_weight_cutoff(0, byref(r))
        assert r.value == 1e-38
        ms = c_int()
        lib._dll.openmc_weight_windows_get_max_split(0, byref(ms))
        assert ms.value == 10
        mi = c_int32()
        lib._dll.openmc_weight_windows_get_mesh(0, byref(mi))
        wm = lib.RegularMesh(mi.value)
        assert wm.id == 2 and wm.dimension == (5, 6, 7)
        assert list(wm.lower_left) == [-240.0] * 3 and list(wm.upper_right) == [240.0] * 3
        tal = lib.tallies[1]
        f = tal.filters
        assert [type(x).__name__ for x in f] == ["MeshFilter", "EnergyFilter"]
        assert f[0].mesh.id == 1 and f[0].mesh.dimension == (5, 10, 15)
        assert f[1].bins.tolist() == [[0.0, 0.5], [0.5, 2.0e7]]
        # the 250 cm vacuum sphere around a 240 cm concrete sphere
        assert lib._dll.cells_size() == 2
        cells = lib.cells()
        assert cells[2].fill == ("material", -1)
    finally:
        lib.finalize()


def test_reflective_plane_geometry(lib, tmp_path):
    """reflective_plane: a reflective 10 cm box of U235 at 4.5 g/cc (its
    eigenvalue settings replaced by a fixed-source box source)."""
    geo = open(os.path.join(FIX, "reflective_plane_geometry.xml")).read()
    mat = open(os.path.join(FIX, "reflective_plane_materials.xml")).read()
    xs, nucs = _library(str(tmp_path), _nuclide_names(mat))
    (tmp_path / "geometry.xml").write_text(geo)
    (tmp_path / "materials.xml").write_text(mat)
    (tmp_path / "settings.xml").write_text(FIXED_SETTINGS.format(box="-4 -4 -4 4 4 4"))
    os.environ["OPENMC_CROSS_SECTIONS"] = xs
    try:
        summ, dens, _ = lib.check_model(str(tmp_path))
    finally:
        del os.environ["OPENMC_CROSS_SECTIONS"]
    assert summ["n_cells"] == 1 and summ["n_surfaces"] == 6 and summ["n_levels"] == 1
    _load(lib, str(tmp_path), xs)
    try:
        m = lib.materials()[1]
        assert m.nuclides == ["U235"] and m.get_density("g/cm3") == pytest.approx(4.5, rel=1e-13)
        assert lib.cells()[1].fill[0] == "material"
    finally:
        lib.finalize()


def test_lattice_geometry(lib, tmp_path):
    """lattice (MIX-COMP-THERM-002): 29x29 lattice of two pin universes inside
    a reflective quarter cylinder.  With its S(a,b) moderator the model is
    refused; without it the geometry reads back: 36 cells, 22 surfaces, two
    levels, each pin cell once per lattice element of its universe."""
    geo = open(os.path.join(FIX, "lattice_geometry.xml")).read()
    mat = open(os.path.join(FIX, "lattice_materials.xml")).read()
    xs, nucs = _library(str(tmp_path), _nuclide_names(mat))
    (tmp_path / "geometry.xml").write_text(geo)
    (tmp_path / "materials.xml").write_text(mat)
    (tmp_path / "settings.xml").write_text(FIXED_SETTINGS.format(box="0 0 3.556 21.336 21.336 94.996"))
    with pytest.raises(lib.OpenMCError, match="S\\(a,b\\)"):
        _load(lib, str(tmp_path), xs)
    (tmp_path / "materials.xml").write_text(re.sub(r"<sab[^>]*/>", "", mat))
    os.environ["OPENMC_CROSS_SECTIONS"] = xs
    try:
        summ, dens, _ = lib.check_model(str(tmp_path))
    finally:
        del os.environ["OPENMC_CROSS_SECTIONS"]
    assert summ["n_cells"] == 36 and summ["n_surfaces"] == 22 and summ["n_lattices"] == 1
    assert summ["n_levels"] == 2 and summ["n_materials"] == 6
    lat = re.search(r"<universes>(.*?)</universes>", geo, re.S).group(1).split()
    n1, n2 = lat.count("1"), lat.count("2")
    assert n1 + n2 == 29 * 29
    _load(lib, str(tmp_path), xs)
    try:
        cells = lib.cells()
        assert cells[3].fill[0] == "lattice" and cells[3].num_instances == 1
        for cid in range(21, 39):
            assert cells[cid].num_instances == n1, cid
        for cid in range(41, 54):
            assert cells[cid].num_instances == n2, cid
        mats = lib.materials()
        assert mats[6].get_density() == pytest.approx(3.2174e-02, rel=1e-14)
    finally:
        lib.finalize()
