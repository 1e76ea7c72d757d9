"""Statepoint files in the reference's layout (openmc_statepoint_write,
src/state_point.cpp:36-360; the reader follows openmc/statepoint.py's paths).

Host side (no GPU): the library's HDF5 reader reads back what its writer
stores (attributes, datasets, fixed-length strings), and the settings
elements round-trip.  GPU side: statepoints written at the <state_point>
batches hold the library's tally results bit for bit, agree with the oracle
within 1e-12, list the model's meshes and filters, and carry the source bank
when asked for it; each problem of a problem set gets its own.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from helpers import ROOT, assert_tallies_close, chdir, oracle_run  # noqa: F401


def _have_hdf5(tmp):
    from openmc_amd import data, lib
    nuc = data.SyntheticNuclide("P1", 10.0, 0.0253, [1e-5, 1.0, 2e7], [1.0, 2.0, 3.0], [0.5, 0.25, 0.0])
    try:
        nuc.export_hdf5(os.path.join(str(tmp), "probe.h5"))
    except lib.OpenMCError as e:
        if "libhdf5" in str(e):
            return False
        raise
    return True


@pytest.fixture
def hdf5(tmp_path):
    if not _have_hdf5(tmp_path):
        pytest.skip("libhdf5 is not available on this machine")
    return tmp_path


def test_h5_read_attributes_datasets_and_strings(hdf5):
    from openmc_amd import data, lib
    from openmc_amd.data import K_BOLTZMANN
    nuc = data.synthetic_nuclide("SP1", 777, n_grid=400)
    path = os.path.join(str(hdf5), "SP1.h5")
    nuc.export_hdf5(path)
    assert lib.h5_read(path, "", "filetype", kind="str") == ["data_neutron"]
    assert list(lib.h5_read(path, "", "version", kind="i8")) == [3, 0]
    T = "%dK" % int(round(nuc.kTs[0] / K_BOLTZMANN))
    assert np.array_equal(lib.h5_read(path, "SP1/energy/" + T), nuc.energy)
    assert lib.h5_read(path, "SP1/reactions/reaction_002", "label", kind="str") == ["(n,elastic)"]
    assert lib.h5_read(path, "SP1/reactions/reaction_102", "mt", kind="i8")[0] == 102
    assert lib.h5_count(path, "SP1/energy/" + T) == nuc.energy.size
    with pytest.raises(lib.OpenMCError, match="dataset"):
        lib.h5_read(path, "SP1/no_such_dataset")
    with pytest.raises(lib.OpenMCError, match="not a string"):
        lib.h5_read(path, "SP1/energy/" + T, kind="str")


def test_settings_elements_for_statepoints():
    """<state_point><batches> and <output><path> as openmc.Settings writes
    them (openmc/settings.py:1462-1483)."""
    from openmc_amd import model as om
    s = om.Settings()
    s.particles, s.batches = 10, 3
    s.statepoint = {"batches": [2, 3]}
    s.output = {"path": "out", "tallies": False}
    el = s.to_xml_element()
    assert el.find("state_point/batches").text == "2 3"
    assert el.find("output/path").text == "out"
    assert el.find("output/tallies").text == "false"
    assert ET.tostring(el).count(b"state_point") == 2


SMALL1 = dict(n_particles=1200, n_batches=3, grid_lo=2000, grid_hi=5000, n_nuclides=9, lattice_n=4,
              mesh=(6, 5, 4), half_height=20.0, energy_bins=[1e-5, 1.0, 1e4, 2e7])


@pytest.mark.gpu
def test_statepoints_hold_the_results(hdf5):
    from oracle.problem import write_oracle_problem
    from openmc_amd import configs, lib
    from openmc_amd.statepoint import StatePoint
    model, nucs = configs.config1(**SMALL1)
    model.settings.statepoint = {"batches": [2, 3]}
    model.settings.output = {"path": "out"}
    d = str(hdf5 / "sp")
    configs.write_problem(model, nucs, d)
    write_oracle_problem(model, nucs, d)
    os.makedirs(os.path.join(d, "out"))
    o = oracle_run(d)
    with chdir(d):
        lib.init(output=False)
        try:
            lib.simulation_init()
            lib.next_batch()
            assert not lib.is_statepoint_batch()
            assert not os.path.exists("out/statepoint.1.h5")
            lib.next_batch()
            assert lib.is_statepoint_batch()
            sp2 = StatePoint("out/statepoint.2.h5")
            lib.next_batch()
            tid = list(lib.tallies.keys())[0]
            res = lib.tallies[tid].results.copy()
            gt = lib.global_tallies()
            lib.statepoint_write("with_source.h5", write_source=True)
            lib.simulation_finalize()
        finally:
            lib.finalize()
        sp = StatePoint("out/statepoint.3.h5")
        spw = StatePoint("with_source.h5")

    assert sp2.current_batch == 2 and sp2.n_realizations == 2
    assert sp.filetype == "statepoint" and sp.version == (18, 1)
    assert sp.run_mode == "fixed source" and sp.energy_mode == "continuous-energy"
    assert sp.n_particles == SMALL1["n_particles"] and sp.n_batches == 3 and sp.current_batch == 3
    assert sp.seed == 1 and sp.stride == 152917
    assert sp.n_realizations == 3 and sp.tallies_present and not sp.source_present
    assert sp.source_sites() == 0

    # meshes and filters as the model defines them
    (mesh_id, mesh), = sp.meshes.items()
    assert mesh.type == "regular"
    assert list(mesh.dimension) == list(SMALL1["mesh"])
    t = sp.tallies[tid]
    assert t.estimator == "tracklength" and t.scores == ["flux"] and t.nuclides == ["total"]
    assert t.num_realizations == 3
    ftypes = [sp.filters[f]["type"] for f in t.filter_ids]
    assert ftypes == ["mesh", "energy"]
    assert int(sp.filters[t.filter_ids[0]]["bins"][0]) == mesh_id
    assert np.array_equal(sp.filters[t.filter_ids[1]]["bins"], SMALL1["energy_bins"])
    assert sp.filters[t.filter_ids[0]]["n_bins"] == int(np.prod(SMALL1["mesh"]))

    # the library's results bit for bit, the oracle's within the tally tolerance
    assert np.array_equal(t.sum[:, 0], res[:, 0, 1])
    assert np.array_equal(t.sum_sq[:, 0], res[:, 0, 2])
    stacked = np.stack([np.zeros_like(t.sum[:, 0]), t.sum[:, 0], t.sum_sq[:, 0]], axis=-1)[:, None, :]
    assert_tallies_close(stacked, o.results())
    assert sp.global_tallies["leakage"]["sum"] == gt[3, 1]
    assert sp.global_tallies["leakage"]["sum_sq"] == gt[3, 2]
    n = 3
    mean = res[:, 0, 1] / n
    assert np.array_equal(t.mean[:, 0], mean)

    # the statepoint at batch 2 has two realizations' sums, a different file
    assert sp2.tallies[tid].num_realizations == 2
    assert not np.array_equal(sp2.tallies[tid].sum, t.sum)

    # explicit write with the source bank: one (unwritten) site per particle
    assert spw.source_present and spw.source_sites() == SMALL1["n_particles"]
    assert np.array_equal(spw.tallies[tid].sum, t.sum)


@pytest.mark.gpu
def test_problem_set_statepoints(hdf5):
    from openmc_amd import configs, lib
    from openmc_amd.statepoint import StatePoint
    d, dirs, problems, nucs = configs.problem_set_dir(n_problems=3, n_batches=2, particles=(300, 900),
                                                       lattice=(2, 3), nuclides=(6, 9), mesh=(3, 4),
                                                       grid_lo=2000, grid_hi=3000)
    for x in dirs:
        p = os.path.join(x, "statepoint.2.h5")
        if os.path.exists(p):
            os.remove(p)
    lib.init_problems(dirs, output=False)
    try:
        lib.run()
        want = []
        for k in range(lib.n_problems()):
            t = lib.problem_tally(k)
            want.append((t.id, t.results.copy(), lib.problem_global_tallies(k)))
            lib.problem_statepoint_write(k, str(hdf5 / ("p%d.h5" % k)))
    finally:
        lib.finalize()
    for k, x in enumerate(dirs):
        tid, res, gt = want[k]
        for path in (os.path.join(x, "statepoint.2.h5"), str(hdf5 / ("p%d.h5" % k))):
            sp = StatePoint(path)
            assert sp.n_particles == int(problems[k][0].settings.particles)
            assert list(sp.tallies) == [tid]
            assert np.array_equal(sp.tallies[tid].sum[:, 0], res[:, 0, 1])
            assert sp.global_tallies["leakage"]["sum"] == gt[3, 1]
