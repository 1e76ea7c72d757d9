"""Nuclear data in the reference's HDF5 layout (openmc/data/neutron.py:374-413,
openmc/data/reaction.py:910-936 write it; src/nuclide.cpp:52-300 and
src/reaction.cpp:25-60 read it).

The writer and reader are this library's; the layout they share is the
reference's.  Host-side checks need no GPU; the GPU check runs one problem on
.nuc and on .h5 data and compares the histories.
"""
import os

import numpy as np
import pytest

from helpers import ROOT  # noqa: F401  (puts the repository on sys.path)


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


def test_hdf5_round_trip_derives_total_and_absorption(hdf5):
    """TOTAL = 0 + MT2 + MT102 (file order) and ABSORPTION = 0 + MT102
    (Nuclide::create_derived, src/nuclide.cpp:359-425); elastic is MT 2."""
    from openmc_amd import data, lib
    nuc = data.synthetic_nuclide("H5A", 4242, n_grid=3000)
    path = os.path.join(str(hdf5), "H5A.h5")
    nuc.export_hdf5(path)
    got = lib.read_nuclide(path, "H5A")
    assert np.array_equal(got["energy"], nuc.energy)
    assert np.array_equal(got["elastic"], nuc.elastic)
    assert np.array_equal(got["total"], (0.0 + nuc.elastic) + nuc.capture)
    assert np.array_equal(got["absorption"], 0.0 + nuc.capture)
    assert got["awr"] == nuc.awr and got["kT"] == nuc.kTs[0]


def test_hdf5_temperature_outside_tolerance_is_an_error(hdf5):
    """Nearest temperature within the tolerance, else the reference's error
    (src/nuclide.cpp:133-166)."""
    from openmc_amd import data, lib
    nuc = data.synthetic_nuclide("H5B", 4243, n_grid=500)
    path = os.path.join(str(hdf5), "H5B.h5")
    nuc.export_hdf5(path)
    lib.read_nuclide(path, "H5B", temperature=300.0)
    with pytest.raises(lib.OpenMCError, match="at or near"):
        lib.read_nuclide(path, "H5B", temperature=600.0)


def test_hdf5_model_reads_like_nuc(hdf5):
    """A model on .h5 data initializes exactly as on .nuc data: energy grid,
    normalized densities and every log-grid bucket."""
    from openmc_amd import configs, lib
    model, nucs = configs.config0(n_particles=100, n_batches=1, n_grid=2000)
    a = os.path.join(str(hdf5), "nuc")
    b = os.path.join(str(hdf5), "h5")
    configs.write_problem(model, nucs, a)
    configs.write_problem(model, nucs, b, data_format="h5")
    n_gi = 4 * 8001
    sa, da, ga = lib.check_model(a, n_gi)
    sb, db, gb = lib.check_model(b, n_gi)
    assert sa == sb
    assert np.array_equal(da, db)
    assert np.array_equal(ga, gb)


@pytest.mark.gpu
def test_hdf5_histories_equal_nuc(hdf5):
    from helpers import gpu_run
    from openmc_amd import configs
    model, nucs = configs.config0(n_particles=2000, n_batches=2, n_grid=3000)
    a = os.path.join(str(hdf5), "nuc")
    b = os.path.join(str(hdf5), "h5")
    configs.write_problem(model, nucs, a)
    configs.write_problem(model, nucs, b, data_format="h5")
    ga = gpu_run(a, options={"trace_hash": 1})
    gb = gpu_run(b, options={"trace_hash": 1})
    assert np.array_equal(ga["hashes"], gb["hashes"])
    assert np.array_equal(ga["results"], gb["results"]) or np.allclose(ga["results"], gb["results"], rtol=1e-12,
                                                                         atol=0)
