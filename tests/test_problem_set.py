"""Problem sets: several independent models sharing one event loop.

Host side (no GPU): the concatenated model keeps each problem's indices,
particle ranges and tally offsets, holds each nuclide data file once and
rejects problems that disagree on run-global settings.  GPU side: every
problem of a set gives, history for history, the hashes of the problem run
alone and of the oracle, and its tally and leakage within 1e-12 relative.
"""
import os
import shutil

import numpy as np
import pytest

from helpers import assert_tallies_close, chdir, gpu_run, oracle_run

SMALL = dict(n_problems=4, n_batches=2, particles=(300, 1500), lattice=(2, 4), nuclides=(6, 12), mesh=(3, 6),
             grid_lo=2000, grid_hi=4000)


@pytest.fixture(scope="module")
def small_set():
    from openmc_amd import configs
    d, dirs, problems, nucs = configs.problem_set_dir(**SMALL)
    return d, dirs, problems, nucs


def test_problem_set_host_layout(small_set):
    from openmc_amd import lib
    d, dirs, problems, nucs = small_set
    head, parts = lib.check_problem_set(dirs)
    used = sorted({n.name for _, ns in problems for n in ns})
    assert head["n_nuclides"] == len(used)
    assert head["n_materials"] == 3 * len(problems)
    assert head["n_levels"] == 2
    npart = [m.settings.particles for m, _ in problems]
    assert head["n_particles"] == sum(npart)
    assert head["n_tallies"] == len(problems)
    first = np.concatenate([[0], np.cumsum(npart)[:-1]])
    cells = surfs = bins = 0
    for k, (p, (m, _)) in enumerate(zip(parts, problems)):
        one, _, _ = lib.check_model(dirs[k])
        assert p["first"] == first[k] and p["n_particles"] == npart[k]
        assert p["mat_begin"] == 3 * k
        assert p["cell_begin"] == cells and p["surf_begin"] == surfs
        assert p["tally"] == k and p["tally_off"] == bins
        cells += one["n_cells"]
        surfs += one["n_surfaces"]
        bins += one["n_tally_bins"]
        assert one["energy_min"] == head["energy_min"] and one["energy_max"] == head["energy_max"]
    assert head["n_cells"] == cells and head["n_surfaces"] == surfs


def test_problem_set_rejects_mismatched_settings(small_set, tmp_path):
    from openmc_amd import lib
    d, dirs, problems, nucs = small_set
    bad = os.path.join(str(tmp_path), "bad")
    shutil.copytree(dirs[1], bad)
    p = os.path.join(bad, "model.xml")
    text = open(p).read()
    assert "<batches>2</batches>" in text
    open(p, "w").write(text.replace("<batches>2</batches>", "<batches>3</batches>"))
    # keep the relative path to the shared library valid from the copy
    text = open(p).read().replace("../data/", os.path.join(d, "data") + "/")
    open(p, "w").write(text)
    with pytest.raises(lib.OpenMCError, match="must share batches"):
        lib.check_problem_set([dirs[0], bad])


def _set_run(dirs, options=None, learned=False):
    from openmc_amd import lib
    lib.init_problems(dirs, output=False)
    try:
        for k, v in (options or {}).items():
            lib.set_option(k, v)
        if learned:
            assert lib.learned_build(256, 100) <= 64
            lib.set_option("learned_index", 1)
        lib.run()
        h = lib.particle_hashes()
        out = []
        for k in range(lib.n_problems()):
            info = lib.problem_info(k)
            t = lib.problem_tally(k)
            out.append({"hashes": h[info["first"]:info["first"] + info["n_particles"]].copy(),
                        "results": t.results.copy(), "n_real": t.num_realizations,
                        "global": lib.problem_global_tallies(k)})
        stats = lib.stats()
    finally:
        lib.finalize()
    return out, stats


@pytest.mark.gpu
@pytest.mark.parametrize("options", [{}, {"max_in_flight": 1000}], ids=["all-in-flight", "chunked"])
def test_problem_set_matches_problems_alone(small_set, options):
    from oracle.problem import write_oracle_problem
    d, dirs, problems, nucs = small_set
    opts = dict(options, trace_hash=1)
    got, _ = _set_run(dirs, opts)
    for k, (model, used) in enumerate(problems):
        alone = gpu_run(dirs[k], options={"trace_hash": 1})
        n = model.settings.particles
        assert got[k]["hashes"].size == n
        assert np.array_equal(got[k]["hashes"], alone["hashes"]), "problem %d histories differ" % k
        assert_tallies_close(got[k]["results"], alone["results"])
        assert got[k]["n_real"] == alone["n_real"]
        leak, leak1 = got[k]["global"][3, 1:], np.asarray(alone["global"]).reshape(-1, 3)[3, 1:]
        assert np.allclose(leak, leak1, rtol=1e-12, atol=0)
        if k < 2:
            write_oracle_problem(model, used, dirs[k], data_dir=os.path.join("..", "data"))
            o = oracle_run(dirs[k])
            assert np.array_equal(got[k]["hashes"], o.hashes(n)), "problem %d differs from the oracle" % k
            assert_tallies_close(got[k]["results"], o.results())


@pytest.mark.gpu
def test_problem_set_64_problems_full_structure():
    """The 64-problem set of the batched workload as it is (lattices 3x3 to
    17x17, 10-60 of the 60 nuclides with 20k-100k point grids, 192 materials,
    meshes 10^3 to 64^3), with 30-120 particles a problem instead of 1e4-1e6:
    the same draws of every other property.  The calculate-xs sort then fills
    the largest key histogram (192 materials x binades x 2^bits buckets up to
    32768 keys).  Every problem's histories equal the oracle's and its tally
    agrees within 1e-12 per bin, on binary search and on the learned index."""
    from openmc_amd import configs, lib
    from oracle.problem import write_oracle_problem
    print("case: problem set of 64, 30-120 particles each", flush=True)
    d, dirs, problems, nucs = configs.problem_set_dir(n_problems=64, n_batches=1, particles=(30, 120))
    head, _ = lib.check_problem_set(dirs)
    assert head["n_materials"] == 192
    binades = int(np.float64(head["energy_max"]).view(np.uint64) >> np.uint64(52)) - int(
        np.float64(head["energy_min"]).view(np.uint64) >> np.uint64(52)) + 1
    want = []
    for k, (model, used) in enumerate(problems):
        write_oracle_problem(model, used, dirs[k], data_dir=os.path.join("..", "data"))
        o = oracle_run(dirs[k], threads=8)
        want.append((o.hashes(model.settings.particles), o.results()))
    for learned in (False, True):
        got, st = _set_run(dirs, {"trace_hash": 1}, learned=learned)
        bits = int(st["sort_bits"])
        assert 1 + 192 * (binades << bits) <= 32768 < 1 + 192 * (binades << (bits + 1))
        for k in range(len(dirs)):
            assert np.array_equal(got[k]["hashes"], want[k][0]), "problem %d differs (learned %s)" % (k, learned)
            assert_tallies_close(got[k]["results"], want[k][1])


@pytest.mark.gpu
def test_problem_set_learned_index_same_histories(small_set):
    """The learned index, built over a problem set's shared nuclide table,
    gives every problem the histories and tallies of binary search, with
    fewer search probes."""
    d, dirs, problems, nucs = small_set
    ref, st_ref = _set_run(dirs, {"trace_hash": 1})
    got, st = _set_run(dirs, {"trace_hash": 1}, learned=True)
    for k in range(len(dirs)):
        assert np.array_equal(got[k]["hashes"], ref[k]["hashes"]), "problem %d histories differ" % k
        assert_tallies_close(got[k]["results"], ref[k]["results"])
        assert got[k]["n_real"] == ref[k]["n_real"]
    assert st["lookups"] == st_ref["lookups"] and st["collisions"] == st_ref["collisions"]
    assert st["search_probes"] < st_ref["search_probes"]


@pytest.mark.gpu
def test_autotune_picks_from_space_and_keeps_histories(small_set):
    """The autotuner times real batches and leaves its choice set; whatever it
    picks, the histories are those of the default launch shape."""
    from openmc_amd import autotune, lib
    d, dirs, problems, nucs = small_set
    ref, _ = _set_run(dirs, {"trace_hash": 1})
    space = [("block", [256, 64]), ("sort_every", [1, 3]), ("cycles_per_check", [16, 4])]
    lib.init_problems(dirs, output=False)
    try:
        best, trials = autotune.tune(space=space)
        assert set(best) >= {"block", "sort_every", "cycles_per_check"}
        for name, values in space:
            assert best[name] in values
        assert len(trials) == 1 + sum(len(v) - 1 for _, v in space)
        assert all(t["seconds"] > 0 for t in trials)
        st = lib.stats()
        assert st["block"] == best["block"] and st["sort_every"] == best["sort_every"]
        lib.set_option("trace_hash", 1)
        lib.run()
        h = lib.particle_hashes()
        for k in range(lib.n_problems()):
            info = lib.problem_info(k)
            assert np.array_equal(h[info["first"]:info["first"] + info["n_particles"]], ref[k]["hashes"])
            t = lib.problem_tally(k)
            assert t.num_realizations == ref[k]["n_real"]
            assert_tallies_close(t.results, ref[k]["results"])
    finally:
        lib.finalize()
