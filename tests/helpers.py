"""Shared helpers of the test suite: problem preparation and GPU runs."""
import contextlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from openmc_amd import configs  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from oracle.problem import write_oracle_problem  # noqa: E402


def prepare(config, **kw):
    """Problem directory with model.xml (product) and oracle_problem.txt (oracle)."""
    d, model, nucs = configs.problem_dir(config, **kw)
    # rewritten every time: it is small (nuclide data are referenced by path)
    write_oracle_problem(model, nucs, d)
    return d, model, nucs


def oracle_run(d, event_based=False, threads=1, n_batches=0, glibc_math=False):
    o = Oracle(os.path.join(d, "oracle_problem.txt"), glibc_math=glibc_math)
    o.run(event_based=event_based, threads=threads, n_batches=n_batches)
    return o


@contextlib.contextmanager
def chdir(d):
    old = os.getcwd()
    os.chdir(d)
    try:
        yield
    finally:
        os.chdir(old)


def gpu_run(d, n_batches=None, options=None, learned=None):
    """Run the problem in directory d on the GPU through openmc_amd.lib.

    Returns dict(results, hashes, stats, n_real, leakage).
    """
    from openmc_amd import lib
    with chdir(d):
        lib.init(output=False)
        try:
            for k, v in (options or {}).items():
                lib.set_option(k, v)
            if learned is not None:
                lib.learned_build(*learned)
                lib.set_option("learned_index", 1)
            lib.simulation_init()
            nb = 0
            while True:
                status = lib.next_batch()
                nb += 1
                if status != 0 or (n_batches is not None and nb >= n_batches):
                    break
            lib.simulation_finalize()
            tid = list(lib.tallies.keys())[0]
            t = lib.tallies[tid]
            out = {"results": t.results.copy(), "hashes": lib.particle_hashes(), "stats": lib.stats(),
                   "n_real": t.num_realizations, "global": lib.global_tallies()}
        finally:
            lib.finalize()
    return out


def assert_tallies_close(gpu, ora, rtol=1e-12):
    """Per-bin relative agreement of SUM (reduction order is the only difference)."""
    g = gpu[:, 0, 1]
    o = ora[:, 0, 1]
    assert g.shape == o.shape
    nz = (o != 0) | (g != 0)
    rel = np.abs(g[nz] - o[nz]) / np.maximum(np.abs(o[nz]), np.abs(g[nz]))
    assert rel.size == 0 or rel.max() <= rtol, "max relative tally difference %g" % rel.max()
    assert np.array_equal(o == 0, g == 0), "bins scored on one side only"
