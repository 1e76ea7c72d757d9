"""Write the golden fixtures of tests/golden/ (run once; outputs are committed).

* G1.nuc: a 2000-point synthetic nuclide whose grid carries three exactly
  duplicated energies (the reference's duplicate-point bump,
  src/nuclide.cpp:740-742) -- fixed data, independent of numpy's libm.
* lookups_G1.npz: energies (random, every grid point, midpoints, bucket
  edges, out of range) with the oracle's i_grid / f / total / absorption /
  elastic.  The oracle is itself pinned by brute force in test_oracle.py;
  the fixture freezes its answers so the GPU path and future changes are
  checked against fixed vectors.
* rng.npz: prn streams for seeds 1, 2^63 - 1 and the first particle seeds of
  batch 1 (init_seed(id, stream) for ids 1..4, streams 0..3), and
  future_seed(n, 1) for n in a table -- computed by an independent Python
  restatement of PCG-RXS-M-XS and Brown's skip-ahead.

Run:  python tests/golden/make_golden.py
"""
import os
import sys
from ctypes import c_double, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

MASK = (1 << 64) - 1
MULT, ADD = 6364136223846793005, 1442695040888963407


def prn(seed):
    seed = (MULT * seed + ADD) & MASK
    word = (((seed >> ((seed >> 59) + 5)) ^ seed) * 12605985483714917081) & MASK
    return seed, float((word >> 43) ^ word) * 2.0 ** -64


def future_seed(n, seed):
    g, c, g_new, c_new = MULT, ADD, 1, 0
    while n > 0:
        if n & 1:
            g_new = (g_new * g) & MASK
            c_new = (c_new * g + c) & MASK
        c = (c * (g + 1)) & MASK
        g = (g * g) & MASK
        n >>= 1
    return (g_new * seed + c_new) & MASK


def make_nuclide():
    from openmc_amd import data
    nuc = data.synthetic_nuclide("G1", 424242, n_grid=2000, n_res=5, pts_per_res=20)
    E = list(nuc.energy)
    el = list(nuc.elastic)
    cap = list(nuc.capture)
    for k in (100, 777, 1500):  # duplicate three points exactly
        E.insert(k, E[k])
        el.insert(k, el[k] * 1.5)
        cap.insert(k, cap[k] * 0.5)
    return data.SyntheticNuclide("G1", nuc.awr, nuc.kTs[0], np.array(E), np.array(el), np.array(cap))


def main():
    from openmc_amd import model as om, stats
    from openmc_amd.configs import write_problem
    from oracle.oracle import Oracle
    from oracle.problem import write_oracle_problem

    nuc = make_nuclide()
    nuc.export(os.path.join(HERE, "G1.nuc"))

    # a one-nuclide problem to load the nuclide into the oracle
    om.reset_auto_ids()
    mat = om.Material()
    mat.add_nuclide("G1", 0.05)
    mat.set_density("sum")
    sph = om.Sphere(r=10.0, boundary_type="vacuum")
    geom = om.Geometry(om.Universe(cells=[om.Cell(fill=mat, region=-sph)]))
    s = om.Settings()
    s.particles = 10
    s.batches = 1
    s.source = om.IndependentSource(space=stats.Point(), energy=stats.Discrete([1e6], [1.0]))
    mesh = om.RegularMesh()
    mesh.dimension = (2, 2, 2)
    mesh.lower_left = (-10.0,) * 3
    mesh.upper_right = (10.0,) * 3
    t = om.Tally()
    t.filters = [om.MeshFilter(mesh)]
    model = om.Model(geom, om.Materials([mat]), s, om.Tallies([t]))
    import tempfile
    d = tempfile.mkdtemp()
    write_problem(model, [nuc], d)
    write_oracle_problem(model, [nuc], d)
    o = Oracle(os.path.join(d, "oracle_problem.txt"))
    e_min, e_max, spacing = o.energy_grid()
    rng = np.random.default_rng(99)
    E = nuc.energy
    energies = np.concatenate([np.exp(rng.uniform(np.log(e_min), np.log(e_max), 4000)), E,
                               0.5 * (E[:-1] + E[1:]), e_min * np.exp(np.arange(0, 8000, 7) * spacing),
                               [E[0] * 0.25, E[-1] * 2.0, np.nextafter(E[-1], 0)]])
    ig, xs = o.lookup(0, energies)
    np.savez_compressed(os.path.join(HERE, "lookups_G1.npz"), energies=energies, i_grid=ig, f=xs[:, 0],
                        total=xs[:, 1], absorption=xs[:, 2], elastic=xs[:, 3],
                        grid_index=o.grid_index(0), energy_grid=np.array([e_min, e_max, spacing]))

    streams = {}
    for s0 in (1, (1 << 63) - 1):
        seed, vals = s0, []
        for _ in range(16):
            seed, v = prn(seed)
            vals.append(v)
        streams["prn_seed_%d" % s0] = np.array(vals)
    init = np.zeros((4, 4), dtype=np.uint64)
    for pid in range(1, 5):
        for st in range(4):
            init[pid - 1, st] = future_seed(pid * 152917, 1 + st)
    ns = np.array([0, 1, 2, 5, 152917, 10**9, 2**40 + 3], dtype=np.uint64)
    fs = np.array([future_seed(int(n), 1) for n in ns], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "rng.npz"), init_seeds=init, future_n=ns, future_seed_1=fs,
                        **streams)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
