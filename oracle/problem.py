"""ORACLE -- test infrastructure only.

Flattens an ``openmc_amd`` model object into the oracle's problem file (a
whitespace-separated token stream read by oracle.cpp:load_problem).  The
product reads the same model from model.xml through its own C++ reader, so
the two sides share only the Python model object and the nuclide files.
"""
import os

import numpy as np

from openmc_amd import model as om
from openmc_amd import stats

SURF_CODES = {"x-plane": 1, "y-plane": 2, "z-plane": 3, "plane": 4, "z-cylinder": 5,
              "x-cylinder": 6, "y-cylinder": 7, "sphere": 8}
BC_CODES = {"transmission": 0, "vacuum": 1, "reflective": 2}
UNITS = {"sum": 0, "atom/b-cm": 1, "atom/cc": 2, "atom/cm3": 2, "g/cc": 3, "g/cm3": 3, "kg/m3": 4}


def _f(x):
    return repr(float(x))


def write_oracle_problem(model, nuclides, directory, data_dir="data"):
    """Write directory/oracle_problem.txt; nuclide files are data_dir/<name>.nuc."""
    out = ["oracle_problem", "1"]
    s = model.settings
    out += ["settings", str(int(s.particles)), str(int(s.batches)), str(int(s.inactive or 0)),
            str(int(s.seed if s.seed is not None else 1)),
            str(int(s.log_grid_bins if s.log_grid_bins else 8000)),
            str(int(s.max_particle_events if s.max_particle_events else 1000000)),
            str(int(s.max_particles_in_flight if s.max_particles_in_flight else 100000))]
    # nuclide index = first appearance over materials (Material constructor order)
    order = []
    for m in model.materials:
        for name, _, _ in m.nuclides:
            if name not in order:
                order.append(name)
    by_name = {n.name: n for n in nuclides}
    out += ["nuclides", str(len(order))]
    for name in order:
        if name not in by_name:
            raise KeyError("no data for nuclide %s" % name)
        out += [name, os.path.join(data_dir, name + ".nuc")]
    out += ["materials", str(len(model.materials))]
    for m in model.materials:
        out += [str(m.id), str(UNITS[m.density_units]), _f(m.density if m.density is not None else 0.0),
                str(len(m.nuclides))]
        for name, pct, ptype in m.nuclides:
            out += [str(order.index(name)), _f(pct if ptype == "ao" else -pct)]

    universes, cells, surfaces, lattices = model.geometry._walk()
    out += ["surfaces", str(len(surfaces))]
    for sid in sorted(surfaces):
        sf = surfaces[sid]
        c = list(sf.coeffs) + [0.0] * (4 - len(sf.coeffs))
        out += [str(sid), str(SURF_CODES[sf._type]), str(BC_CODES[sf.boundary_type])] + [_f(v) for v in c]
    # universes in order of first appearance over cells (sorted by id, as written to XML)
    uorder = []
    for cid in sorted(cells):
        uid = cells[cid][1].id
        if uid not in uorder:
            uorder.append(uid)
    out += ["universes", str(len(uorder))] + [str(u) for u in uorder]
    out += ["lattices", str(len(lattices))]
    for lid in sorted(lattices):
        lat = lattices[lid]
        shape = list(lat.shape) + [1] * (3 - len(lat.shape))
        ll = list(lat.lower_left) + [0.0] * (3 - len(lat.lower_left))
        pitch = list(lat.pitch) + [0.0] * (3 - len(lat.pitch))
        out += [str(lid)] + [str(v) for v in shape] + ["1" if len(lat.pitch) == 3 else "0"]
        out += [_f(v) for v in ll] + [_f(v) for v in pitch]
        out += [str(lat.outer.id if lat.outer is not None else -1)]
        out += [str(u.id) for u in lat.flat_universes()]
    out += ["cells", str(len(cells))]
    for cid in sorted(cells):
        c, u = cells[cid]
        if c.fill is None:
            ft, fid = 0, -1
        elif isinstance(c.fill, om.Material):
            ft, fid = 0, c.fill.id
        elif isinstance(c.fill, om.RectLattice):
            ft, fid = 2, c.fill.id
        else:
            ft, fid = 1, c.fill.id
        toks = c.region.tokens() if c.region is not None else []
        out += [str(cid), str(u.id), str(ft), str(fid), str(len(toks))] + [str(t) for t in toks]
    out += ["root", str(model.geometry.root_universe.id), str(_levels(model.geometry.root_universe))]

    src = s.source[0] if isinstance(s.source, (list, tuple)) else s.source
    sp = src.space
    if isinstance(sp, stats.Point):
        out += ["source", "0"] + [_f(v) for v in sp.xyz] + ["0", "0", "0"]
    else:
        out += ["source", "1"] + [_f(v) for v in sp.lower_left] + [_f(v) for v in sp.upper_right]
    if isinstance(src.angle, stats.Monodirectional):
        out += ["1"] + [_f(v) for v in src.angle.reference_uvw]
    else:
        out += ["0", "0", "0", "0"]
    en = src.energy
    if isinstance(en, stats.Discrete):
        out += ["0", str(len(en.x))]
        for x, p in zip(en.x, en.p):
            out += [_f(x), _f(p)]
    else:
        out += ["1", _f(en.a), _f(en.b)]
    doms = src.constraints.get("domains", []) if src.constraints else []
    out += [str(len(doms))] + [str(c.id) for c in doms]

    tally = model.tallies[0] if len(model.tallies) else None
    if tally is None:
        out += ["tally", "0"]
    else:
        meshes = [f.mesh for f in tally.filters if isinstance(f, om.MeshFilter)]
        ef = [f for f in tally.filters if isinstance(f, om.EnergyFilter)]
        if meshes:
            dim = list(meshes[0].dimension)
            nd = len(dim)
            shape = dim + [1] * (3 - nd)
            ll = list(meshes[0].lower_left) + [0.0] * (3 - nd)
            ur = list(meshes[0].upper_right) + [0.0] * (3 - nd)
        else:  # no mesh filter: n_dim 0
            nd, shape, ll, ur = 0, [1, 1, 1], [0.0] * 3, [0.0] * 3
        out += ["tally", "1", str(nd)] + [str(v) for v in shape] + [_f(v) for v in ll] + [_f(v) for v in ur]
        if ef:
            out += [str(len(ef[0].values))] + [_f(v) for v in ef[0].values]
        else:
            out += ["0"]
    out += _weight_windows_tokens(s)
    out += ["end"]
    path = os.path.join(directory, "oracle_problem.txt")
    with open(path, "w") as f:
        f.write(" ".join(out) + "\n")
    return path


def _weight_windows_tokens(s):
    """ww 0 | ww N <set>*N ckpt_collision ckpt_surface max_history_splits, a set being
    <mesh: nd shape[3] ll[3] ur[3]> <n_e+1 bounds> <n lower> <n upper> survival max_lb_ratio
    max_split cutoff"""
    wws = [w for w in (s.weight_windows or []) if getattr(w, "particle_type", "neutron") == "neutron"]
    on = s.weight_windows_on if s.weight_windows_on is not None else len(wws) > 0
    if not wws or not on:
        return ["ww", "0"]
    out = ["ww", str(len(wws))]
    for ww in wws:
        m = ww.mesh
        dim = list(m.dimension)
        nd = len(dim)
        shape = dim + [1] * (3 - nd)
        ll = list(m.lower_left) + [0.0] * (3 - nd)
        ur = list(m.upper_right) + [0.0] * (3 - nd)
        eb = list(ww.energy_bounds) if ww.energy_bounds is not None else [0.0, float("inf")]
        lo, hi = ww.flat_bounds()
        out += [str(nd)] + [str(v) for v in shape] + [_f(v) for v in ll] + [_f(v) for v in ur]
        out += [str(len(eb))] + [_f(v) for v in eb]
        out += [str(lo.size)] + [_f(v) for v in lo] + [_f(v) for v in hi]
        out += [_f(ww.survival_ratio),
                _f(ww.max_lower_bound_ratio if ww.max_lower_bound_ratio is not None else 1.0),
                str(int(ww.max_split)), _f(ww.weight_cutoff)]
    ck = s.weight_window_checkpoints or {}
    out += ["1" if ck.get("collision", True) else "0", "1" if ck.get("surface", False) else "0",
            str(int(s.max_history_splits if s.max_history_splits is not None else 10_000_000))]
    return out


def _levels(univ):
    below = 0
    for c in univ.cells:
        if isinstance(c.fill, om.Universe):
            below = max(below, _levels(c.fill))
        elif isinstance(c.fill, om.RectLattice):
            for u in set(c.fill.flat_universes()):
                below = max(below, _levels(u))
            if c.fill.outer is not None:
                below = max(below, _levels(c.fill.outer))
    return 1 + below
