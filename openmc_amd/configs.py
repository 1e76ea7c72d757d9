"""Builders of the benchmark problems of the transport path.

Each builder returns (model, nuclides) and can write itself into a directory
(model.xml + cross_sections.xml + nuclide files) with ``write_problem``.
Sizes default to the full benchmark definitions; tests shrink them through
the keyword arguments.  Nuclide data are drawn from fixed seeds
(``openmc_amd.data.synthetic_nuclide``), so a problem is reproducible from
its arguments.

Choices the benchmark text leaves open (recorded here and in the design notes):
  * atomic weight ratios ~ U[2, 240] from each nuclide's seed; T = 294 K;
  * config 0 uses the atom densities themselves ("sum" density units); configs
    1-2 draw atom fractions ~ U[0.005, 0.05] and scale each material to
    0.1 atom/b-cm so that 20 or 300 nuclides give a realistic total density;
  * config 1's source is a Watt spectrum sampled uniformly in the fuel
    cylinders: a box over the lattice with a cell-domain constraint;
  * config 2's source is a Watt spectrum sampled uniformly in the cube;
  * config 3's problems split their 10-60 nuclides into the three lattice
    materials in thirds (in the order drawn), each material scaled to
    0.1 atom/b-cm like config 1, 10 batches each.
"""
import hashlib
import json
import os
import tempfile

import numpy as np

from . import data, stats
from . import model as om

DATA_SEED_BASE = 1000


def _nuclide_set(prefix, count, seed, grid_lo, grid_hi, n_res=30, pts_per_res=40):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(grid_lo, grid_hi + 1, size=count)
    return [data.synthetic_nuclide("%s%03d" % (prefix, k), seed * 1000 + k, n_grid=int(sizes[k]),
                                   n_res=n_res, pts_per_res=pts_per_res) for k in range(count)]


def config0(n_particles=100000, n_batches=10, n_grid=10000, mesh=(10, 10, 10), event_based=True,
            energy_bins=None):
    """20 cm cube, one material of 4 nuclides, 1 MeV point source, 10^3 flux mesh
    (mesh None: a tally with no mesh filter; energy_bins: an energy filter)."""
    om.reset_auto_ids()
    nucs = [data.synthetic_nuclide("N%d" % k, DATA_SEED_BASE + k, n_grid=n_grid) for k in range(4)]
    rng = np.random.default_rng(DATA_SEED_BASE)
    mat = om.Material(name="mix")
    for n in nucs:
        mat.add_nuclide(n.name, float(rng.uniform(0.005, 0.05)))
    mat.set_density("sum")
    planes = [om.XPlane(-10.0, boundary_type="vacuum"), om.XPlane(10.0, boundary_type="vacuum"),
              om.YPlane(-10.0, boundary_type="vacuum"), om.YPlane(10.0, boundary_type="vacuum"),
              om.ZPlane(-10.0, boundary_type="vacuum"), om.ZPlane(10.0, boundary_type="vacuum")]
    region = +planes[0] & -planes[1] & +planes[2] & -planes[3] & +planes[4] & -planes[5]
    cell = om.Cell(fill=mat, region=region)
    geom = om.Geometry(om.Universe(cells=[cell]))
    s = om.Settings()
    s.particles = n_particles
    s.batches = n_batches
    s.event_based = event_based
    s.source = om.IndependentSource(space=stats.Point((0.0, 0.0, 0.0)), angle=stats.Isotropic(),
                                    energy=stats.Discrete([1.0e6], [1.0]))
    t = om.Tally(name="flux")
    if mesh is not None:
        m = om.RegularMesh()
        m.dimension = mesh
        m.lower_left = (-10.0, -10.0, -10.0)
        m.upper_right = (10.0, 10.0, 10.0)
        t.filters = [om.MeshFilter(m)]
    if energy_bins is not None:
        t.filters.append(om.EnergyFilter(energy_bins))
    t.scores = ["flux"]
    model = om.Model(geom, om.Materials([mat]), s, om.Tallies([t]))
    return model, nucs


def config1(n_particles=10_000_000, n_batches=20, grid_lo=20000, grid_hi=100000, n_nuclides=60,
            lattice_n=17, mesh=(100, 100, 100), event_based=True, half_height=50.0,
            energy_bins=None, tally=True):
    """17x17 pin lattice, three materials of 20 nuclides, Watt source in the fuel
    (tally False: no tally, for measuring what the mesh tally costs)."""
    om.reset_auto_ids()
    nucs = _nuclide_set("L", n_nuclides, DATA_SEED_BASE + 1, grid_lo, grid_hi)
    per = n_nuclides // 3
    rng = np.random.default_rng(DATA_SEED_BASE + 11)
    mats = []
    for k, name in enumerate(["fuel", "clad", "moderator"]):
        mt = om.Material(name=name)
        for n in nucs[k * per:(k + 1) * per]:
            mt.add_nuclide(n.name, float(rng.uniform(0.005, 0.05)))
        mt.set_density("atom/b-cm", 0.1)
        mats.append(mt)
    pitch = 1.26
    half = 0.5 * pitch * lattice_n
    cyl1 = om.ZCylinder(r=0.4)
    cyl2 = om.ZCylinder(r=0.47)
    fuel = om.Cell(fill=mats[0], region=-cyl1, name="fuel")
    clad = om.Cell(fill=mats[1], region=+cyl1 & -cyl2, name="clad")
    mod = om.Cell(fill=mats[2], region=+cyl2, name="moderator")
    pin = om.Universe(cells=[fuel, clad, mod])
    lat = om.RectLattice()
    lat.lower_left = (-half, -half)
    lat.pitch = (pitch, pitch)
    lat.universes = [[pin] * lattice_n for _ in range(lattice_n)]
    xm = om.XPlane(-half, boundary_type="reflective")
    xp = om.XPlane(half, boundary_type="reflective")
    ym = om.YPlane(-half, boundary_type="reflective")
    yp = om.YPlane(half, boundary_type="reflective")
    zm = om.ZPlane(-half_height, boundary_type="vacuum")
    zp = om.ZPlane(half_height, boundary_type="vacuum")
    root = om.Cell(fill=lat, region=+xm & -xp & +ym & -yp & +zm & -zp)
    geom = om.Geometry(om.Universe(cells=[root]))
    s = om.Settings()
    s.particles = n_particles
    s.batches = n_batches
    s.event_based = event_based
    s.source = om.IndependentSource(
        space=stats.Box((-half, -half, -half_height), (half, half, half_height)),
        angle=stats.Isotropic(), energy=stats.Watt(0.988e6, 2.249e-6),
        constraints={"domains": [fuel]})
    m = om.RegularMesh()
    m.dimension = mesh
    m.lower_left = (-half, -half, -half_height)
    m.upper_right = (half, half, half_height)
    t = om.Tally(name="flux")
    t.filters = [om.MeshFilter(m)]
    if energy_bins is not None:
        t.filters.append(om.EnergyFilter(energy_bins))
    model = om.Model(geom, om.Materials(mats), s, om.Tallies([t] if tally else []))
    return model, nucs


def config2(n_particles=10_000_000, n_batches=20, n_nuclides=300, n_grid=200000, mesh=(100, 100, 100),
            event_based=True, half=50.0):
    """One homogeneous material of 300 nuclides with 200k-point grids in a 1 m cube."""
    om.reset_auto_ids()
    nucs = [data.synthetic_nuclide("H%03d" % k, (DATA_SEED_BASE + 2) * 1000 + k, n_grid=n_grid)
            for k in range(n_nuclides)]
    rng = np.random.default_rng(DATA_SEED_BASE + 22)
    mat = om.Material(name="mix300")
    for n in nucs:
        mat.add_nuclide(n.name, float(rng.uniform(0.005, 0.05)))
    mat.set_density("atom/b-cm", 0.1)
    planes = [om.XPlane(-half, boundary_type="vacuum"), om.XPlane(half, boundary_type="vacuum"),
              om.YPlane(-half, boundary_type="vacuum"), om.YPlane(half, boundary_type="vacuum"),
              om.ZPlane(-half, boundary_type="vacuum"), om.ZPlane(half, boundary_type="vacuum")]
    region = +planes[0] & -planes[1] & +planes[2] & -planes[3] & +planes[4] & -planes[5]
    cell = om.Cell(fill=mat, region=region)
    geom = om.Geometry(om.Universe(cells=[cell]))
    s = om.Settings()
    s.particles = n_particles
    s.batches = n_batches
    s.event_based = event_based
    s.source = om.IndependentSource(space=stats.Box((-half,) * 3, (half,) * 3), angle=stats.Isotropic(),
                                    energy=stats.Watt(0.988e6, 2.249e-6))
    m = om.RegularMesh()
    m.dimension = mesh
    m.lower_left = (-half,) * 3
    m.upper_right = (half,) * 3
    t = om.Tally(name="flux")
    t.filters = [om.MeshFilter(m)]
    model = om.Model(geom, om.Materials([mat]), s, om.Tallies([t]))
    return model, nucs


def config4(n_particles=10_000_000, n_batches=16, n_grid=200000, n_layers=5, layer=20.0, per_layer=20,
            half_width=50.0, density=0.01, ww_mesh=(100, 50, 50), ww_length=8.0, upper_ratio=5.0,
            tally_mesh=(200, 200, 200), energy_groups=10, event_based=True, ww_split_sets=False):
    """Deep-penetration slab with mesh weight windows.

    Layers of `layer` cm along x (x = 0 and x = n_layers * layer vacuum,
    the lateral faces reflective), each a distinct mix of `per_layer`
    nuclides drawn (seeded) from config 2's 300, at `density` atom/b-cm.  A
    14 MeV monodirectional (+x) plane source sits 10 um inside the x = 0 face
    (a point exactly on the face has no cell for a direction along the face).
    Weight windows on a `ww_mesh` mesh over the slab: lower bound
    a0 * exp(-x / ww_length) at the bin centre, a0 ~ U[0.3, 0.6] from the seed,
    upper = upper_ratio * lower, one energy group; splitting and roulette at
    birth and collisions.  Tally: track-length flux on `tally_mesh` with
    `energy_groups` log-spaced groups over [1e-5 eV, 20 MeV].
    ww_split_sets: the same windows as two sets, the first valid (lower >= 0)
    only in the front half of the slab and -1 behind it, the second over the
    whole slab at 1.5x the bounds: a particle in the back half takes the
    second set's window (the first set's are invalid there).
    """
    om.reset_auto_ids()
    rng = np.random.default_rng(DATA_SEED_BASE + 44)
    pick = sorted(int(k) for k in rng.choice(300, size=n_layers * per_layer, replace=False))
    nucs = [data.synthetic_nuclide("H%03d" % k, (DATA_SEED_BASE + 2) * 1000 + k, n_grid=n_grid) for k in pick]
    mats = []
    for i in range(n_layers):
        mt = om.Material(name="layer%d" % i)
        for n in nucs[i * per_layer:(i + 1) * per_layer]:
            mt.add_nuclide(n.name, float(rng.uniform(0.005, 0.05)))
        mt.set_density("atom/b-cm", density)
        mats.append(mt)
    xs = [om.XPlane(i * layer, boundary_type="vacuum" if i in (0, n_layers) else "transmission")
          for i in range(n_layers + 1)]
    ym = om.YPlane(-half_width, boundary_type="reflective")
    yp = om.YPlane(half_width, boundary_type="reflective")
    zm = om.ZPlane(-half_width, boundary_type="reflective")
    zp = om.ZPlane(half_width, boundary_type="reflective")
    cells = [om.Cell(fill=mats[i], region=+xs[i] & -xs[i + 1] & +ym & -yp & +zm & -zp, name="layer%d" % i)
             for i in range(n_layers)]
    geom = om.Geometry(om.Universe(cells=cells))
    depth = n_layers * layer
    s = om.Settings()
    s.particles = n_particles
    s.batches = n_batches
    s.event_based = event_based
    s.source = om.IndependentSource(
        space=stats.Box((1.0e-3, -half_width, -half_width), (1.0e-3, half_width, half_width)),
        angle=stats.Monodirectional((1.0, 0.0, 0.0)), energy=stats.Discrete([14.0e6], [1.0]))
    wm = om.RegularMesh()
    wm.dimension = ww_mesh
    wm.lower_left = (0.0, -half_width, -half_width)
    wm.upper_right = (depth, half_width, half_width)
    a0 = float(rng.uniform(0.3, 0.6))
    xc = (np.arange(ww_mesh[0]) + 0.5) * depth / ww_mesh[0]
    lower = np.broadcast_to((a0 * np.exp(-xc / ww_length))[:, None, None, None],
                            tuple(ww_mesh) + (1,)).copy()
    if ww_split_sets:
        front = lower.copy()
        front[ww_mesh[0] // 2:] = -1.0
        s.weight_windows = [om.WeightWindows(wm, front, np.where(front < 0, -10.0, front * upper_ratio)),
                            om.WeightWindows(wm, 1.5 * lower, 1.5 * lower * upper_ratio)]
    else:
        s.weight_windows = [om.WeightWindows(wm, lower, lower * upper_ratio)]
    tm = om.RegularMesh()
    tm.dimension = tally_mesh
    tm.lower_left = (0.0, -half_width, -half_width)
    tm.upper_right = (depth, half_width, half_width)
    t = om.Tally(name="flux")
    t.filters = [om.MeshFilter(tm)]
    if energy_groups:
        t.filters.append(om.EnergyFilter(list(np.geomspace(1.0e-5, 2.0e7, energy_groups + 1))))
    model = om.Model(geom, om.Materials(mats), s, om.Tallies([t]))
    return model, nucs


def config3_problem(nucs, rng, lattice_n, n_nuc, n_particles, mesh_m, n_batches=10, half_height=50.0,
                    event_based=True):
    """One problem of the batched set: a lattice_n x lattice_n pin lattice like
    config 1, its three materials sharing n_nuc nuclides drawn from the
    config-1 set, a Watt source in the fuel and an m^3 flux mesh."""
    om.reset_auto_ids()
    pick = sorted(rng.choice(len(nucs), size=n_nuc, replace=False))
    chosen = [nucs[k] for k in pick]
    cut = [0, n_nuc // 3, 2 * (n_nuc // 3), n_nuc]
    mats = []
    for k, name in enumerate(["fuel", "clad", "moderator"]):
        mt = om.Material(name=name)
        for n in chosen[cut[k]:cut[k + 1]]:
            mt.add_nuclide(n.name, float(rng.uniform(0.005, 0.05)))
        mt.set_density("atom/b-cm", 0.1)
        mats.append(mt)
    pitch = 1.26
    half = 0.5 * pitch * lattice_n
    cyl1 = om.ZCylinder(r=0.4)
    cyl2 = om.ZCylinder(r=0.47)
    fuel = om.Cell(fill=mats[0], region=-cyl1, name="fuel")
    clad = om.Cell(fill=mats[1], region=+cyl1 & -cyl2, name="clad")
    mod = om.Cell(fill=mats[2], region=+cyl2, name="moderator")
    pin = om.Universe(cells=[fuel, clad, mod])
    lat = om.RectLattice()
    lat.lower_left = (-half, -half)
    lat.pitch = (pitch, pitch)
    lat.universes = [[pin] * lattice_n for _ in range(lattice_n)]
    xm = om.XPlane(-half, boundary_type="reflective")
    xp = om.XPlane(half, boundary_type="reflective")
    ym = om.YPlane(-half, boundary_type="reflective")
    yp = om.YPlane(half, boundary_type="reflective")
    zm = om.ZPlane(-half_height, boundary_type="vacuum")
    zp = om.ZPlane(half_height, boundary_type="vacuum")
    root = om.Cell(fill=lat, region=+xm & -xp & +ym & -yp & +zm & -zp)
    geom = om.Geometry(om.Universe(cells=[root]))
    s = om.Settings()
    s.particles = int(n_particles)
    s.batches = n_batches
    s.event_based = event_based
    s.source = om.IndependentSource(
        space=stats.Box((-half, -half, -half_height), (half, half, half_height)),
        angle=stats.Isotropic(), energy=stats.Watt(0.988e6, 2.249e-6),
        constraints={"domains": [fuel]})
    m = om.RegularMesh()
    m.dimension = (mesh_m,) * 3
    m.lower_left = (-half, -half, -half_height)
    m.upper_right = (half, half, half_height)
    t = om.Tally(name="flux")
    t.filters = [om.MeshFilter(m)]
    return om.Model(geom, om.Materials(mats), s, om.Tallies([t])), chosen


def config3(n_problems=64, n_batches=10, particles=(1e4, 1e6), lattice=(3, 17), nuclides=(10, 60),
            mesh=(10, 64), grid_lo=20000, grid_hi=100000, seed=DATA_SEED_BASE + 3):
    """Batched uneven problems: n_problems lattice problems drawn from a seed
    (lattice n x n with n uniform in `lattice`, 10-60 of the config-1
    nuclides, particles per batch log-uniform in `particles`, an m^3 mesh with
    m uniform in `mesh`).  Returns (list of (model, nuclides used), all nuclides)."""
    nucs = _nuclide_set("L", 60, DATA_SEED_BASE + 1, grid_lo, grid_hi)
    rng = np.random.default_rng(seed)
    problems = []
    for _ in range(n_problems):
        n = int(rng.integers(lattice[0], lattice[1] + 1))
        k = int(rng.integers(nuclides[0], nuclides[1] + 1))
        npart = int(round(float(np.exp(rng.uniform(np.log(particles[0]), np.log(particles[1]))))))
        m = int(rng.integers(mesh[0], mesh[1] + 1))
        problems.append(config3_problem(nucs, rng, n, k, npart, m, n_batches=n_batches))
    return problems, nucs


def write_problem_set(problems, nuclides, directory):
    """Write one shared nuclide library and one model directory per problem
    (directory/p000, p001, ...); returns the problem directories."""
    os.makedirs(directory, exist_ok=True)
    xs = data.write_library(nuclides, os.path.join(directory, "data"))
    dirs = []
    for k, (model, _) in enumerate(problems):
        d = os.path.join(directory, "p%03d" % k)
        os.makedirs(d, exist_ok=True)
        model.materials.cross_sections = os.path.relpath(xs, d)
        model.export_to_model_xml(os.path.join(d, "model.xml"))
        dirs.append(d)
    return dirs


def problem_set_dir(**kwargs):
    """Build (or reuse from the cache) a config-3 problem set.
    Returns (set directory, problem directories, problems, nuclides)."""
    d = os.path.join(cache_root(), problem_key(3, **kwargs))
    problems, nucs = config3(**kwargs)
    stamp = os.path.join(d, "complete")
    if not os.path.exists(stamp):
        dirs = write_problem_set(problems, nucs, d)
        with open(stamp, "w") as f:
            f.write(json.dumps({"config": 3, "kwargs": kwargs, "problems": len(dirs)}, default=str))
    dirs = [os.path.join(d, "p%03d" % k) for k in range(len(problems))]
    return d, dirs, problems, nucs


BUILDERS = {0: config0, 1: config1, 2: config2, 4: config4}


def cache_root():
    return os.environ.get("OPENMC_AMD_DATA", os.path.join(tempfile.gettempdir(), "openmc_amd_data"))


def write_problem(model, nuclides, directory, data_format="nuc"):
    """Write nuclide files (data_format "nuc" or "h5"), cross_sections.xml and
    model.xml into directory."""
    os.makedirs(directory, exist_ok=True)
    xs = data.write_library(nuclides, os.path.join(directory, "data"), fmt=data_format)
    model.materials.cross_sections = os.path.relpath(xs, directory)
    model.export_to_model_xml(os.path.join(directory, "model.xml"))
    return directory


def problem_key(config, **kwargs):
    return "cfg%d_%s" % (config, hashlib.sha1(
        json.dumps([config, sorted(kwargs.items())], default=str).encode()).hexdigest()[:16])


def problem_dir(config, **kwargs):
    """Build (or reuse from the cache) the directory of a configured problem.

    Returns (directory, model, nuclides); the model and nuclides are rebuilt
    from their seeds (cheap next to transport) so callers can inspect them.
    """
    d = os.path.join(cache_root(), problem_key(config, **kwargs))
    model, nucs = BUILDERS[config](**kwargs)
    stamp = os.path.join(d, "complete")
    if not os.path.exists(stamp):
        write_problem(model, nucs, d)
        with open(stamp, "w") as f:
            f.write(json.dumps({"config": config, "kwargs": kwargs, "nuclides": [n.name for n in nucs]},
                               default=str))
    return d, model, nucs
