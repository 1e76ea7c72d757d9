"""Model-building subset of the ``openmc`` Python API for the transport path.

Objects mirror the reference's (openmc/material.py, openmc/surface.py,
openmc/cell.py, openmc/universe.py, openmc/lattice.py, openmc/settings.py,
openmc/source.py, openmc/mesh.py, openmc/filter.py, openmc/tallies.py,
openmc/model/model.py) for the features the path covers, and
``Model.export_to_model_xml`` writes the same ``model.xml`` layout
(docs/source/io_formats/{settings,geometry,materials,tallies}.rst) that
``openmc_init`` reads.
"""
import itertools
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import stats

_ids = {}


def _next_id(kind):
    c = _ids.setdefault(kind, itertools.count(1))
    return next(c)


def reset_auto_ids():
    _ids.clear()


def _fmt(values):
    return " ".join(repr(float(v)) for v in values)


# ----------------------------------------------------------------------------
# Materials
# ----------------------------------------------------------------------------

class Material:
    def __init__(self, material_id=None, name=""):
        self.id = material_id if material_id is not None else _next_id("material")
        self.name = name
        self.nuclides = []  # (name, percent, percent_type)
        self.density_units = "sum"
        self.density = None
        self.temperature = None

    def add_nuclide(self, nuclide, percent, percent_type="ao"):
        if percent_type not in ("ao", "wo"):
            raise ValueError("percent_type must be 'ao' or 'wo'")
        self.nuclides.append((nuclide, float(percent), percent_type))

    def set_density(self, units, density=None):
        if units not in ("sum", "atom/b-cm", "atom/cm3", "atom/cc", "g/cm3", "g/cc", "kg/m3"):
            raise ValueError("unsupported density units %r" % units)
        if units != "sum" and density is None:
            raise ValueError("density value required")
        self.density_units = units
        self.density = density

    def to_xml_element(self):
        el = ET.Element("material")
        el.set("id", str(self.id))
        if self.name:
            el.set("name", self.name)
        d = ET.SubElement(el, "density")
        d.set("units", self.density_units)
        if self.density is not None:
            d.set("value", repr(float(self.density)))
        for name, pct, ptype in self.nuclides:
            n = ET.SubElement(el, "nuclide")
            n.set("name", name)
            n.set(ptype, repr(pct))
        return el


class Materials(list):
    def __init__(self, materials=(), cross_sections=None):
        super().__init__(materials)
        self.cross_sections = cross_sections

    def to_xml_element(self):
        el = ET.Element("materials")
        if self.cross_sections is not None:
            cs = ET.SubElement(el, "cross_sections")
            cs.text = str(self.cross_sections)
        for m in self:
            el.append(m.to_xml_element())
        return el


# ----------------------------------------------------------------------------
# Surfaces and regions
# ----------------------------------------------------------------------------

class Halfspace:
    def __init__(self, surface, side):
        self.surface = surface
        self.side = side

    def __and__(self, other):
        return Intersection([self]) & other

    def tokens(self):
        return [self.surface.id if self.side == "+" else -self.surface.id]

    def surfaces(self):
        return [self.surface]


class Intersection:
    def __init__(self, nodes):
        self.nodes = list(nodes)

    def __and__(self, other):
        if isinstance(other, Intersection):
            return Intersection(self.nodes + other.nodes)
        return Intersection(self.nodes + [other])

    def tokens(self):
        out = []
        for n in self.nodes:
            out.extend(n.tokens())
        return out

    def surfaces(self):
        out = []
        for n in self.nodes:
            out.extend(n.surfaces())
        return out


class Surface:
    _type = None
    _coeff_names = ()

    def __init__(self, surface_id=None, boundary_type="transmission", **coeffs):
        self.id = surface_id if surface_id is not None else _next_id("surface")
        if boundary_type not in ("transmission", "vacuum", "reflective"):
            raise ValueError("unsupported boundary type %r" % boundary_type)
        self.boundary_type = boundary_type
        self.coeffs = [float(coeffs[k]) for k in self._coeff_names]

    def __neg__(self):
        return Halfspace(self, "-")

    def __pos__(self):
        return Halfspace(self, "+")

    def to_xml_element(self):
        el = ET.Element("surface")
        el.set("id", str(self.id))
        el.set("type", self._type)
        if self.boundary_type != "transmission":
            el.set("boundary", self.boundary_type)
        el.set("coeffs", _fmt(self.coeffs))
        return el


class XPlane(Surface):
    _type = "x-plane"
    _coeff_names = ("x0",)

    def __init__(self, x0=0.0, **kw):
        super().__init__(x0=x0, **kw)


class YPlane(Surface):
    _type = "y-plane"
    _coeff_names = ("y0",)

    def __init__(self, y0=0.0, **kw):
        super().__init__(y0=y0, **kw)


class ZPlane(Surface):
    _type = "z-plane"
    _coeff_names = ("z0",)

    def __init__(self, z0=0.0, **kw):
        super().__init__(z0=z0, **kw)


class Plane(Surface):
    _type = "plane"
    _coeff_names = ("a", "b", "c", "d")

    def __init__(self, a=1.0, b=0.0, c=0.0, d=0.0, **kw):
        super().__init__(a=a, b=b, c=c, d=d, **kw)


class ZCylinder(Surface):
    _type = "z-cylinder"
    _coeff_names = ("x0", "y0", "r")

    def __init__(self, x0=0.0, y0=0.0, r=1.0, **kw):
        super().__init__(x0=x0, y0=y0, r=r, **kw)


class Sphere(Surface):
    _type = "sphere"
    _coeff_names = ("x0", "y0", "z0", "r")

    def __init__(self, x0=0.0, y0=0.0, z0=0.0, r=1.0, **kw):
        super().__init__(x0=x0, y0=y0, z0=z0, r=r, **kw)


# ----------------------------------------------------------------------------
# Cells, universes, lattices, geometry
# ----------------------------------------------------------------------------

class Cell:
    def __init__(self, cell_id=None, fill=None, region=None, name=""):
        self.id = cell_id if cell_id is not None else _next_id("cell")
        self.fill = fill
        self.region = region
        self.name = name


class Universe:
    def __init__(self, universe_id=None, cells=(), name=""):
        self.id = universe_id if universe_id is not None else _next_id("universe")
        self.cells = list(cells)
        self.name = name

    def add_cell(self, cell):
        self.cells.append(cell)


class RectLattice:
    """Rectangular lattice; ``universes[row][col]`` with row 0 at the top (max y)."""

    def __init__(self, lattice_id=None, name=""):
        self.id = lattice_id if lattice_id is not None else _next_id("universe")
        self.name = name
        self.lower_left = None
        self.pitch = None
        self.outer = None
        self.universes = None

    @property
    def ndim(self):
        return len(self.pitch)

    @property
    def shape(self):
        u = np.asarray(self.universes, dtype=object)
        if self.ndim == 2:
            return (u.shape[1], u.shape[0])
        return (u.shape[2], u.shape[1], u.shape[0])

    def flat_universes(self):
        """Universes in XML order: for each z (bottom up) rows top to bottom."""
        u = np.asarray(self.universes, dtype=object)
        if self.ndim == 2:
            return list(u.ravel())
        return list(u.ravel())


class Geometry:
    def __init__(self, root):
        if isinstance(root, Universe):
            self.root_universe = root
        else:
            self.root_universe = Universe(cells=list(root))

    def _walk(self):
        universes, cells, surfaces, lattices = {}, {}, {}, {}

        def visit_universe(u):
            if u.id in universes:
                return
            universes[u.id] = u
            for c in u.cells:
                cells[c.id] = (c, u)
                if c.region is not None:
                    for s in c.region.surfaces():
                        surfaces[s.id] = s
                if isinstance(c.fill, Universe):
                    visit_universe(c.fill)
                elif isinstance(c.fill, RectLattice):
                    visit_lattice(c.fill)

        def visit_lattice(lat):
            if lat.id in lattices:
                return
            lattices[lat.id] = lat
            for u in lat.flat_universes():
                visit_universe(u)
            if lat.outer is not None:
                visit_universe(lat.outer)

        visit_universe(self.root_universe)
        return universes, cells, surfaces, lattices

    def to_xml_element(self):
        universes, cells, surfaces, lattices = self._walk()
        el = ET.Element("geometry")
        for cid in sorted(cells):
            c, u = cells[cid]
            ce = ET.SubElement(el, "cell")
            ce.set("id", str(c.id))
            if c.name:
                ce.set("name", c.name)
            if c.fill is None:
                ce.set("material", "void")
            elif isinstance(c.fill, Material):
                ce.set("material", str(c.fill.id))
            else:
                ce.set("fill", str(c.fill.id))
            if c.region is not None:
                ce.set("region", " ".join(str(t) for t in c.region.tokens()))
            ce.set("universe", str(u.id))
        for sid in sorted(surfaces):
            el.append(surfaces[sid].to_xml_element())
        for lid in sorted(lattices):
            lat = lattices[lid]
            le = ET.SubElement(el, "lattice")
            le.set("id", str(lat.id))
            ET.SubElement(le, "pitch").text = _fmt(lat.pitch)
            if lat.outer is not None:
                ET.SubElement(le, "outer").text = str(lat.outer.id)
            ET.SubElement(le, "dimension").text = " ".join(str(s) for s in lat.shape)
            ET.SubElement(le, "lower_left").text = _fmt(lat.lower_left)
            ET.SubElement(le, "universes").text = " ".join(str(u.id) for u in lat.flat_universes())
        return el


# ----------------------------------------------------------------------------
# Settings and sources
# ----------------------------------------------------------------------------

class IndependentSource:
    def __init__(self, space=None, angle=None, energy=None, strength=1.0, constraints=None):
        self.space = space
        self.angle = angle if angle is not None else stats.Isotropic()
        self.energy = energy if energy is not None else stats.Watt()
        self.strength = float(strength)
        self.constraints = constraints or {}

    def to_xml_element(self):
        el = ET.Element("source")
        el.set("type", "independent")
        el.set("strength", repr(self.strength))
        el.set("particle", "neutron")
        if self.space is not None:
            el.append(self.space.to_xml_element("space"))
        el.append(self.angle.to_xml_element("angle"))
        el.append(self.energy.to_xml_element("energy"))
        if self.constraints.get("domains"):
            ce = ET.SubElement(el, "constraints")
            doms = self.constraints["domains"]
            ET.SubElement(ce, "domain_type").text = "cell"
            ET.SubElement(ce, "domain_ids").text = " ".join(str(c.id) for c in doms)
        return el


class Settings:
    def __init__(self):
        self.run_mode = "fixed source"
        self.particles = None
        self.batches = None
        self.inactive = 0
        self.seed = None
        self.source = []
        self.event_based = False
        self.max_particles_in_flight = None
        self.log_grid_bins = None
        self.max_particle_events = None
        self.cross_sections = None
        self.weight_windows = []
        self.weight_windows_on = None
        self.weight_window_checkpoints = {}
        self.max_history_splits = None
        self.statepoint = {}  # {"batches": [...]} (openmc/settings.py:1477-1483)
        self.output = None    # {"path": ...} (openmc/settings.py:1462-1470)

    def to_xml_element(self):
        el = ET.Element("settings")
        ET.SubElement(el, "run_mode").text = self.run_mode
        ET.SubElement(el, "particles").text = str(int(self.particles))
        ET.SubElement(el, "batches").text = str(int(self.batches))
        if self.inactive:
            ET.SubElement(el, "inactive").text = str(int(self.inactive))
        srcs = self.source if isinstance(self.source, (list, tuple)) else [self.source]
        for s in srcs:
            el.append(s.to_xml_element())
        if self.seed is not None:
            ET.SubElement(el, "seed").text = str(int(self.seed))
        if self.event_based:
            ET.SubElement(el, "event_based").text = "true"
        if self.max_particles_in_flight is not None:
            ET.SubElement(el, "max_particles_in_flight").text = str(int(self.max_particles_in_flight))
        if self.log_grid_bins is not None:
            ET.SubElement(el, "log_grid_bins").text = str(int(self.log_grid_bins))
        if self.max_particle_events is not None:
            ET.SubElement(el, "max_particle_events").text = str(int(self.max_particle_events))
        # weight windows and their meshes (openmc/settings.py:1797-1862)
        written = set()
        for ww in self.weight_windows:
            el.append(ww.to_xml_element())
            if ww.mesh.id not in written:
                el.append(ww.mesh.to_xml_element())
                written.add(ww.mesh.id)
        if self.weight_windows_on is not None:
            ET.SubElement(el, "weight_windows_on").text = str(bool(self.weight_windows_on)).lower()
        if self.weight_window_checkpoints:
            ck = ET.SubElement(el, "weight_window_checkpoints")
            for key in ("collision", "surface"):
                if key in self.weight_window_checkpoints:
                    ET.SubElement(ck, key).text = str(bool(self.weight_window_checkpoints[key])).lower()
        if self.max_history_splits is not None:
            ET.SubElement(el, "max_history_splits").text = str(int(self.max_history_splits))
        if self.output is not None:
            oe = ET.SubElement(el, "output")
            for key, value in sorted(self.output.items()):
                ET.SubElement(oe, key).text = str(value).lower() if key in ("summary", "tallies") else value
        if self.statepoint:
            se = ET.SubElement(el, "state_point")
            if "batches" in self.statepoint:
                ET.SubElement(se, "batches").text = " ".join(str(x) for x in self.statepoint["batches"])
        return el


class WeightWindows:
    """Mesh-based weight windows (openmc/weight_windows.py:316-357 writes the
    same element).  lower/upper bounds have shape mesh.dimension + (n_e,) and
    are written in Fortran order, i.e. flat index e * n_mesh + mesh_bin."""

    def __init__(self, mesh, lower_ww_bounds, upper_ww_bounds=None, energy_bounds=None,
                 upper_bound_ratio=None, particle_type="neutron", survival_ratio=3.0,
                 max_lower_bound_ratio=None, max_split=10, weight_cutoff=1.0e-38, ww_id=None):
        self.id = ww_id if ww_id is not None else _next_id("weight_windows")
        self.mesh = mesh
        self.energy_bounds = None if energy_bounds is None else np.asarray(energy_bounds, dtype=float)
        n_e = 1 if self.energy_bounds is None else self.energy_bounds.size - 1
        shape = tuple(int(d) for d in mesh.dimension) + (n_e,)
        self.lower_ww_bounds = np.asarray(lower_ww_bounds, dtype=float).reshape(shape)
        if upper_ww_bounds is None:
            upper_ww_bounds = self.lower_ww_bounds * float(upper_bound_ratio)
        self.upper_ww_bounds = np.asarray(upper_ww_bounds, dtype=float).reshape(shape)
        self.particle_type = particle_type
        self.survival_ratio = survival_ratio
        self.max_lower_bound_ratio = max_lower_bound_ratio
        self.max_split = max_split
        self.weight_cutoff = weight_cutoff

    def flat_bounds(self):
        """(lower, upper) flattened as written: mesh bin fastest, energy slowest."""
        return self.lower_ww_bounds.ravel("F"), self.upper_ww_bounds.ravel("F")

    def to_xml_element(self):
        el = ET.Element("weight_windows")
        el.set("id", str(self.id))
        ET.SubElement(el, "mesh").text = str(self.mesh.id)
        ET.SubElement(el, "particle_type").text = self.particle_type
        if self.energy_bounds is not None:
            ET.SubElement(el, "energy_bounds").text = _fmt(self.energy_bounds)
        lo, hi = self.flat_bounds()
        ET.SubElement(el, "lower_ww_bounds").text = _fmt(lo)
        ET.SubElement(el, "upper_ww_bounds").text = _fmt(hi)
        ET.SubElement(el, "survival_ratio").text = repr(float(self.survival_ratio))
        if self.max_lower_bound_ratio is not None:
            ET.SubElement(el, "max_lower_bound_ratio").text = repr(float(self.max_lower_bound_ratio))
        ET.SubElement(el, "max_split").text = str(int(self.max_split))
        ET.SubElement(el, "weight_cutoff").text = repr(float(self.weight_cutoff))
        return el


# ----------------------------------------------------------------------------
# Tallies
# ----------------------------------------------------------------------------

class RegularMesh:
    def __init__(self, mesh_id=None, name=""):
        self.id = mesh_id if mesh_id is not None else _next_id("mesh")
        self.name = name
        self.dimension = None
        self.lower_left = None
        self.upper_right = None

    def to_xml_element(self):
        el = ET.Element("mesh")
        el.set("id", str(self.id))
        ET.SubElement(el, "dimension").text = " ".join(str(int(d)) for d in self.dimension)
        ET.SubElement(el, "lower_left").text = _fmt(self.lower_left)
        ET.SubElement(el, "upper_right").text = _fmt(self.upper_right)
        return el


class MeshFilter:
    def __init__(self, mesh, filter_id=None):
        self.id = filter_id if filter_id is not None else _next_id("filter")
        self.mesh = mesh

    @property
    def num_bins(self):
        return int(np.prod(self.mesh.dimension))

    def to_xml_element(self):
        el = ET.Element("filter")
        el.set("id", str(self.id))
        el.set("type", "mesh")
        ET.SubElement(el, "bins").text = str(self.mesh.id)
        return el


class EnergyFilter:
    def __init__(self, values, filter_id=None):
        self.id = filter_id if filter_id is not None else _next_id("filter")
        self.values = [float(v) for v in values]

    @property
    def num_bins(self):
        return len(self.values) - 1

    def to_xml_element(self):
        el = ET.Element("filter")
        el.set("id", str(self.id))
        el.set("type", "energy")
        ET.SubElement(el, "bins").text = _fmt(self.values)
        return el


class Tally:
    def __init__(self, tally_id=None, name=""):
        self.id = tally_id if tally_id is not None else _next_id("tally")
        self.name = name
        self.filters = []
        self.scores = ["flux"]
        self.estimator = "tracklength"

    def to_xml_element(self):
        el = ET.Element("tally")
        el.set("id", str(self.id))
        if self.name:
            el.set("name", self.name)
        if self.filters:
            ET.SubElement(el, "filters").text = " ".join(str(f.id) for f in self.filters)
        ET.SubElement(el, "scores").text = " ".join(self.scores)
        ET.SubElement(el, "estimator").text = self.estimator
        return el


class Tallies(list):
    def to_xml_element(self):
        el = ET.Element("tallies")
        meshes, filters = {}, {}
        for t in self:
            for f in t.filters:
                filters[f.id] = f
                if isinstance(f, MeshFilter):
                    meshes[f.mesh.id] = f.mesh
        for mid in sorted(meshes):
            el.append(meshes[mid].to_xml_element())
        for fid in sorted(filters):
            el.append(filters[fid].to_xml_element())
        for t in self:
            el.append(t.to_xml_element())
        return el


# ----------------------------------------------------------------------------
# Model
# ----------------------------------------------------------------------------

class Model:
    def __init__(self, geometry=None, materials=None, settings=None, tallies=None):
        self.geometry = geometry
        self.materials = materials if materials is not None else Materials()
        self.settings = settings if settings is not None else Settings()
        self.tallies = tallies if tallies is not None else Tallies()

    def export_to_model_xml(self, path="model.xml"):
        root = ET.Element("model")
        root.append(self.materials.to_xml_element())
        root.append(self.geometry.to_xml_element())
        root.append(self.settings.to_xml_element())
        if len(self.tallies):
            root.append(self.tallies.to_xml_element())
        if os.path.isdir(path):
            path = os.path.join(path, "model.xml")
        ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
        return path
