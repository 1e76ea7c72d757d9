"""MI355X-native event-based continuous-energy fixed-source transport.

The model-building subset of the ``openmc`` Python API lives at the top level
(``openmc_amd.Material``, ``openmc_amd.ZCylinder``, ``openmc_amd.Model``, ...);
``openmc_amd.lib`` is the ``openmc.lib``-compatible binding to libopenmc.so.
"""
from . import stats, data
from .model import (Material, Materials, XPlane, YPlane, ZPlane, Plane, ZCylinder, Sphere, Cell,
                    Universe, RectLattice, Geometry, IndependentSource, Settings, RegularMesh,
                    MeshFilter, EnergyFilter, Tally, Tallies, Model, reset_auto_ids)

__version__ = "0.1.0"
