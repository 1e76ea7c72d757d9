"""Source distributions (subset of ``openmc.stats``) used by the transport path.

Each class writes the XML element the reference's C++ reader expects
(docs/source/io_formats/settings.rst, ``<source>`` element; univariate
distributions in src/distribution.cpp, spatial in src/distribution_spatial.cpp,
angular in src/distribution_multi.cpp).
"""
import xml.etree.ElementTree as ET


def _fmt(values):
    return " ".join(repr(float(v)) for v in values)


class Point:
    """Point source location (src/distribution_spatial.cpp:385-388)."""

    def __init__(self, xyz=(0.0, 0.0, 0.0)):
        self.xyz = tuple(float(v) for v in xyz)

    def to_xml_element(self, tag="space"):
        el = ET.Element(tag)
        el.set("type", "point")
        el.set("parameters", _fmt(self.xyz))
        return el


class Box:
    """Uniform box (src/distribution_spatial.cpp:363-367)."""

    def __init__(self, lower_left, upper_right, only_fissionable=False):
        if only_fissionable:
            raise NotImplementedError("fission-only boxes are outside this path")
        self.lower_left = tuple(float(v) for v in lower_left)
        self.upper_right = tuple(float(v) for v in upper_right)

    def to_xml_element(self, tag="space"):
        el = ET.Element(tag)
        el.set("type", "box")
        el.set("parameters", _fmt(self.lower_left + self.upper_right))
        return el


class Isotropic:
    """Isotropic direction (src/distribution_multi.cpp:106-112)."""

    def to_xml_element(self, tag="angle"):
        el = ET.Element(tag)
        el.set("type", "isotropic")
        return el


class Monodirectional:
    """Fixed direction (src/distribution_multi.cpp:123-126)."""

    def __init__(self, reference_uvw=(1.0, 0.0, 0.0)):
        self.reference_uvw = tuple(float(v) for v in reference_uvw)

    def to_xml_element(self, tag="angle"):
        el = ET.Element(tag)
        el.set("type", "monodirectional")
        el.set("reference_uvw", _fmt(self.reference_uvw))
        return el


class Discrete:
    """Discrete distribution sampled by the alias method (src/distribution.cpp:87-101)."""

    def __init__(self, x, p):
        self.x = [float(v) for v in x]
        self.p = [float(v) for v in p]
        if len(self.x) != len(self.p):
            raise ValueError("x and p must have the same length")

    def to_xml_element(self, tag="energy"):
        el = ET.Element(tag)
        el.set("type", "discrete")
        el.set("parameters", _fmt(self.x + self.p))
        return el


class Watt:
    """Watt fission spectrum, a in eV and b in 1/eV (src/random_dist.cpp:34-39)."""

    def __init__(self, a=0.988e6, b=2.249e-6):
        self.a = float(a)
        self.b = float(b)

    def to_xml_element(self, tag="energy"):
        el = ET.Element(tag)
        el.set("type", "watt")
        el.set("parameters", _fmt([self.a, self.b]))
        return el
