"""Synthetic continuous-energy nuclear data for the transport path.

The benchmark problems describe their nuclides procedurally: a log-uniform
energy grid from 1e-5 eV to 20 MeV, extra points clustered around resonances
with Breit-Wigner peaks, elastic scattering (MT2) and radiative capture (MT102)
with a 1/v component, isotropic scattering in the centre-of-mass frame and one
temperature (294 K).  This module draws such a nuclide from a seed and writes it
in a small binary layout that libopenmc's loader and the oracle read.

On-disk layout (little endian, ``*.nuc``)::

    char[8]  magic "OMCNUC1\\0"
    int32    len(name), name bytes
    float64  awr (atomic weight ratio)
    int32    n_temperatures, float64 kT[n] (eV)
    int32    n, float64 energy[n], float64 elastic[n], float64 capture[n]

The reaction content mirrors what ``IncidentNeutron.export_to_hdf5`` writes for
such a nuclide (openmc/data/neutron.py:354-430): MT2 and MT102 on the full
grid (threshold index 0), no angular data (isotropic in CM), and the derived
total/absorption columns are built by the reader exactly as
``Nuclide::create_derived`` does (src/nuclide.cpp:359-500).
"""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

K_BOLTZMANN = 8.617333262e-5  # eV/K, include/openmc/constants.h:102
E_THERMAL = 0.0253


class SyntheticNuclide:
    def __init__(self, name, awr, kT, energy, elastic, capture):
        self.name = name
        self.awr = float(awr)
        self.kTs = [float(kT)]
        self.energy = np.ascontiguousarray(energy, dtype=np.float64)
        self.elastic = np.ascontiguousarray(elastic, dtype=np.float64)
        self.capture = np.ascontiguousarray(capture, dtype=np.float64)

    @property
    def n_points(self):
        return self.energy.size

    def export(self, path):
        with open(path, "wb") as f:
            f.write(b"OMCNUC1\0")
            nb = self.name.encode()
            f.write(struct.pack("<i", len(nb)))
            f.write(nb)
            f.write(struct.pack("<d", self.awr))
            f.write(struct.pack("<i", len(self.kTs)))
            f.write(struct.pack("<%dd" % len(self.kTs), *self.kTs))
            f.write(struct.pack("<i", self.n_points))
            f.write(self.energy.astype("<f8").tobytes())
            f.write(self.elastic.astype("<f8").tobytes())
            f.write(self.capture.astype("<f8").tobytes())

    def export_hdf5(self, path):
        """Write the nuclide in the reference's HDF5 layout
        (openmc.data.IncidentNeutron.export_to_hdf5): elastic scattering (MT 2)
        and capture (MT 102) at its temperature, through libopenmc.so."""
        from . import lib
        lib.write_nuclide_hdf5(path, self.name, self.awr, self.kTs[0], int(round(self.kTs[0] / K_BOLTZMANN)),
                               self.energy, self.elastic, self.capture)

    @classmethod
    def load(cls, path):
        with open(path, "rb") as f:
            buf = f.read()
        if buf[:7] != b"OMCNUC1":
            raise ValueError("not a nuclide file: %s" % path)
        off = 8
        (nlen,) = struct.unpack_from("<i", buf, off)
        off += 4
        name = buf[off:off + nlen].decode()
        off += nlen
        (awr,) = struct.unpack_from("<d", buf, off)
        off += 8
        (nt,) = struct.unpack_from("<i", buf, off)
        off += 4
        kts = struct.unpack_from("<%dd" % nt, buf, off)
        off += 8 * nt
        (n,) = struct.unpack_from("<i", buf, off)
        off += 4
        arr = np.frombuffer(buf, dtype="<f8", count=3 * n, offset=off)
        return cls(name, awr, kts[0], arr[:n].copy(), arr[n:2 * n].copy(), arr[2 * n:].copy())


def synthetic_nuclide(name, seed, n_grid=10000, n_res=30, pts_per_res=40,
                      e_min=1.0e-5, e_max=2.0e7, temperature=294.0):
    """Draw one synthetic nuclide from ``seed``.

    * grid: ``n_grid`` log-uniform points over [e_min, e_max] plus
      ``pts_per_res`` points clustered (Lorentzian quantiles) around each of
      ``n_res`` resonances; sorted, exact duplicates removed.
    * resonances: E_r log-uniform in [1 eV, 1e5 eV], width Gamma log-uniform in
      [0.01, 10] eV, Breit-Wigner (Lorentzian) shape.
    * elastic: potential scattering s_p ~ U[1, 50] b, plus a 1/v term
      s_p * 0.05 * sqrt(0.0253/E), plus resonance peaks with heights ~ U[1, 50] b.
    * capture: 1/v with s_c(0.0253 eV) ~ U[0.1, 10] b, plus resonance peaks
      with heights log-uniform in [1, 1000] b.
    * awr ~ U[2, 240]; kT = k_B * 294 K.
    """
    rng = np.random.default_rng(seed)
    awr = rng.uniform(2.0, 240.0)
    e_r = np.exp(rng.uniform(np.log(1.0), np.log(1.0e5), n_res))
    gam = np.exp(rng.uniform(np.log(0.01), np.log(10.0), n_res))
    h_el = rng.uniform(1.0, 50.0, n_res)
    h_c = np.exp(rng.uniform(np.log(1.0), np.log(1000.0), n_res))
    s_pot = rng.uniform(1.0, 50.0)
    s_c0 = rng.uniform(0.1, 10.0)

    base = np.geomspace(e_min, e_max, n_grid)
    q = (np.arange(pts_per_res) + 0.5) / pts_per_res
    offs = np.tan(np.pi * (q - 0.5) * 0.96)  # Lorentzian quantiles within +-~16 half-widths
    clus = (e_r[:, None] + 0.5 * gam[:, None] * offs[None, :]).ravel()
    clus = clus[(clus > e_min) & (clus < e_max)]
    energy = np.unique(np.concatenate([base, clus]))

    lor = np.zeros((n_res, energy.size))
    for i in range(n_res):
        hw = 0.5 * gam[i]
        lor[i] = hw * hw / ((energy - e_r[i]) ** 2 + hw * hw)
    inv_v = np.sqrt(E_THERMAL / energy)
    elastic = s_pot + s_pot * 0.05 * inv_v + h_el @ lor
    capture = s_c0 * inv_v + h_c @ lor
    kT = K_BOLTZMANN * temperature
    return SyntheticNuclide(name, awr, kT, energy, elastic, capture)


class DataLibrary:
    """``cross_sections.xml`` writer (src/cross_sections.cpp:45-89 reads it)."""

    def __init__(self):
        self.libraries = []

    def register_file(self, path, materials):
        self.libraries.append({"path": path, "materials": materials, "type": "neutron"})

    def export_to_xml(self, path):
        root = ET.Element("cross_sections")
        for lib in self.libraries:
            el = ET.SubElement(root, "library")
            el.set("materials", lib["materials"])
            el.set("path", lib["path"])
            el.set("type", lib["type"])
        ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def write_library(nuclides, directory, fmt="nuc"):
    """Export nuclides into ``directory`` and write its cross_sections.xml;
    fmt "nuc" (this package's binary files) or "h5" (the reference's HDF5)."""
    os.makedirs(directory, exist_ok=True)
    lib = DataLibrary()
    for nuc in nuclides:
        fn = nuc.name + "." + fmt
        if fmt == "h5":
            nuc.export_hdf5(os.path.join(directory, fn))
        else:
            nuc.export(os.path.join(directory, fn))
        lib.register_file(fn, nuc.name)
    xml_path = os.path.join(directory, "cross_sections.xml")
    lib.export_to_xml(xml_path)
    return xml_path
