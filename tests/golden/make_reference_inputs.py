"""Copy the model inputs of the reference's regression tests that fall on the
fixed-source transport path into tests/golden/reference_inputs/ (data files,
committed; the reference's Python API wrote them):

  fixed_source_model.xml          tests/regression_tests/fixed_source/inputs_true.dat
  weightwindows_model.xml         tests/regression_tests/weightwindows/inputs_true.dat
  reflective_plane_{geometry,materials}.xml   tests/regression_tests/reflective_plane/
  lattice_{geometry,materials}.xml            tests/regression_tests/lattice/

(the last two cases are eigenvalue runs: only their geometry and materials
are used, under a fixed-source settings.xml the tests write).

Run:  python tests/golden/make_reference_inputs.py [/path/to/reference]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "fixed_source_model.xml": "fixed_source/inputs_true.dat",
    "weightwindows_model.xml": "weightwindows/inputs_true.dat",
    "reflective_plane_geometry.xml": "reflective_plane/geometry.xml",
    "reflective_plane_materials.xml": "reflective_plane/materials.xml",
    "lattice_geometry.xml": "lattice/geometry.xml",
    "lattice_materials.xml": "lattice/materials.xml",
}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(HERE, "reference_inputs")
    os.makedirs(out, exist_ok=True)
    for name, src in FILES.items():
        shutil.copyfile(os.path.join(ref, "tests", "regression_tests", src), os.path.join(out, name))
        print(name)


if __name__ == "__main__":
    main()
