"""Writes the golden fixtures of this directory from the data the reference's
own tests hold (run once in a container that has the reference tree):

* bessel_j0_mathematica.json -- the 201 Mathematica values of
  J0(i^{3/2} x), x = 0, 0.1, ..., 20 (Code/tests/util/BesselTests.cc:27-221),
  checked there with Catch2's default Approx (relative 100 * FLT_EPSILON).
* four_cube.gmy -- Code/tests/resources/four_cube.gmy, the geometry file that
  Code/tests/geometry/GeometryReaderTests.cc:41-76 compares with the
  programmatic FourCube domain.
* four_cube.xml, config.xml, config-velocity-iolet.xml,
  config_new_velocity_inlets.xml -- the settings files of
  Code/tests/resources that Code/tests/configuration/SimConfigTests.cc and
  Code/tests/lb/iolets/InOutLetTests.cc:60-200 read (their expected values
  are restated in tests/test_config.py).
* mrt_basis_diag.json -- BASIS_TIMES_BASIS_TRANSPOSED expected by
  Code/tests/lb/KernelTests.cc:358-370.
* xtr_headers.json -- the main and field header bytes that
  Code/tests/extraction/LocalPropertyOutputTests.cc:131-190 expects for the
  dummy data source of Code/tests/extraction/DummyDataSource.h:22-40
  (voxel 0.3e-3 m, origin (0.034, 0.001, 0.074), 64 sites; a float Pressure
  field with one offset, 8 mmHg, and a float Velocity field).
"""
import json
import os
import re
import shutil

REF = "/root/reference/Code/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = open(os.path.join(REF, "util", "BesselTests.cc")).read()
    src = src[src.index("mathematica[] = {"):]
    src = src[:src.index("};")]
    vals = re.findall(r"Complex\(\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\)", src)
    vals = [(float(a), float(b)) for a, b in vals]
    assert len(vals) == 201, len(vals)
    json.dump({"x": [round(0.1 * i, 10) for i in range(201)],
               "re": [v[0] for v in vals], "im": [v[1] for v in vals],
               "source": "Code/tests/util/BesselTests.cc:27-221",
               "relative_tolerance": 100 * 1.1920928955078125e-07},
              open(os.path.join(HERE, "bessel_j0_mathematica.json"), "w"), indent=1)
    shutil.copyfile(os.path.join(REF, "resources", "four_cube.gmy"),
                    os.path.join(HERE, "four_cube.gmy"))
    for name in ("four_cube.xml", "config.xml", "config-velocity-iolet.xml",
                 "config_new_velocity_inlets.xml"):
        shutil.copyfile(os.path.join(REF, "resources", name), os.path.join(HERE, name))
    kt = open(os.path.join(REF, "lb", "KernelTests.cc")).read()
    d15 = re.search(r"std::array<double, 11>\{([^}]*)\}", kt).group(1)
    d19 = re.search(r"std::array<double, 15>\{([^}]*)\}", kt).group(1)
    json.dump({"15": [float(v) for v in d15.split(",")],
               "19": [float(v) for v in d19.split(",")],
               "source": "Code/tests/lb/KernelTests.cc:358-370"},
              open(os.path.join(HERE, "mrt_basis_diag.json"), "w"), indent=1)

    lp = open(os.path.join(REF, "extraction", "LocalPropertyOutputTests.cc")).read()
    json.dump({"main_header": _c_bytes(lp, "expectedMainHeader[] =").hex(),
               "field_header": _c_bytes(lp, "expectedFieldHeader[] =").hex(),
               "voxel_size": 0.3e-3, "origin": [0.034, 0.001, 0.074], "sites": 64,
               "reference_pressure_mmHg": 8.0,
               "source": "Code/tests/extraction/LocalPropertyOutputTests.cc:131-190"},
              open(os.path.join(HERE, "xtr_headers.json"), "w"), indent=1)


def _c_bytes(src: str, marker: str) -> bytes:
    """Bytes of the concatenated C string literals after `marker`."""
    body = src[src.index(marker) + len(marker):]
    body = body[:body.index(";")]
    out = bytearray()
    for lit in re.findall(r'"((?:[^"\\]|\\.)*)"', body):
        i = 0
        while i < len(lit):
            if lit[i] == "\\" and lit[i + 1] == "x":
                j = i + 2
                while j < len(lit) and lit[j] in "0123456789abcdefABCDEF":
                    j += 1
                out.append(int(lit[i + 2:j], 16))
                i = j
            else:
                out.append(ord(lit[i]))
                i += 1
    return bytes(out)


if __name__ == "__main__":
    main()
