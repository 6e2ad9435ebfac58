"""Writes the fixtures of the reference's tests/08_steady_flow_sources into
tests/golden/ref:
  * test08.msh (input/test08.msh, as it is);
  * test08_pressure.txt: pressure_elements of
    ref_output/flow_vtk.con/flow/flow-000000.vtu, the values as printed."""
import os
import re
import shutil

SRC = "/root/reference/tests/08_steady_flow_sources"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")


def main():
    shutil.copyfile(os.path.join(SRC, "input", "test08.msh"), os.path.join(OUT, "test08.msh"))
    txt = open(os.path.join(SRC, "ref_output", "flow_vtk.con", "flow", "flow-000000.vtu")).read()
    m = re.search(r'Name="pressure_elements_\[[^"]*\]"[^>]*>\s*(.*?)\s*</DataArray>', txt, re.S)
    with open(os.path.join(OUT, "test08_pressure.txt"), "w") as fh:
        fh.write("\n".join(m.group(1).split()) + "\n")


if __name__ == "__main__":
    main()
