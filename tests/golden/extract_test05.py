"""Writes the fixtures of the reference's tests/05_flow_boundary_conditions into
tests/golden/ref:
  * test05.msh (the input mesh, as it is);
  * test05_<case>_fields.txt for case in dirichlet, neumann, robin: the flow output
    of ref_output/flow_<case>.con/test05/test05-000000.vtu, one array a line
    ("name v1 v2 ..." with the values as printed): pressure_nodes,
    pressure_elements, velocity_elements (3 components per element);
  * test05_<case>_water_balance.out (ref_output/flow_<case>.con/water_balance.out)."""
import os
import re
import shutil

SRC = "/root/reference/tests/05_flow_boundary_conditions"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")
ARRAYS = ["pressure_nodes", "pressure_elements", "velocity_elements"]


def main():
    shutil.copyfile(os.path.join(SRC, "input", "test05.msh"), os.path.join(OUT, "test05.msh"))
    for case in ("dirichlet", "neumann", "robin"):
        d = os.path.join(SRC, "ref_output", "flow_%s.con" % case)
        txt = open(os.path.join(d, "test05", "test05-000000.vtu")).read()
        rows = []
        for name in ARRAYS:
            m = re.search(r'Name="%s_\[[^"]*\]"[^>]*>\s*(.*?)\s*</DataArray>' % name, txt, re.S)
            rows.append(" ".join([name] + m.group(1).split()))
        with open(os.path.join(OUT, "test05_%s_fields.txt" % case), "w") as fh:
            fh.write("\n".join(rows) + "\n")
        shutil.copyfile(os.path.join(d, "water_balance.out"), os.path.join(OUT, "test05_%s_water_balance.out" % case))


if __name__ == "__main__":
    main()
