"""Writes tests/golden/ref/test10_lmh_pressure.txt: the element pressures
(pressure_elements_[L], the values as printed) of the reference's
tests/10_unsteady_2d/ref_output/flow_LMH.con/test10/test10-00000{0..5}.vtu,
one output time per line ("time v1 v2 ..."), times from test10.pvd."""
import os
import re

SRC = "/root/reference/tests/10_unsteady_2d/ref_output/flow_LMH.con"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref", "test10_lmh_pressure.txt")


def main():
    pvd = open(os.path.join(SRC, "test10.pvd")).read()
    rows = []
    for t, f in re.findall(r'timestep="([^"]+)" group="" part="0" file="([^"]+)"', pvd):
        txt = open(os.path.join(SRC, f)).read()
        m = re.search(r'Name="pressure_elements_\[L\]" format="ascii">\s*(.*?)\s*</DataArray>', txt, re.S)
        rows.append(" ".join([t] + m.group(1).split()))
    with open(OUT, "w") as fh:
        fh.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main()
