"""Writes tests/golden/ref/test15_transport_conc.txt and test15_flow_pressure.txt:
the element concentrations (conc_mobile_[M/L^3]) and element pressures
(pressure_elements_[L]) of the reference's
tests/15_unsteady_and_transport/ref_output/flow_trans_explicit.con VTK output,
values as printed, one output time per line ("time v1 v2 ..."), times from the
.pvd files."""
import os
import re

SRC = "/root/reference/tests/15_unsteady_and_transport/ref_output/flow_trans_explicit.con"
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")


def extract(pvd, name, out):
    rows = []
    for t, f in re.findall(r'timestep="([^"]+)" group="" part="0" file="([^"]+)"', open(os.path.join(SRC, pvd)).read()):
        txt = open(os.path.join(SRC, f)).read()
        m = re.search(r'Name="%s" format="ascii">\s*(.*?)\s*</DataArray>' % re.escape(name), txt, re.S)
        rows.append(" ".join([t] + m.group(1).split()))
    with open(os.path.join(HERE, out), "w") as fh:
        fh.write("\n".join(rows) + "\n")


def main():
    extract("transport.pvd", "conc_mobile_[M/L^3]", "test15_transport_conc.txt")
    extract("test15.pvd", "pressure_elements_[L]", "test15_flow_pressure.txt")


if __name__ == "__main__":
    main()
