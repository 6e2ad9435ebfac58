"""Writes tests/golden/ref/test17_transport_conc.txt: the concentrations of the five
substances (A..E, values as printed) on the three elements of the reference's
tests/17_decay_pade_approximation/ref_output/flow_decay_vtk.con/transport_decay/
transport_decay-0000NN.vtu, one output time per line:
"time A1 A2 A3 B1 .. E3", times from transport_decay.pvd."""
import os
import re

SRC = "/root/reference/tests/17_decay_pade_approximation/ref_output/flow_decay_vtk.con"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref", "test17_transport_conc.txt")


def main():
    pvd = open(os.path.join(SRC, "transport_decay.pvd")).read()
    rows = []
    for t, f in re.findall(r'timestep="([^"]+)" group="" part="0" file="([^"]+)"', pvd):
        txt = open(os.path.join(SRC, f)).read()
        vals = [t]
        for sub in "ABCDE":
            m = re.search(r'Name="%s_mobile_\[M/L\^3\]" format="ascii">\s*(.*?)\s*</DataArray>' % sub, txt, re.S)
            vals += m.group(1).split()
        rows.append(" ".join(vals))
    with open(OUT, "w") as fh:
        fh.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main()
