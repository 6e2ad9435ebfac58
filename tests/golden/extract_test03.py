"""Copies the fixtures of the reference's tests/03_transport_small_12d
(flow_old_gmsh.con) into tests/golden/ref:
  * input/test3.msh, test3.fbc, test3.tbc -> test03.msh, test03.fbc, test03.tbc;
  * ref_output/flow_old_gmsh.con/test3.msh (flow output) and water_balance.txt ->
    test03_flow.msh, test03_water_balance.txt, as they are;
  * of test3-transport.msh and mass_balance.txt (output every 0.01 to t = 1) the
    sections and blocks at t = 0.1, 0.5 and 1 only, unchanged:
    test03_transport_sel.msh, test03_mass_balance_sel.txt."""
import os
import shutil

SRC = "/root/reference/tests/03_transport_small_12d"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")
TIMES = ("0.1", "0.5", "1")
MB_TIMES = ("# Time: 0.100000", "# Time: 0.500000", "# Time: 1.000000")


def main():
    for src, dst in [("input/test3.msh", "test03.msh"), ("input/test3.fbc", "test03.fbc"),
                     ("input/test3.tbc", "test03.tbc"), ("ref_output/flow_old_gmsh.con/test3.msh", "test03_flow.msh"),
                     ("ref_output/flow_old_gmsh.con/water_balance.txt", "test03_water_balance.txt")]:
        shutil.copyfile(os.path.join(SRC, src), os.path.join(OUT, dst))
    d = os.path.join(SRC, "ref_output", "flow_old_gmsh.con")
    txt = open(os.path.join(d, "test3-transport.msh")).read()
    secs = ["$ElementData\n" + p.split("$EndElementData\n")[0] + "$EndElementData\n"
            for p in txt.split("$ElementData\n")[1:]]
    keep = [s for s in secs if s.split("\n")[4] in TIMES]
    with open(os.path.join(OUT, "test03_transport_sel.msh"), "w") as fh:
        fh.write("".join(keep))
    head = "# ----------------------------------------------- MASS BALANCE"
    mb = open(os.path.join(d, "mass_balance.txt")).read()
    blocks = [head + b for b in mb.split(head)[1:]]
    keep = [b for b in blocks if b.split("\n")[1].strip() in MB_TIMES]
    with open(os.path.join(OUT, "test03_mass_balance_sel.txt"), "w") as fh:
        fh.write("".join(keep))


if __name__ == "__main__":
    main()
