"""Copies the fixtures of the reference's tests/02_transport_12d (flow_gmsh.con)
into tests/golden/ref, as they are:
  * input/test2.msh, test2.fbc, test2.tbc -> test02.msh, test02.fbc, test02.tbc;
  * ref_output/flow_gmsh.con/test2.msh (flow output) -> test02_flow.msh;
  * ref_output/flow_gmsh.con/transport.msh, mass_balance.txt, water_balance.txt
    -> test02_transport.msh, test02_mass_balance.txt, test02_water_balance.txt."""
import os
import shutil

SRC = "/root/reference/tests/02_transport_12d"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")
FILES = [("input/test2.msh", "test02.msh"), ("input/test2.fbc", "test02.fbc"), ("input/test2.tbc", "test02.tbc"),
         ("ref_output/flow_gmsh.con/test2.msh", "test02_flow.msh"),
         ("ref_output/flow_gmsh.con/transport.msh", "test02_transport.msh"),
         ("ref_output/flow_gmsh.con/mass_balance.txt", "test02_mass_balance.txt"),
         ("ref_output/flow_gmsh.con/water_balance.txt", "test02_water_balance.txt")]


def main():
    for src, dst in FILES:
        shutil.copyfile(os.path.join(SRC, src), os.path.join(OUT, dst))


if __name__ == "__main__":
    main()
