"""Writes the fixtures of the reference's tests/20_limited_sorption (test_20.con)
into tests/golden/ref:
  * test20.msh (input/test_20.msh, as it is);
  * test20_A_mobile.txt: the substance A (no sorption; B and C are equal to it in
    the reference's output) of ref_output/test_20.con/transport_test_20/*.vtu at
    t = 0, 2.5, .., 15, one output time a line ("time v1 v2 ..." with the values
    as printed);
  * test20_sorbing.txt: the sorbing substances As-lin, Bs-lang, Cs-freund at
    t = 7.5 and 15, one a line ("name time v1 v2 ...")."""
import os
import re
import shutil

SRC = "/root/reference/tests/20_limited_sorption"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref")


def main():
    shutil.copyfile(os.path.join(SRC, "input", "test_20.msh"), os.path.join(OUT, "test20.msh"))
    d = os.path.join(SRC, "ref_output", "test_20.con")
    pvd = open(os.path.join(d, "transport_test_20.pvd")).read()
    rows = []
    for k, (t, f) in enumerate(re.findall(r'timestep="([^"]+)" group="" part="0" file="([^"]+)"', pvd)):
        if k % 5:
            continue
        txt = open(os.path.join(d, f)).read()
        m = re.search(r'Name="A_mobile_\[M/L\^3\]"[^>]*>\s*(.*?)\s*</DataArray>', txt, re.S)
        rows.append(" ".join([t] + m.group(1).split()))
    with open(os.path.join(OUT, "test20_A_mobile.txt"), "w") as fh:
        fh.write("\n".join(rows) + "\n")
    rows = []
    for k, (t, f) in enumerate(re.findall(r'timestep="([^"]+)" group="" part="0" file="([^"]+)"', pvd)):
        if k not in (15, 30):
            continue
        txt = open(os.path.join(d, f)).read()
        for name in ("As-lin", "Bs-lang", "Cs-freund"):
            m = re.search(r'Name="%s_mobile_\[M/L\^3\]"[^>]*>\s*(.*?)\s*</DataArray>' % re.escape(name), txt, re.S)
            rows.append(" ".join([name, t] + m.group(1).split()))
    with open(os.path.join(OUT, "test20_sorbing.txt"), "w") as fh:
        fh.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main()
