"""Water- and mass-balance files in the reference's layouts (fl_write_water_balance,
fl_write_mass_balance and the handle-level writers), compared with the reference's
own output files the way its test harness compares them (ndiff: the same
non-numeric tokens, numbers within a tolerance), plus a column check: every
token starts at the column it starts at in the reference's file.

Values come from the oracle (CPU tests) or the GPU handles (gpu tests); the
layout from the library's writer.  Files: tests/01 water_balance.out, tests/13
and tests/16 mass_balance.txt, tests/16 water_balance.txt."""
import re

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_case16 import load_case16

_TOK = re.compile(r"\S+")


def _num(s):
    try:
        return float(s)
    except ValueError:
        return None


def ndiff(ours, gold, rtol, atol, columns=True):
    """Line by line: token texts equal unless both are numbers within rtol*|gold|+atol;
    with `columns` (fixed-width tables) also the token start columns equal.  Returns
    the number of numeric tokens compared."""
    a, b = ours.split("\n"), gold.split("\n")
    assert len(a) == len(b), (len(a), len(b))
    n = 0
    for k, (la, lb) in enumerate(zip(a, b)):
        ta, tb = list(_TOK.finditer(la)), list(_TOK.finditer(lb))
        assert len(ta) == len(tb), (k, la, lb)
        if columns:
            assert [m.start() for m in ta] == [m.start() for m in tb], (k, la, lb)
        for ma, mb in zip(ta, tb):
            x, y = _num(ma.group()), _num(mb.group())
            if x is None or y is None or ma.group() == mb.group():
                assert ma.group() == mb.group(), (k, la, lb)
                n += x is not None
                continue
            assert abs(x - y) <= rtol * abs(y) + atol, (k, la, lb)
            n += 1
    return n


def _darcy_oracle(name, mesh_file):
    om = O.Mesh(cases.ref_path(mesh_file))
    d = cases.darcy_case(name, om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                aniso=d["aniso"])
    D.solve(0, rtol=1e-15)
    return om, d, D


def test_water_balance_test01(tmp_path):
    """tests/01 water_balance.out: four boundary regions, four bulk regions, no sources."""
    om, d, D = _darcy_oracle("test01", "test01_mesh.msh")
    gm = G.Mesh(cases.ref_path("test01_mesh.msh"))
    assert [r["id"] for r in gm.regions] == [r["id"] for r in om.regions]
    assert not d["f"].any()
    p = str(tmp_path / "water_balance.out")
    G.write_water_balance(gm, p, float("inf"), *D.water_balance(), np.zeros(gm.n_bulk_regions))
    n = ndiff(open(p).read(), open(cases.ref_path("test01_water_balance.out")).read(), 2e-5, 1e-9)
    assert n >= 4 * 5 + 3 + 4 * 3 + 1  # ids, balances, times and totals of both tables


def test_water_balance_test16(tmp_path):
    """tests/16 water_balance.txt; its total balance -1.35e-6 is the reference's
    iterative-solve residual, hence the absolute tolerance."""
    c = load_case16()
    om = O.Mesh(cases.ref_path("test16.msh"))
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    D.solve(0, rtol=1e-15)
    p = str(tmp_path / "water_balance.txt")
    G.write_water_balance(c["gm"], p, float("inf"), *D.water_balance(), np.zeros(1))
    ndiff(open(p).read(), open(cases.ref_path("test16_water_balance.txt")).read(), 2e-5, 2e-6)


def test_mass_balance_test13(tmp_path):
    """tests/13 mass_balance.txt: 201 blocks (t = 0, 0.5, ..., 100), three bulk regions."""
    om, d, D = _darcy_oracle("test13", "test13.msh")
    gm = G.Mesh(cases.ref_path("test13.msh"))
    E = om.n_elements
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.ones(E), cases.test13_bc_conc(om), np.zeros(E), 100.0, 0.5)
    p = str(tmp_path / "mass_balance.txt")
    for k in range(201):
        T.update(0.5 * k)
        G.write_mass_balance(gm, p, T.time, 0.0, ["A"], T.mass_balance())
    n = ndiff(open(p).read(), open(cases.ref_path("test13_mass_balance.txt")).read(), 2e-5, 1e-9)
    assert n > 201 * 15


def test_mass_balance_test16(tmp_path):
    """tests/16 mass_balance.txt: nine blocks, sources from $ElementData, including the
    t = 0 block (source evaluated with the zero previous concentration)."""
    c = load_case16()
    om = O.Mesh(cases.ref_path("test16.msh"))
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    D.solve(0, rtol=1e-15)
    T = O.Transport(om, D.solution[:D.S], c["cs"], c["por"], c["bcc"], c["init"], 5.7, 0.8)
    T.set_sources(*c["src"])
    p = str(tmp_path / "mass_balance.txt")
    for t in [0.0] + [0.8 * k for k in range(1, 8)] + [5.7]:
        T.update(t)
        G.write_mass_balance(c["gm"], p, T.time, 0.0, ["conc"], T.mass_balance())
    ndiff(open(p).read(), open(cases.ref_path("test16_mass_balance.txt")).read(), 2e-5, 1e-9)


def test_writer_rejects_bad_path(tmp_path):
    gm = G.Mesh(cases.ref_path("test13.msh"))
    with pytest.raises(G.FlowError):
        G.write_water_balance(gm, str(tmp_path / "no" / "x.txt"), 0.0, [0.0], [0.0], [0.0], np.zeros(3))


@pytest.mark.gpu
def test_balance_files_gpu(tmp_path):
    """The handle-level writers on GPU results: tests/13 flow + transport and tests/16
    flow + transport with sources, against the reference's files."""
    om = O.Mesh(cases.ref_path("test13.msh"))
    d = cases.darcy_case("test13", om)
    gm = G.Mesh(cases.ref_path("test13.msh"))
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    E = gm.n_elements
    T = G.ConvectionTransport(gm, np.ones(E), np.ones(E), cases.test13_bc_conc(om), np.zeros(E), 100.0, 0.5)
    T.set_velocity_field(darcy=g)
    p = str(tmp_path / "mb13.txt")
    for k in range(201):
        T.update_solution(0.5 * k)
        T.write_mass_balance(p, ["A"])
    ndiff(open(p).read(), open(cases.ref_path("test13_mass_balance.txt")).read(), 2e-5, 1e-9)

    c = load_case16()
    g16 = G.DarcyFlowMH_Steady(c["gm"])
    g16.set_data(c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    g16.update_solution(r_tol=1e-14, a_tol=1e-14)
    pw = str(tmp_path / "wb16.txt")
    g16.write_water_balance(pw)
    ndiff(open(pw).read(), open(cases.ref_path("test16_water_balance.txt")).read(), 2e-5, 2e-6)
    T16 = G.ConvectionTransport(c["gm"], c["por"], c["cs"], c["bcc"], c["init"], 5.7, 0.8)
    T16.set_velocity_field(darcy=g16)
    T16.set_sources(*c["src"])
    pm = str(tmp_path / "mb16.txt")
    for t in [0.0] + [0.8 * k for k in range(1, 8)] + [5.7]:
        T16.update_solution(t)
        T16.write_mass_balance(pm, ["conc"])
    ndiff(open(pm).read(), open(cases.ref_path("test16_mass_balance.txt")).read(), 2e-5, 1e-9)
