"""tests/14_trans_time_bc of the reference from its input files (trans_explicit.con):
steady flow from test14.fbc, then explicit transport to t = 10 with output every 1.0
and boundary concentrations that change at t = 1, 3, 6 and 7 (four legacy .tbc
files).  New boundary data take effect at the first step that starts at their
time (ConvectionTransport::compute_one_step, transport.cc:598-627); the balance
written at that output time still shows the previous data, as the reference's.
Checked against ref_output/trans_explicit.con/mass_balance.txt (all 11 blocks)
and water_balance.txt on the oracle and on the GPU, and the GPU concentrations
bitwise against the oracle's on the same fluxes."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff

CHANGES = {1: 0, 3: 1, 6: 2, 7: 3}  # output time -> tbc file taking effect from there


def load_case14():
    ref = cases.ref_path
    gm = G.Mesh(ref("test14.msh"))
    b = G.OldBcdInput(gm)
    bc = b.read_flow(ref("test14.fbc"))
    tbc = [b.read_transport(ref("test14_tbc_%03d" % k), 1) for k in range(4)]
    E = gm.n_elements
    aniso = np.zeros((E, 9))
    aniso[:, [0, 4, 8]] = 0.1  # anisotropy 0.1 on region 100
    return dict(gm=gm, bc=bc, tbc=tbc, aniso=aniso, E=E)


def run(T, update, gm, path, tbc, conc=None):
    hist = []
    for k in range(11):
        update(float(k))
        if hasattr(T, "write_mass_balance"):
            T.write_mass_balance(path, ["A"])
        else:
            G.write_mass_balance(gm, path, T.time, 0.0, ["A"], T.mass_balance())
        if conc is not None:
            hist.append(conc().copy())
        if k in CHANGES:
            T.set_bc_conc(tbc[CHANGES[k]])
    return hist


def test_case14_oracle_against_reference_output(tmp_path):
    c14 = load_case14()
    gm, E = c14["gm"], c14["E"]
    om = O.Mesh(cases.ref_path("test14.msh"))
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c14["bc"], aniso=c14["aniso"])
    D.solve(0, rtol=1e-15)
    pw = str(tmp_path / "water_balance.txt")
    G.write_water_balance(gm, pw, float("inf"), *D.water_balance(), np.zeros(1))
    ndiff(open(pw).read(), open(cases.ref_path("test14_water_balance.txt")).read(), 2e-5, 1e-9)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.ones(E), np.zeros((1, gm.n_bc_elements)), np.zeros(E),
                    10.0, 1.0)
    p = str(tmp_path / "mass_balance.txt")
    run(T, T.update, gm, p, c14["tbc"])
    n = ndiff(open(p).read(), open(cases.ref_path("test14_mass_balance.txt")).read(), 2e-5, 1e-9)
    assert n > 11 * 15


@pytest.mark.gpu
def test_case14_gpu_against_reference_output(tmp_path):
    c14 = load_case14()
    gm, E = c14["gm"], c14["E"]
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c14["bc"], aniso=c14["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    T = G.ConvectionTransport(gm, np.ones(E), np.ones(E), np.zeros((1, gm.n_bc_elements)), np.zeros(E), 10.0, 1.0)
    T.set_velocity_field(darcy=g)
    p = str(tmp_path / "mb.txt")
    run(T, T.update_solution, gm, p, c14["tbc"])
    ndiff(open(p).read(), open(cases.ref_path("test14_mass_balance.txt")).read(), 2e-5, 1e-9)
    # bitwise against the oracle on the oracle's fluxes
    om = O.Mesh(cases.ref_path("test14.msh"))
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c14["bc"], aniso=c14["aniso"])
    D.solve(0, rtol=1e-15)
    flux = D.solution[:D.S]
    zero = np.zeros((1, gm.n_bc_elements))
    T0 = O.Transport(om, flux, np.ones(E), np.ones(E), zero, np.zeros(E), 10.0, 1.0)
    T1 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), zero, np.zeros(E), 10.0, 1.0)
    T1.set_velocity_field(side_flux=flux)
    h0 = run(T0, T0.update, gm, str(tmp_path / "a.txt"), c14["tbc"], conc=T0.conc)
    h1 = run(T1, T1.update_solution, gm, str(tmp_path / "b.txt"), c14["tbc"], conc=T1.conc)
    for k, (a, b) in enumerate(zip(h0, h1)):
        assert np.array_equal(a, b), k
