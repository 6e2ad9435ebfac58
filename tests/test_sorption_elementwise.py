"""The reaction module's sorption where the parameters are not constant on a
region: Sorption::compute_reaction re-initializes the isotherm from the
element's own values and solves it directly, without a table
(src/reaction/sorption.cc:172-213, 236-260; Isotherm::compute_reaction,
src/reaction/isotherm.cc:184-200), with scale_sorbed = phi (1 - por_m - por_imm)
rock_density molar_mass (sorption.cc:205).

Pins of the oracle's restatement:
* closed forms on a field without flow (every step leaves the mobile
  concentration as it is, so each element's total mass por_m c0 stays fixed and
  the first sorption step reaches the isotherm): linear c = por c0 / (por + ss k),
  Langmuir from its quadratic, Freundlich from scipy's brentq, with a different
  porosity per element and the immobile porosity and phi of the dual-porosity
  data in scale_sorbed; a dropped phi, por_imm, rock density or molar mass
  factor misses them by 10% or more;
* tests/20 of the reference through the direct path (porosity perturbed by
  1e-12 so no region is constant) against the reference's output, which its
  tables produced, within the same bounds as the table path;
* the element-wise entry point with the same values equals the per-region one
  bit for bit (both solve directly);
* failures the reference stops on: a negative total mass (isotherm.cc:62) and
  an empty root bracket without a solubility limit (toms748_solve throws).
GPU: the same cases through the C ABI, within 1e-12 of the oracle (device pow),
and the same failures as FL_E_SORPTION."""
import numpy as np
import pytest
from scipy.optimize import brentq

import cases
import flowgpu as G
import oracle as O
from test_case20 import NAMES, SORB, SORB_TOL, load20, run6, sorbing_gold

K_LIN, (W_LANG, A_LANG), (K_FR, N_FR) = 0.6, (0.4, 0.6), (0.6, 0.4)
ROCK, MOLAR, PHI, POR_IMM, C0 = 2.0, np.array([1.0, 0.5, 2.0]), 0.5, 0.05, 0.5


def porosity(E):
    return 0.25 + 0.01 * (np.arange(E) % 3)


def no_flow(om, gm=None):
    """test20's mesh, zero fluxes, three sorbing substances starting at C0"""
    E, NB, S = om.n_elements, om.n_bc_elements, om.n_sides
    por = porosity(E)
    args = (np.ones(E), por, np.zeros((3, NB)), np.full((3, E), C0), 2.0, 0.5)
    if gm is None:
        T = O.Transport(om, np.zeros(S), *args)
    else:
        T = G.ConvectionTransport(gm, por, np.ones(E), np.zeros((3, NB)), np.full((3, E), C0), 2.0, 0.5)
        T.set_velocity_field(side_flux=np.zeros(S))
    # dual porosity and ConvectionTransport's own sorption off; their por_imm and phi
    # enter the reaction module's scale_sorbed
    if gm is None:
        T.set_phases(False, False, np.full(E, POR_IMM), phi=np.full(E, PHI))
    else:
        T.set_dual_porosity_sorption(False, False, np.full(E, POR_IMM), phi=np.full(E, PHI))
    return T, por


def sorb_args(nreg=1):
    return dict(subst=[0, 1, 2], types=[[1, 3, 2]] * nreg, mult=[[K_LIN, W_LANG, K_FR]] * nreg,
                second=[[0.0, A_LANG, N_FR]] * nreg, rock_density=ROCK, molar_mass=MOLAR, solubility=[1.0] * 3)


def closed_forms(por):
    ss = PHI * (1 - por - POR_IMM) * ROCK * MOLAR[:, None]
    c_lin = por * C0 / (por + ss[0] * K_LIN)
    B = por + ss[1] * W_LANG * A_LANG - por * A_LANG * C0
    c_lang = (-B + np.sqrt(B * B + 4 * por * A_LANG * por * C0)) / (2 * por * A_LANG)
    c_fr = np.array([brentq(lambda c: p * c + s * K_FR * c ** N_FR - p * C0, 0.0, C0, xtol=1e-17, rtol=1e-15)
                     for p, s in zip(por, ss[2])])
    return np.array([c_lin, c_lang, c_fr])


def test_direct_isotherms_against_closed_forms():
    om, E, NB, _, _ = load20()
    T, por = no_flow(om)
    T.set_reaction_sorption(**sorb_args())
    want = closed_forms(por)
    for t in (0.5, 2.0):
        T.update(t)
        c, s = T.conc(), T.reaction_sorbed()
        assert np.max(np.abs(c - want) / want) < 1e-12, t
        # the sorbed amounts sit exactly on the isotherms
        assert np.allclose(s[0], K_LIN * c[0], rtol=1e-13, atol=0)
        assert np.allclose(s[1], W_LANG * A_LANG * c[1] / (A_LANG * c[1] + 1), rtol=1e-13, atol=0)
        assert np.allclose(s[2], K_FR * c[2] ** N_FR, rtol=1e-13, atol=0)
    # what the closed forms exclude: scale_sorbed without phi and por_imm
    por_only = (1 - por) * ROCK * MOLAR[0]
    assert np.min(np.abs(por * C0 / (por + por_only * K_LIN) - want[0]) / want[0]) > 0.1


def test_elementwise_entry_equals_per_region_direct_path():
    om, E, NB, _, _ = load20()
    T1, _ = no_flow(om)
    T1.set_reaction_sorption(**sorb_args())
    T2, _ = no_flow(om)
    a = sorb_args()
    T2.set_reaction_sorption_elementwise(a["subst"], np.array(a["types"][0])[:, None] + np.zeros((3, E), int),
                                         np.array(a["mult"][0])[:, None] + np.zeros((3, E)),
                                         np.array(a["second"][0])[:, None] + np.zeros((3, E)), np.full(E, ROCK),
                                         MOLAR, solubility=[1.0] * 3)
    for t in (0.5, 1.0):
        T1.update(t)
        T2.update(t)
        assert np.array_equal(T1.conc(), T2.conc())
        assert np.array_equal(T1.reaction_sorbed(), T2.reaction_sorbed())


def case20_direct(om, E, NB, flux, gm=None):
    por = 0.25 * (1 + 1e-12 * (np.arange(E) % 2))
    if gm is None:
        T = O.Transport(om, flux, np.ones(E), por, np.zeros((6, NB)), np.zeros((6, E)), 15.0, 0.5)
        T.set_bc_function(lambda t: np.full((6, NB), 0.05 * t))
    else:
        T = G.ConvectionTransport(gm, por, np.ones(E), np.zeros((6, NB)), np.zeros((6, E)), 15.0, 0.5)
        T.set_velocity_field(side_flux=flux)
        T.set_bc_conc_function(lambda t: np.full((6, NB), 0.05 * t))
    T.set_reaction_sorption(**SORB)
    return T


def test_case20_direct_path_against_reference_output():
    om, E, NB, bc, goldA = load20()
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    T = case20_direct(om, E, NB, D.solution[:D.S])
    res = run6(T.update, T.conc)
    for (name, t), g in sorbing_gold().items():
        assert np.max(np.abs(res[t][NAMES.index(name)] - g)) < SORB_TOL[name], (name, t)
    c, s = T.conc(), T.reaction_sorbed()
    assert np.allclose(s[2], 0.4 * (0.6 * c[3]) / (0.6 * c[3] + 1), rtol=1e-12, atol=1e-15)  # Bs-lang exact


def test_sorption_failures_oracle():
    om, E, NB, _, _ = load20()
    # no solubility on the direct path: the reference's bracket is empty
    T, _ = no_flow(om)
    a = sorb_args()
    a["solubility"] = None
    T.set_reaction_sorption(**a)
    with pytest.raises(ValueError, match="isotherm"):
        T.update(0.5)
    # a negative total mass in a tabulated region
    S = om.n_sides
    T = O.Transport(om, np.zeros(S), np.ones(E), np.full(E, 0.25), np.zeros((3, NB)), np.full((3, E), -0.1), 2.0, 0.5)
    T.set_reaction_sorption(**sorb_args())
    with pytest.raises(ValueError, match="negative"):
        T.update(0.5)


@pytest.mark.gpu
def test_direct_isotherms_gpu_against_oracle():
    om, E, NB, bc, goldA = load20()
    gm = G.Mesh(cases.ref_path("test20.msh"))
    T0, por = no_flow(om)
    T1, _ = no_flow(om, gm)
    T0.set_reaction_sorption(**sorb_args())
    T1.set_reaction_sorption(**sorb_args())
    for t in (0.5, 2.0):
        assert T0.update(t) == T1.update_solution(t)
        assert np.max(np.abs(T1.conc() - T0.conc())) <= 1e-12
        assert np.max(np.abs(T1.reaction_sorbed() - T0.reaction_sorbed())) <= 1e-12
    assert np.max(np.abs(T1.conc() - closed_forms(por)) / closed_forms(por)) < 1e-12
    # tests/20 through the direct path on the oracle's fluxes
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    flux = D.solution[:D.S]
    T0 = case20_direct(om, E, NB, flux)
    T1 = case20_direct(om, E, NB, flux, gm)
    for k in range(31):
        assert T0.update(0.5 * k) == T1.update_solution(0.5 * k)
        assert np.max(np.abs(T1.conc() - T0.conc())) <= 1e-12, k
    assert np.max(np.abs(T1.reaction_sorbed() - T0.reaction_sorbed())) <= 1e-12


@pytest.mark.gpu
def test_sorption_failures_gpu():
    om, E, NB, _, _ = load20()
    gm = G.Mesh(cases.ref_path("test20.msh"))
    T, _ = no_flow(om, gm)
    a = sorb_args()
    a["solubility"] = None
    T.set_reaction_sorption(**a)
    with pytest.raises(G.FlowError, match="bracket") as ei:
        T.update_solution(0.5)
    assert ei.value.status == G.E_SORPTION
    S = om.n_sides
    T = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), np.zeros((3, NB)), np.full((3, E), -0.1), 2.0, 0.5)
    T.set_velocity_field(side_flux=np.zeros(S))
    T.set_reaction_sorption(**sorb_args())
    with pytest.raises(G.FlowError, match="negative total mass") as ei:
        T.update_solution(0.5)
    assert ei.value.status == G.E_SORPTION
    print(str(ei.value), flush=True)


@pytest.mark.gpu
def test_bc_function_failure_stops_the_update():
    om, E, NB, _, _ = load20()
    gm = G.Mesh(cases.ref_path("test20.msh"))
    T = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), np.zeros((1, NB)), np.zeros((1, E)), 2.0, 0.5)
    T.set_velocity_field(side_flux=np.zeros(om.n_sides))

    def bad(t):
        if t >= 0.5:
            return np.zeros((1, NB + 1))  # wrong shape from the second step on
        return np.full((1, NB), 1.0)

    T.set_bc_conc_function(bad)
    with pytest.raises(ValueError) as ei:
        T.update_solution(2.0)
    assert isinstance(ei.value.__cause__, G.FlowError) and ei.value.__cause__.status == G.E_CALLBACK
    t, _, _ = T.times()
    assert t == 0.5  # the first step ran, the second was not taken
