"""Fluid presets: component tables and EOS settings as nn_fluid_desc.

Component data follow the reference's case tables:
  H2O, CO2, nC10   case/h2o_co2_nc10_2d_benchmark/case_config.hpp:49-58
  C1 (CH4)         case/3p4c_pr_reservoir/case_config.hpp:90-99
  nC4, nC16        case/scw_co2_nc4_nc16_3d_migration/case_config.hpp:46-55
  N2, C2, C3       case/panfili2025_case3_fullphysics/case_config.hpp:111-147
EOS constants (Omega_a = 0.45724, Omega_b = 0.07780, eos type 5 = PR78 m(omega)
for omega > 0.49, m1 = 1+sqrt2, m2 = 1-sqrt2) follow case_config.hpp:60-63,77.
Soreide-Whitson aqueous H2O-solute k_ij: CO2 -0.06609, nC10 -0.14229
(case_config.hpp:95-100); components without one fall back to k_ij as in
cubic_eos.hpp:2389-2404.

Binary interaction matrices: the H2O/CO2/nC10 preset uses k(H2O,CO2)=0.19,
k(H2O,nC10)=0.5 and zero otherwise (the benchmark workload statement); the
reference's own locked case values (0.1896, 0.5, 0.1141) are available as
``bips="case"``.
"""
from __future__ import annotations

from ._abi import FluidDesc, NN_MAX_NC, NN_MODEL_CPA, NN_MODEL_PR, NN_MODEL_SW

R = 8.31446261815324
_PSI = (0.45359237 * 9.80665) / (0.0254 * 0.0254)   # common/include/common/units.hpp

# name: (Tc [K], Pc [Pa], Vc [m3/mol], omega, M [kg/mol])
COMPONENTS = {
    "H2O": (647.096, 22.064e6, 5.5948e-5, 0.3443, 0.018015268),
    "CO2": (304.1282, 7.3773e6, 9.4118e-5, 0.22394, 0.0440098),
    "C1": (190.564, 46.00e5, 9.9e-5, 0.008, 0.016043),
    "nC4": (425.125, 3.7960e6, 2.5492e-4, 0.20081, 0.0581222),
    "nC10": (617.70, 2.103e6, 6.10e-4, 0.4920, 0.14228168),
    "nC16": (723.0, 1.410e6, 9.0e-4, 0.742, 0.2264412),
    # N2, C2, C3: Tc, Pc, omega, M of case/panfili2025_case3_fullphysics/case_config.hpp:111-147
    # (degR / psia converted); Vc from the Zc = 0.27 estimate that case uses.
    "N2": (227.16 * 5.0 / 9.0, 492.31 * _PSI, 0.27 * R * 227.16 * 5.0 / 9.0 / (492.31 * _PSI),
           0.04000, 0.028013),
    "C2": (549.77 * 5.0 / 9.0, 708.34 * _PSI, 0.27 * R * 549.77 * 5.0 / 9.0 / (708.34 * _PSI),
           0.09860, 0.03007),
    "C3": (665.64 * 5.0 / 9.0, 618.70 * _PSI, 0.27 * R * 665.64 * 5.0 / 9.0 / (618.70 * _PSI),
           0.15240, 0.044097),
}

# Pairwise k_ij used for the six-component mixture (symmetric).
_PAIR_KIJ_NC6 = {
    ("H2O", "CO2"): 0.1896, ("H2O", "C1"): 0.4850, ("H2O", "nC4"): 0.5,
    ("H2O", "nC10"): 0.5, ("H2O", "nC16"): 0.5,
    ("CO2", "C1"): 0.12, ("CO2", "nC4"): 0.1277, ("CO2", "nC10"): 0.1141,
    ("CO2", "nC16"): 0.09,
}

SW_AQUEOUS_KIJ = {"CO2": -0.06609, "nC10": -0.14229}

# Corey curves of case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:97-126:
# swc, sorw, sgc, sorg, krw_max, krg_max, kro_max
BENCHMARK_RELPERM = (0.10, 0.15, 0.02, 0.10, 0.30, 0.80, 0.80)

PR_M1 = 2.4142135623730951
PR_M2 = -0.4142135623730951


def make_desc(names, kij, model: str = "PR", water: str = "H2O",
              eos_type: int = 5, omega_a: float = 0.45724, omega_b: float = 0.07780,
              salinity: float = 0.0) -> FluidDesc:
    nc = len(names)
    if not 2 <= nc <= NN_MAX_NC:
        raise ValueError("nc must be in [2, 8]")
    d = FluidDesc()
    d.nc = nc
    d.thermo_model = {"PR": NN_MODEL_PR, "SW": NN_MODEL_SW, "CPA": NN_MODEL_CPA}[model]
    d.eos_type = eos_type
    d.water_component = names.index(water) if water in names else -1
    d.omega_a, d.omega_b, d.m1, d.m2 = omega_a, omega_b, PR_M1, PR_M2
    for i, n in enumerate(names):
        tc, pc, vc, om, mw = COMPONENTS[n]
        d.tc[i], d.pc[i], d.vc[i], d.acentric[i], d.molar_mass[i] = tc, pc, vc, om, mw
        for j in range(nc):
            d.kij[i * NN_MAX_NC + j] = kij[i][j]
    if model == "SW":
        d.sw_salinity = salinity
        for i, n in enumerate(names):
            if n in SW_AQUEOUS_KIJ:
                d.sw_aqueous_kij[i] = SW_AQUEOUS_KIJ[n]
                d.sw_aqueous_kij_set[i] = 1
    for k, v in enumerate(BENCHMARK_RELPERM):
        d.relperm[k] = v
    return d


def with_aqueous_closures(d: FluidDesc, viscosity: str = "mcbride_wright",
                          volume: bool = True, co2: int = 1,
                          viscosity_max_unsupported: float = 1.0e-4) -> FluidDesc:
    """Switch on the Water-role closures of the reference's benchmark case
    (case/h2o_co2_nc10_2d_benchmark/case_config.hpp:70-74): IAPWS-IF97 + Garcia
    molar volume and McBride-Wright (or IAPWS-2008, `viscosity="iapws2008"`)
    viscosity, water molar mass 0.018015268 kg/mol, unsupported-solute limits
    1e-4 (the viscosity closure's: `viscosity_max_unsupported`; IAPWS-2008 counts
    every non-H2O fraction, CO2 included).  `co2`: index of CO2 (1 in the presets)."""
    from ._abi import NN_AQ_VISCOSITY_IAPWS2008, NN_AQ_VISCOSITY_MCBRIDE_WRIGHT, NN_AQ_VISCOSITY_NONE
    d.aqueous_volume = 1 if volume else 0
    d.aqueous_viscosity = {"mcbride_wright": NN_AQ_VISCOSITY_MCBRIDE_WRIGHT,
                           "iapws2008": NN_AQ_VISCOSITY_IAPWS2008,
                           "none": NN_AQ_VISCOSITY_NONE}[viscosity]
    d.co2_component = co2
    d.water_molar_mass = 0.018015268
    d.aqueous_volume_max_unsupported = 1.0e-4
    d.aqueous_viscosity_max_unsupported = float(viscosity_max_unsupported)
    return d


def h2o_co2_nc10(model: str = "PR", bips: str = "workload") -> FluidDesc:
    """Three-component H2O/CO2/nC10 fluid (PR78, eos type 5)."""
    names = ["H2O", "CO2", "nC10"]
    if bips == "workload":
        kij = [[0.0, 0.19, 0.5], [0.19, 0.0, 0.0], [0.5, 0.0, 0.0]]
    elif bips == "case":
        kij = [[0.0, 0.1896, 0.5], [0.1896, 0.0, 0.1141], [0.5, 0.1141, 0.0]]
    else:
        raise ValueError(bips)
    if model == "CPA":
        return cpa_h2o_co2_nc10()
    return make_desc(names, kij, model=model, eos_type=5 if model == "PR" else 1)


def six_component(model: str = "PR") -> FluidDesc:
    """H2O, CO2, C1, nC4, nC10, nC16 (PR78 / SW)."""
    names = ["H2O", "CO2", "C1", "nC4", "nC10", "nC16"]
    nc = len(names)
    kij = [[0.0] * nc for _ in range(nc)]
    for (a, b), v in _PAIR_KIJ_NC6.items():
        i, j = names.index(a), names.index(b)
        kij[i][j] = kij[j][i] = v
    return make_desc(names, kij, model=model, eos_type=5 if model == "PR" else 1)


def five_component(model: str = "PR") -> FluidDesc:
    """H2O, CO2, C1, nC4, nC10 (PR78 / SW), k_ij as the six-component preset."""
    names = ["H2O", "CO2", "C1", "nC4", "nC10"]
    nc = len(names)
    kij = [[0.0] * nc for _ in range(nc)]
    for (a, b), v in _PAIR_KIJ_NC6.items():
        if a in names and b in names:
            i, j = names.index(a), names.index(b)
            kij[i][j] = kij[j][i] = v
    return make_desc(names, kij, model=model, eos_type=5 if model == "PR" else 1)


# Pairwise k_ij of the eight-component preset (symmetric; unlisted pairs are 0).
_PAIR_KIJ_NC8 = dict(_PAIR_KIJ_NC6)
_PAIR_KIJ_NC8.update({
    ("H2O", "N2"): 0.4778, ("H2O", "C2"): 0.4920, ("H2O", "C3"): 0.5525,
    ("CO2", "N2"): -0.0170, ("CO2", "C2"): 0.1300, ("CO2", "C3"): 0.1350,
    ("N2", "C1"): 0.0311, ("N2", "C2"): 0.0515, ("N2", "C3"): 0.0852,
    ("N2", "nC4"): 0.0800, ("N2", "nC10"): 0.1100,
})


def eight_component(model: str = "PR") -> FluidDesc:
    """H2O, CO2, N2, C1, C2, C3, nC4, nC10 (PR78 / SW): a synthetic eight-component
    preset for the serving workload (N2/C2/C3 data of the Panfili case)."""
    names = ["H2O", "CO2", "N2", "C1", "C2", "C3", "nC4", "nC10"]
    nc = len(names)
    kij = [[0.0] * nc for _ in range(nc)]
    for (a, b), v in _PAIR_KIJ_NC8.items():
        if a in names and b in names:
            i, j = names.index(a), names.index(b)
            kij[i][j] = kij[j][i] = v
    return make_desc(names, kij, model=model, eos_type=5 if model == "PR" else 1)


def cpa_h2o_co2_nc10() -> FluidDesc:
    """PR-CPA H2O/CO2/nC10 of case_config.hpp:106-142 (4C water, CO2 B2 cross-association)."""
    names = ["H2O", "CO2", "nC10"]
    kij = [[0.0, 0.10566, 0.0], [0.10566, 0.0, 0.1141], [0.0, 0.1141, 0.0]]
    d = make_desc(names, kij, model="CPA", eos_type=5)
    d.cpa_physical_term = 1
    tc = [COMPONENTS[n][0] for n in names]
    pc = [COMPONENTS[n][1] for n in names]
    om = [COMPONENTS[n][3] for n in names]
    a0 = [0.15782, 0.45724 * R * R * tc[1] * tc[1] / pc[1],
          0.45724 * R * R * tc[2] * tc[2] / pc[2]]
    b = [1.4788e-5, 0.07780 * R * tc[1] / pc[1], 0.07780 * R * tc[2] / pc[2]]
    c1 = [0.6736,
          0.37464 + 1.54226 * om[1] - 0.26992 * om[1] * om[1],
          0.379642 + 1.48503 * om[2] - 0.164423 * om[2] * om[2]
          + 0.016666 * om[2] * om[2] * om[2]]
    eps = [16123.0, 8061.5, 0.0]
    beta = [0.069662, 0.15182, 0.0]
    donors = [2, 0, 0]
    acceptors = [2, 1, 0]
    for i in range(3):
        d.cpa_a0[i], d.cpa_b[i], d.cpa_c1[i] = a0[i], b[i], c1[i]
        d.cpa_epsilon[i], d.cpa_beta[i] = eps[i], beta[i]
        d.cpa_donors[i], d.cpa_acceptors[i] = donors[i], acceptors[i]
    d.cpa_cross_epsilon[0 * NN_MAX_NC + 1] = 8061.5
    d.cpa_cross_beta[0 * NN_MAX_NC + 1] = 0.15182
    d.cpa_max_iterations = 2000
    d.cpa_tolerance = 1e-12
    d.cpa_damping = 0.2
    return d


def cpa_water_methanol() -> FluidDesc:
    """sCPA water/methanol of test/src/benchmark/cpa_phase_benchmark.cpp:69-91
    (SRK physical term, 4C water + 2B methanol: the general multi-donor site
    solver), the state of the reference's published CPA micro-benchmark."""
    d = FluidDesc()
    d.nc = 2
    d.thermo_model = NN_MODEL_CPA
    d.eos_type = 1
    d.water_component = 0
    d.omega_a, d.omega_b, d.m1, d.m2 = 0.42748, 0.08664, 1.0, 0.0
    tc, pc, vc = (647.3, 512.6), (22.0483e6, 8.0959e6), (5.60e-5, 1.18e-4)
    om, mw = (0.344, 0.559), (0.018015, 0.032042)
    for i in range(2):
        d.tc[i], d.pc[i], d.vc[i], d.acentric[i], d.molar_mass[i] = tc[i], pc[i], vc[i], om[i], mw[i]
    d.kij[0 * NN_MAX_NC + 1] = d.kij[1 * NN_MAX_NC + 0] = -0.09
    d.cpa_physical_term = 0
    a0, b, c1 = (0.12277, 0.40531), (1.4515e-5, 3.0978e-5), (0.67359, 0.43102)
    eps, beta, sites = (16655.0, 24591.0), (0.0692, 0.0161), (2, 1)
    for i in range(2):
        d.cpa_a0[i], d.cpa_b[i], d.cpa_c1[i] = a0[i], b[i], c1[i]
        d.cpa_epsilon[i], d.cpa_beta[i] = eps[i], beta[i]
        d.cpa_donors[i] = d.cpa_acceptors[i] = sites[i]
    d.cpa_max_iterations, d.cpa_tolerance, d.cpa_damping = 2000, 1e-12, 0.2
    d.relperm[:] = BENCHMARK_RELPERM
    return d


def n_primary(nc: int, has_well: bool) -> int:
    """Indices::numPrimaryVariables: p, 3(nc-1) compositions, 3 saturations [, BHP]."""
    return 3 * nc + 1 + (1 if has_well else 0)


def n_phase_state(nc: int) -> int:
    """Indices::numPhaseStateVariables for the fully compositional mode: 3nc+8."""
    return 3 * nc + 8
