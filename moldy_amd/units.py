"""Moldy's program units and the conversions its input path applies.

Program units are amu, Angstrom, picosecond and the charge unit for which
Coulomb's law reads q1 q2 / r (src/defs.h:183-245, CODATA-1986 constants).
Input quantities are converted with ``unit_scale`` exactly as
src/convert.c:63-76 does (through logarithms).
"""
import math

PI = 3.14159265358979323846
KB_SI = 1.380658e-23
EPS0_SI = 8.854187817e-12
ELCHG = 1.60217733e-19
ROOT_4_PI_EPS = 1.05482230112e-05
AVOGAD = 6.0221367e23
AMU = 1.6605402e-27
RTAMU = 4.07497263794495e-14

MUNIT = AMU
LUNIT = 1.0e-10
TUNIT = 1.0e-12
EUNIT = MUNIT * (LUNIT / TUNIT) * (LUNIT / TUNIT)
QUNIT = RTAMU * 1.e-3 * ROOT_4_PI_EPS
KB = KB_SI / EUNIT                      # Boltzmann constant, program units
EPS0 = 0.25 / PI                        # permittivity, program units
CONV_E = 0.001 * AVOGAD * EUNIT         # program energy -> kJ/mol
CONV_P = MUNIT / (LUNIT * TUNIT * TUNIT) / 1.0e6   # program pressure -> MPa
CONV_Q = QUNIT / ELCHG                  # program charge -> electron charges

PROGRAM_UNITS = (MUNIT, LUNIT, TUNIT, QUNIT)
# Default input units of a Moldy run: amu, A, 0.1 ps (energies in kJ/mol), e
# (src/startup.c:198-201).
DEFAULT_INPUT_UNITS = (AMU, 1.0e-10, 1.0e-13, ELCHG)

# Dimensions (m, l, t) of each potential parameter, src/kernel.c:75-84.
POT_DIM = {
    0: [(1, 2, -2), (0, 1, 0)],
    1: [(1, 8, -2), (1, 2, -2), (0, -1, 0)],
    2: [(1, 2, -2), (0, -1, 0), (1, 2, -2), (0, -1, 0)],
    3: [(1, 2, -2), (0, -1, 0), (1, 14, -2), (1, 6, -2), (1, 8, -2), (1, 10, -2)],
    4: [(1, 6, -2), (1, 8, -2), (1, 14, -2)],
    5: [(0, 0, 0)],
    6: [(1, 2, -2), (0, 1, 0), (0, -1, 0), (1, 8, -2), (1, 2, -2), (0, -1, 0), (0, 1, 0)],
}
POT_NAMES = ["lennard-jones", "buckingham", "mcy", "generic", "hiw",
             "reserved for developer", "morse"]
POT_NPAR = [2, 3, 4, 6, 3, 1, 7]


def unit_scale(dim, unit_from, unit_to):
    """Scale factor for a quantity of dimension ``dim`` = (m, l, t[, q])."""
    m, l, t = dim[:3]
    q = dim[3] if len(dim) > 3 else 0
    ln = (m * (math.log(unit_from[0]) - math.log(unit_to[0]))
          + l * (math.log(unit_from[1]) - math.log(unit_to[1]))
          + t * (math.log(unit_from[2]) - math.log(unit_to[2]))
          + q * (math.log(unit_from[3]) - math.log(unit_to[3])))
    return math.exp(ln)


def density_to_program(rho_g_cm3):
    """g/cm^3 -> amu/A^3 (conv_control entry for density, src/convert.c:54)."""
    return rho_g_cm3 * unit_scale((1, -3, 0, 0), (0.001, 0.01, 1.0, 1.0), PROGRAM_UNITS)


def charge_to_program(q_e, unit_from=DEFAULT_INPUT_UNITS):
    return q_e * unit_scale((0, 0, 0, 1), unit_from, PROGRAM_UNITS)


def potpar_to_program(ptype, params, unit_from=DEFAULT_INPUT_UNITS):
    out = list(params)
    for i, d in enumerate(POT_DIM[ptype]):
        if i < len(out):
            out[i] = out[i] * unit_scale(d, unit_from, PROGRAM_UNITS)
    return out


def energy_kjmol_to_program(e):
    return e / CONV_E


def ns_per_day(evals_per_s, dt_fs):
    return evals_per_s * dt_fs * 86400.0 / 1.0e6
