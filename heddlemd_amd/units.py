"""SI <-> Hartree atomic units.

Constants are the CODATA 2018 values the reference generates into
src/units/atomic_constants.rs:10-35; inside the engine k_B = 1 and k_C = 1
(src/gpu/kernels.rs:176), so a temperature is stored as k_B*T in Hartree.
"""
LENGTH_BOHR_TO_M = 5.29177210903e-11
MASS_ME_TO_KG = 9.1093837015e-31
CHARGE_E_TO_C = 1.602176634e-19
ENERGY_HARTREE_TO_J = 4.359744722207101e-18
TIME_AU_TO_S = 2.4188843265857195e-17
FORCE_AU_TO_N = 8.238723498254079e-08
PRESSURE_AU_TO_PA = 29421015696522.1
TEMPERATURE_AU_TO_K = 315775.0248040668
VELOCITY_AU_TO_M_PER_S = 2187691.2636411153
AMU_TO_KG = 1.66053906660e-27
AVOGADRO = 6.02214076e23


def length(m):
    return m / LENGTH_BOHR_TO_M


def mass(kg):
    return kg / MASS_ME_TO_KG


def charge(c):
    return c / CHARGE_E_TO_C


def energy(j):
    return j / ENERGY_HARTREE_TO_J


def time(s):
    return s / TIME_AU_TO_S


def temperature(k):
    return k / TEMPERATURE_AU_TO_K


def pressure(pa):
    return pa / PRESSURE_AU_TO_PA


def inv_pressure(per_pa):
    return per_pa * PRESSURE_AU_TO_PA
