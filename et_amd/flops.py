"""Algorithmic operation counts of eT's CCSD residual factorisation.

F_omega counts the floating-point operations of eT's own algorithm for one
residual (src/wavefunctions/ccsd/omega_ccsd.F90 and the routines it calls),
term by term; it is the work figure roofline fractions are quoted against,
independently of how this implementation schedules the contractions.
"""


def _p(x):
    return x * (x + 1) // 2


def _m(x):
    return x * (x - 1) // 2


def omega_terms(o, v, J):
    """Dict term -> FLOP for one CCSD residual (eT factorisation)."""
    ov = o * v
    return {
        'a2_ladder': 2 * (_p(o) * _p(v) ** 2 + _m(o) * _m(v) ** 2),     # omega_ccsd.F90:345-448
        'a2_g_vvvv': J * v * v * (v * v + 1),                          # get_packed('vv') / dsfrk
        'c2_d2': 8 * ov ** 3,                                          # omega_ccsd.F90:647-688
        'e2': 4 * o * o * v ** 3 + 4 * o ** 3 * v * v,                 # omega_ccsd.F90:783-859
        'b2': 4 * o ** 4 * v * v,                                      # omega_ccsd.F90:549-566
        'a1': 2 * J * ov ** 2 + 2 * J * o * v * v,                     # omega_doubles.F90:36-139
        'b1_c1': 2 * o ** 3 * v * v + 2 * ov ** 2,                     # omega_doubles.F90:141-250
        'integrals': 2 * J * (6 * ov ** 2 + o ** 4 + o ** 3 * v),      # eri_tool get('..')
    }


def omega_flops(o, v, J):
    return float(sum(omega_terms(o, v, J).values()))
