"""Pinning of the MBWR equation of state (reference cmmdft/eos.py:295-447,
Johnson, Zollweg & Gubbins, Mol. Phys. 78, 591 (1993)) in the oracle and in
the product, independently of both:

* the 32 regression coefficients equal the reference's literals
  (tests/golden/mbwr_ref.json, written from eos.py:313-345 by
  tests/golden/make_mbwr_ref.py);
* the mapping of the coefficients onto a_i(T*), b_i(T*) reproduces what the
  published equation fixes: its vapour-liquid critical point
  T*c = 1.313, rho*c = 0.310, P*c = 0.13 (quoted in the paper), and the
  Lennard-Jones second virial coefficient B2*(T*) from quadrature of the
  Mayer function (the low-density limit (P/(rho kT) - 1)/rho* = a1(T*)/T*);
* a planted sign flip or swap of neighbouring coefficients in the oracle's
  table fails at least one of these checks (so they pin every coefficient).
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.integrate as integ
import scipy.optimize as opt

from cmmdft_amd import units as U, workloads as W
from cmmdft_amd import eos as E
from oracle import cdft_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, 'golden', 'mbwr_ref.json')))

SIG, EPS = W.GUEST_SIGMA, W.GUEST_EPSILON
TEMPS = [0.8, 1.0, 1.5, 2.0, 3.0, 5.0, 10.0]


def lj_b2(Tr):
    """B2* = -2 pi int_0^inf (exp(-u(r)/T*) - 1) r^2 dr, u = 4 (r^-12 - r^-6)."""
    f = lambda r: (np.exp(-4.0 * (r ** -12 - r ** -6) / Tr) - 1.0) * r * r
    return -2 * np.pi * (integ.quad(f, 1e-6, 1.0, limit=200)[0] + integ.quad(f, 1.0, 50.0, limit=400)[0])


B2_REF = [lj_b2(t) for t in TEMPS]


def oracle_pressure(table=None):
    cls = O.MBWREOS
    if table is not None:
        cls = type('PlantedMBWR', (O.MBWREOS,), {'X': list(table)})
    e = cls(W.GUEST_MASS, SIG, EPS)

    def P(rr, Tr):
        e.set_temperature(Tr * EPS / U.boltzmann)
        rho = rr / SIG ** 3
        return (U.boltzmann * e.temperature * rho + rho ** 2 * e.dfex_particle(rho)) * SIG ** 3 / EPS
    return P


def product_pressure():
    e = E.ModifiedBenedictWebbRubinEOS(W.GUEST_MASS, SIG, EPS)

    def P(rr, Tr):
        e.set_temperature(Tr * EPS / U.boltzmann)
        return e.compute_pressure(rr / SIG ** 3) * SIG ** 3 / EPS
    return P


def critical_point(P):
    """(rho*, T*, P*) where dP/drho = d2P/drho2 = 0 (central differences)."""
    d1 = lambda r, t, h=1e-5: (P(r + h, t) - P(r - h, t)) / (2 * h)
    d2 = lambda r, t, h=1e-4: (P(r + h, t) - 2 * P(r, t) + P(r - h, t)) / h ** 2
    (r, t), info, ier, _ = opt.fsolve(lambda x: [d1(x[0], x[1]), d2(x[0], x[1])], [0.31, 1.313], full_output=True)
    return r, t, P(r, t)


def virial_errors(P, h=1e-6):
    """|(P/(rho kT) - 1)/rho* - B2*(T*)| at each test temperature."""
    return [abs((P(h, t) / (h * t) - 1.0) / h - b) for t, b in zip(TEMPS, B2_REF)]


def passes(P):
    r, t, p = critical_point(P)
    crit = abs(t - 1.313) <= 5e-4 and abs(r - 0.310) <= 5e-4 and abs(p - 0.13) <= 5e-3
    return crit and max(virial_errors(P)) <= 5e-3


def test_fixture_is_self_consistent():
    lits = REF['literals']
    assert len(lits) == 32
    assert hashlib.sha256('\n'.join(lits).encode()).hexdigest() == REF['sha256_literals']
    assert [float(s) for s in lits] == REF['values']


def test_coefficient_tables_equal_the_reference_literals():
    assert list(O.MBWREOS.X) == REF['values']
    assert [float(v) for v in E._MBWR_X] == REF['values']


@pytest.mark.parametrize('which', ['oracle', 'product'])
def test_critical_point_is_the_published_one(which):
    P = oracle_pressure() if which == 'oracle' else product_pressure()
    r, t, p = critical_point(P)
    assert abs(t - 1.313) <= 5e-4, t
    assert abs(r - 0.310) <= 5e-4, r
    assert abs(p - 0.13) <= 5e-3, p


@pytest.mark.parametrize('which', ['oracle', 'product'])
def test_low_density_limit_is_the_lj_second_virial_coefficient(which):
    P = oracle_pressure() if which == 'oracle' else product_pressure()
    # the regression reproduces B2* to ~2e-3 over 0.8 <= T* <= 10
    assert max(virial_errors(P)) <= 5e-3


def test_planted_coefficient_mistakes_fail_the_pins():
    assert passes(oracle_pressure())
    base = list(REF['values'])
    undetected = []
    for i in range(32):
        flipped = list(base)
        flipped[i] = -flipped[i]
        if passes(oracle_pressure(flipped)):
            undetected.append(('sign', i + 1))
        if i < 31:
            swapped = list(base)
            swapped[i], swapped[i + 1] = swapped[i + 1], swapped[i]
            if passes(oracle_pressure(swapped)):
                undetected.append(('swap', i + 1))
    assert not undetected, undetected
