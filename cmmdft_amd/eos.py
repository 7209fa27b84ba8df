"""Bulk equations of state used by the WDA functional and for mu(p).

Mirrors the interface of cmmdft/eos.py (``EquationOfState`` and the mean-field,
Carnahan-Starling, MBWR and sum classes, eos.py:22-577): per-particle excess
free energy ``excess_free_energy_particle`` and its density derivative, the
derived per-volume quantities, ``compute_chempot``/``compute_pressure`` and
the rough-grid + brentq root finding of ``calculate_mu`` (eos.py:122-157,
209-215).  This host code works on scalars/small arrays (state-point setup).
On the grid, the same formulas run inside the HIP functional kernel; each
class therefore also packs itself into the device EOS description
(``device_terms``).
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import brentq

from . import units as U

EOS_MFA, EOS_CS, EOS_MBWR = 0, 1, 2


class EquationOfState:
    name = 'EOS'

    def __init__(self, mass):
        self.mass = mass
        self.temperature = None

    def set_temperature(self, temperature, **kwargs):
        self.temperature = temperature
        self.wvl = U.planck / np.sqrt(2 * np.pi * self.mass * U.boltzmann * temperature)

    # per particle ----------------------------------------------------------
    def excess_free_energy_particle(self, rho):
        raise NotImplementedError

    def derivative_excess_free_energy_particle(self, rho):
        raise NotImplementedError

    # per volume (eos.py:47-63) --------------------------------------------
    def excess_free_energy_volume(self, rho):
        return rho * self.excess_free_energy_particle(rho)

    def derivative_excess_free_energy_volume(self, rho):
        return rho * self.derivative_excess_free_energy_particle(rho) + self.excess_free_energy_particle(rho)

    def compute_chempot(self, rho):
        kT = U.boltzmann * self.temperature
        return kT * np.log(self.wvl ** 3 * rho) + self.derivative_excess_free_energy_volume(rho)

    def compute_excess_chempot(self, rho):
        return self.derivative_excess_free_energy_volume(rho)

    def compute_pressure(self, rho):
        kT = U.boltzmann * self.temperature
        return kT * rho + rho ** 2 * self.derivative_excess_free_energy_particle(rho)

    def get_rough_density_grid(self, npoints):
        return np.logspace(-10, 0, npoints) / U.angstrom ** 3

    def _bracket_solve(self, target, f_grid, fun, grid):
        sols = []
        inside = np.nonzero((f_grid[:-1] <= target) & (target <= f_grid[1:]))[0]
        for j in inside:
            sols.append(brentq(fun, grid[j], grid[j + 1]))
        return sorted(sols)

    def solve_densities_from_pressures(self, pressures, n_rough_gridpoints=10000):
        grid = self.get_rough_density_grid(n_rough_gridpoints)
        kT = U.boltzmann * self.temperature
        pgrid = kT * grid + grid ** 2 * self.derivative_excess_free_energy_particle(grid)
        out = np.full((len(pressures), 3), np.nan)
        for i, p in enumerate(pressures):
            sols = self._bracket_solve(p, pgrid, lambda r: kT * r + r ** 2 * self.derivative_excess_free_energy_particle(r) - p, grid)
            if len(sols) > 3:
                raise ValueError('Solving densities from EOS only supports max 3 branches, found %i' % len(sols))
            out[i, :len(sols)] = sols
        return out

    def solve_densities_from_chempots(self, chempots, n_rough_gridpoints=1000):
        grid = self.get_rough_density_grid(n_rough_gridpoints)
        kT = U.boltzmann * self.temperature
        mgrid = (kT * np.log(self.wvl ** 3 * grid) + self.excess_free_energy_particle(grid)
                 + grid * self.derivative_excess_free_energy_particle(grid))
        out = np.full((len(chempots), 2), np.nan)
        for i, mu in enumerate(chempots):
            f = lambda r: (kT * np.log(self.wvl ** 3 * r) + self.excess_free_energy_particle(r)
                           + r * self.derivative_excess_free_energy_particle(r) - mu)
            sols = self._bracket_solve(mu, mgrid, f, grid)
            out[i, :len(sols[:2])] = sols[:2]
        return out

    def calculate_mu(self, temp, pressure):
        self.set_temperature(temp)
        rho = self.solve_densities_from_pressures([pressure])[0][0]
        return self.compute_chempot(rho)

    def calculate_pressure(self, temp, chempot):
        self.set_temperature(temp)
        rho = self.solve_densities_from_chempots([chempot])[0][0]
        return self.compute_pressure(rho)

    def device_terms(self):
        """[(kind, factor, params[16])] for the HIP WDA kernel."""
        raise NotImplementedError


class MFAEOS(EquationOfState):
    """Mean-field EOS, a_ex = a rho with a = -16/9 pi eps sigma^3 (eos.py:523-546)."""
    name = 'MFA'

    def __init__(self, mass, sigma=None, epsilon=None, a=None):
        super().__init__(mass)
        if a is not None:
            self.a = a
        elif sigma is not None and epsilon is not None:
            self.a = -16 / 9 * np.pi * epsilon * sigma ** 3
        else:
            raise IOError('Either argument a should be defined or BOTH epsilon and sigma!')

    def excess_free_energy_particle(self, rho):
        return self.a * rho

    def derivative_excess_free_energy_particle(self, rho):
        return self.a + 0.0 * rho

    def device_terms(self):
        p = [0.0] * 16
        p[0] = self.a
        return [(EOS_MFA, 1.0, p)]


class CarnahanStarlingEOS(EquationOfState):
    """Carnahan-Starling hard-sphere EOS with eta = 4 pi R^3 / 3 (eos.py:471-520)."""
    name = 'CS'

    def __init__(self, mass, Rhs):
        super().__init__(mass)
        self.Rfun = None
        if callable(Rhs):
            self.Rfun, self.R, self.eta = Rhs, None, None
        elif isinstance(Rhs, float):
            self.R = Rhs
            self.eta = 4 * np.pi * Rhs ** 3 / 3
        else:
            raise TypeError('Rhs should be a float or a callable')

    def set_temperature(self, temperature, **kwargs):
        super().set_temperature(temperature)
        if self.Rfun is not None:
            self.R = self.Rfun(temperature, **kwargs)
            self.eta = 4 * np.pi * self.R ** 3 / 3

    def get_rough_density_grid(self, npoints):
        end = np.log(U.angstrom ** 3 / self.eta) / np.log(10) - 0.01
        return np.logspace(-10, end, npoints) / U.angstrom ** 3

    def excess_free_energy_particle(self, rho):
        kT = U.boltzmann * self.temperature
        x = self.eta * rho
        return kT * (4 * x - 3 * x ** 2) / (1 - x) ** 2

    def derivative_excess_free_energy_particle(self, rho):
        kT = U.boltzmann * self.temperature
        x = self.eta * rho
        return 2 * kT * self.eta * (2 - x) / (1 - x) ** 3

    def device_terms(self):
        p = [0.0] * 16
        p[0] = self.eta
        p[1] = U.boltzmann * self.temperature
        return [(EOS_CS, 1.0, p)]


_MBWR_X = np.array([
    0.8623085097507421, 2.976218765822098, -8.402230115796038, 0.1054136629203555, -0.8564583828174598,
    1.582759470107601, 0.7639421948305453, 1.753173414312048, 2.798291772190376e+3, -4.8394220260857657e-2,
    0.9963265197721935, -3.698000291272493e+1, 2.084012299434647e+1, 8.305402124717285e+1, -9.574799715203068e+2,
    -1.477746229234994e+2, 6.398607852471505e+1, 1.603993673294834e+1, 6.805916615864377e+1, -2.791293578795945e+3,
    -6.245128304568454, -8.116836104958410e+3, 1.488735559561229e+1, -1.059346754655084e+4, -1.131607632802822e+2,
    -8.867771540418822e+3, -3.986982844450543e+1, -4.689270299917261e+3, 2.593535277438717e+2, -2.694523589434903e+3,
    -7.218487631550215e+2, 1.721802063863269e+2])


class ModifiedBenedictWebbRubinEOS(EquationOfState):
    """LJ-fluid MBWR EOS of Johnson, Zollweg and Gubbins (Mol. Phys. 78, 591
    (1993), table 10 regression parameters as used at eos.py:311-345)."""
    name = 'MBWR'
    gamma = 3.0

    def __init__(self, mass, sigma, epsilon, logging=False):
        super().__init__(mass)
        self.sigma = sigma
        self.epsilon = epsilon

    def set_temperature(self, temperature, **kwargs):
        super().set_temperature(temperature)
        x = np.concatenate([[0.0], _MBWR_X])
        Tr = U.boltzmann * temperature / self.epsilon
        self.a = [x[1] * Tr + x[2] * np.sqrt(Tr) + x[3] + x[4] / Tr + x[5] / Tr ** 2,
                  x[6] * Tr + x[7] + x[8] / Tr + x[9] / Tr ** 2,
                  x[10] * Tr + x[11] + x[12] / Tr,
                  x[13],
                  x[14] / Tr + x[15] / Tr ** 2,
                  x[16] / Tr,
                  x[17] / Tr + x[18] / Tr ** 2,
                  x[19] / Tr ** 2]
        self.b = [x[20] / Tr ** 2 + x[21] / Tr ** 3,
                  x[22] / Tr ** 2 + x[23] / Tr ** 4,
                  x[24] / Tr ** 2 + x[25] / Tr ** 3,
                  x[26] / Tr ** 2 + x[27] / Tr ** 4,
                  x[28] / Tr ** 2 + x[29] / Tr ** 3,
                  x[30] / Tr ** 2 + x[31] / Tr ** 3 + x[32] / Tr ** 4]

    def get_rough_density_grid(self, npoints):
        return np.logspace(-10, 0, npoints) * 1.5 / self.sigma ** 3

    def excess_free_energy_particle(self, rho):
        rr = rho * self.sigma ** 3
        F = np.exp(-self.gamma * rr ** 2)
        ig = 1.0 / (2 * self.gamma)
        A = sum(ai / (i + 1) * rr ** (i + 1) for i, ai in enumerate(self.a))
        G = ig * (1 - F)
        A = A + self.b[0] * G
        for t in range(1, 6):
            G = -ig * (F * rr ** (2 * t) - 2 * t * G)
            A = A + self.b[t] * G
        return A * self.epsilon

    def derivative_excess_free_energy_particle(self, rho):
        rr = rho * self.sigma ** 3
        s3 = self.sigma ** 3
        F = np.exp(-self.gamma * rr ** 2)
        dA = sum(ai * rr ** i * s3 for i, ai in enumerate(self.a))
        dA = dA + sum(bt * s3 * rr ** (2 * t + 1) * F for t, bt in enumerate(self.b))
        return dA * self.epsilon

    def device_terms(self):
        p = [float(v) for v in self.a] + [float(v) for v in self.b] + [self.sigma ** 3, self.epsilon]
        return [(EOS_MBWR, 1.0, p)]


class SumOfEOS(EquationOfState):
    """Weighted sum of EOS terms (eos.py:226-265)."""

    def __init__(self, mass, list_eos, factors=None):
        assert isinstance(list_eos, list) and len(list_eos) > 1
        super().__init__(mass)
        self.list_eos = list_eos
        self.factors = list(factors) if factors is not None else [1.0] * len(list_eos)
        assert len(self.factors) == len(list_eos)

    def set_temperature(self, temperature, **kwargs):
        for e in self.list_eos:
            e.set_temperature(temperature)
        super().set_temperature(temperature)

    def excess_free_energy_particle(self, rho):
        return sum(f * e.excess_free_energy_particle(rho) for e, f in zip(self.list_eos, self.factors)) + 0.0 * rho

    def derivative_excess_free_energy_particle(self, rho):
        return sum(f * e.derivative_excess_free_energy_particle(rho) for e, f in zip(self.list_eos, self.factors)) + 0.0 * rho

    def excess_free_energy_volume(self, rho):
        return sum(f * e.excess_free_energy_volume(rho) for e, f in zip(self.list_eos, self.factors)) + 0.0 * rho

    def derivative_excess_free_energy_volume(self, rho):
        return sum(f * e.derivative_excess_free_energy_volume(rho) for e, f in zip(self.list_eos, self.factors)) + 0.0 * rho

    def device_terms(self):
        out = []
        for e, f in zip(self.list_eos, self.factors):
            for kind, fac, p in e.device_terms():
                out.append((kind, fac * f, p))
        if len(out) > 4:
            raise NotImplementedError('the device WDA supports at most 4 EOS terms')
        return out
