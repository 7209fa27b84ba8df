"""Properties of converged densities (cmmdft/calculator.py:22-492): loadings
(absolute, masked, excess), the Henry coefficient and the helium void
fraction from the external potential, isotherm files (CSV, AIF), and the
free-energy contributions and grand potential of a stored density.

Densities are read from the solver's ``rho_<mu>kJmol_<T>K.npy`` files in the
program's work directory (``Program.solve`` writes them), moved to the
device, and integrated there (``cmm_field_reduce``); the helium potential
is the device LJ grid sum.

Not built: ``loading_MWBR_unreliable`` (its WDA construction passes the radius
and the grid in swapped order, calculator.py:105, and fails in the
reference), and the diffusion-path analysis (collective variables, density
projections, free-energy paths, diffusion coefficients, contribution
approximations; calculator.py:494-1269).
"""
from __future__ import annotations

import datetime
import getpass
import os
import re
from pathlib import Path

import numpy as np

from . import units as U
from ._device import torch_mod, as_device
from ._lib import field_reduce
from .eos import EquationOfState
from .program import get_file_suffix

__all__ = ['Calculator', 'convert_units', 'write_aif']

VERSION = '1.0'
HE_SIGMA = 2.576 * U.angstrom       # Hirschfelder, Curtiss and Bird (calculator.py:160)
HE_EPSILON = 10.22 * U.boltzmann


class convert_units:
    """Adsorption-unit conversion factors (tools.py:368-399): masses of one
    guest and of the host cell (atomic units), host cell volume."""

    def __init__(self, mass_guest, mass_host, volume_host):
        rho_stp = (mass_guest / U.amu) * 1e-3 / 22.414
        rho_host = (mass_host / U.gram) / (volume_host / U.centimeter ** 3)
        self.output_dict = {'wt%': mass_host / mass_guest / 100, 'mg/g': mass_host / mass_guest / 1000,
                            'cm3/cm3': mass_guest * rho_host / mass_host / rho_stp, 'mol/mol': 1,
                            'mol/g': mass_host / U.amu, 'mol/kg': mass_host / U.amu / 1000, 'au/uc': 1}
        self.input_dict = dict(self.output_dict)

    def conversion_factor(self, input='mol/mol', output='mol/mol'):
        units = ['wt%', 'cm3/cm3', 'mol/mol', 'mol/g', 'mol/kg', 'au/uc', 'mg/g']
        assert input in units and output in units, 'unknown adsorption unit'
        return self.input_dict[input] / self.output_dict[output]


def write_aif(fn, block, pairs, loop_prefix, loop_keys, columns):
    """AIF text in the layout the reference writes (tools.py:401-456):
    ``data_<block>``, one ``key value`` line per pair, then a loop of
    columns (each row's values followed by a space)."""
    with open(fn, 'w') as fh:
        fh.write('data_%s\n' % block)
        for key, value in pairs:
            fh.write('%s %s\n' % (key, value))
        fh.write('\nloop_\n')
        for key in loop_keys:
            fh.write('%s%s\n' % (loop_prefix, key))
        for i in range(len(columns[0])):
            fh.write(''.join(col[i] + ' ' for col in columns) + '\n')


class Calculator:
    def __init__(self, program, label=None):
        self.program = program
        self.name_dict = program.name_dict
        self.workdir = Path(program.workdir)
        self.grid = program.grid
        self.fener = program.fener
        self.host = program.system.host
        self.guest = program.system.guest
        self.label = label

    # -- densities on file -------------------------------------------------
    def _rho_file(self, temp, chempot):
        fn = self.workdir / ('rho_%s.npy' % get_file_suffix(chempot, temp))
        assert fn.is_file(), 'No file found at %s' % fn
        return fn

    def _rho(self, temp, chempot):
        return as_device(np.load(self._rho_file(temp, chempot)))

    def _integrate(self, field, op='sum', mask=None, beta=1.0, lam3=1.0):
        m = None if mask is None else as_device(np.asarray(mask)).to(torch_mod().uint8)
        return field_reduce(field, op, beta=beta, lam3=lam3, mask=m) * self.grid.dr

    def density_statistics(self, temp, chempot, mask=None):
        """(mean, min, max, std) of the stored density (calculator.py:37-58)."""
        rho = self._rho(temp, chempot)
        n = rho.numel()
        mean = field_reduce(rho, 'sum') / n
        var = max(field_reduce(rho, 'sumsq') / n - mean * mean, 0.0)
        return mean, field_reduce(rho, 'min'), field_reduce(rho, 'max'), float(np.sqrt(var))

    def loading(self, temp, chempot, mask=None):
        """Guests per cell: the integral of the stored density, inside
        ``mask`` (a boolean array of the grid shape) when given
        (calculator.py:60-88)."""
        return self._integrate(self._rho(temp, chempot), mask=mask)

    def return_loading(self, temp, chempots, excess=False, eos=None, He_frac=None):
        """Loadings for a list of chemical potentials (NaN where no density
        is stored); excess: minus the bulk density times the helium-accessible
        volume (calculator.py:114-138)."""
        loads = np.zeros(len(chempots))
        for i, mu in enumerate(chempots):
            try:
                loads[i] = self.loading(temp, mu)
            except AssertionError:
                loads[i] = np.nan
        if excess:
            assert eos is not None, 'Must provide an equation of state object when calculating excess uptake'
            if He_frac is None:
                He_frac = self.get_helium_fraction(temp)
            if isinstance(eos, EquationOfState):
                eos.set_temperature(temp)
                bulk = np.nanmin(eos.solve_densities_from_chempots(chempots), axis=1)
            else:
                bulk = np.array([eos.calculate_rho(temp, mu) for mu in chempots])
            loads -= bulk * He_frac * self.host.cell.volume
        return loads

    def get_chemical_potential(self, temperature):
        """Chemical potentials of the densities stored for ``temperature``,
        ascending (calculator.py:140-150)."""
        num = re.compile(r'[-+]? (?: (?: \d* \. \d+ ) | (?: \d+ \.? ) )(?: [Ee] [+-]? \d+ ) ?', re.VERBOSE)
        tail = '%#7.5fK.npy' % temperature
        names = [f.name for f in self.workdir.iterdir() if f.name.startswith('rho') and f.name.endswith(tail)]
        return np.sort(np.array([float(num.findall(f)[0]) for f in names]) * U.kjmol)

    # -- potentials --------------------------------------------------------
    def _extpot(self):
        for name in self.fener.part_names:
            if 'ExtPot' in name:
                return self.fener.part_dict[name]
        return None

    def get_helium_fraction(self, temperature):
        """Helium void fraction: (1/V) int min(exp(-beta V_He), 1) with the
        He-host LJ potential (calculator.py:152-167); the potential is cached
        in ``workdir/ExtPots/He_potential.npy``."""
        from .extpot import get_external_potential
        fn = self.workdir / 'ExtPots' / 'He_potential.npy'
        if fn.is_file():
            V = as_device(np.load(fn))
        else:
            V = get_external_potential(self.grid, self.host.ff, HE_SIGMA, HE_EPSILON, self.host.mol,
                                       cutoff=12 * U.angstrom, as_tensor=True)
            if self.program.write_files:
                os.makedirs(fn.parent, exist_ok=True)
                np.save(fn, V.cpu().numpy())
        vol = self._integrate(V, 'boltzmann_clip', beta=1.0 / (U.boltzmann * temperature))
        return vol / self.host.cell.volume

    def get_Henry_Coefficient(self, temperature):
        """K_H = int exp(-beta V) / (N_A k_B T V_cell) from the external
        potential part (calculator.py:169-184)."""
        ep = self._extpot()
        if ep is None:
            raise ValueError('No external potential found in the functional')
        integral = self._integrate(ep.potential_dev, 'boltzmann', beta=1.0 / (U.boltzmann * temperature))
        return 1 / U.avogadro / U.boltzmann / temperature / self.host.cell.volume * integral

    # -- isotherm files ----------------------------------------------------
    def save_loadings(self, temperature, chempots=None, pressure=False, excess=False, eos=None, fn=None):
        """CSV of loadings against chemical potential or pressure
        (calculator.py:186-228)."""
        if chempots is None:
            chempots = self.get_chemical_potential(temperature)
        data = np.zeros((2, len(chempots)))
        if pressure:
            header = 'pressures [au], loadings [molecules/uc]'
            data[0] = np.array([eos.calculate_pressure(temperature, mu) for mu in chempots])
        else:
            header = 'chempot [au], loadings [molecules/uc]'
            data[0] = chempots
        data[1] = self.return_loading(temperature, chempots, excess=excess, eos=eos)
        if fn is None:
            suffix = '_vs_P' if pressure else ''
            prefix = 'excess_' if excess else ''
            fn = self.workdir / ('%sloads_%#3.0fK%s.csv' % (prefix, temperature, suffix))
        np.savetxt(Path(fn), data.T, delimiter=',', header=header, comments='')

    def save_loadings_AIF(self, temp, chempots=None, pressures=None, eos=None, input_fn=None, user=None, excess=False,
                          loading_unit='au/uc', fn=None, He_frac=None):
        """Isotherm as an adsorption information file (calculator.py:230-341)."""
        if user is None:
            user = getpass.getuser()
        inputs = [str(f) for f in (input_fn, getattr(self.host, 'chk', None), getattr(self.host, 'par', None),
                                   getattr(self.guest, 'chk', None), getattr(self.guest, 'par', None))
                  if f is not None]
        pairs = [('_audit_aif_version', '6acf6ef'), ('_exptl_operator', user),
                 ('_adsnt_material_id', self.host.name), ('_exptl_adsorptive', self.guest.name),
                 ('_exptl_temperature', '%0.3fK' % temp), ('_exptl_method', 'simulation'),
                 ('_exptl_isotherm_type', 'excess' if excess else 'absolute'),
                 ('_simltn_date', datetime.datetime.now().strftime('%Y-%m-%d')),
                 ('_simltn_code', 'CmmDFT-%s' % VERSION), ('_simltn_sampling', 'cDFT'),
                 ('_simltn_input_files', inputs)]
        ffs = self.program.name_dict['ff_suffix']
        if hasattr(self.host, 'mol'):
            parts = ffs.split('_')
            pairs += [('_simltn_forcefield_adsorptive', parts[0]),
                      ('_simltn_forcefield_adsorbent', parts[1] if len(parts) > 1 else '')]
        else:
            pairs += [('_simltn_forcefield_adsorbent', ffs)]
        pairs += [('_units_temperature', 'K'), ('_units_energy', 'kJ/mol'), ('_units_loading', loading_unit),
                  ('_units_pressure', 'bar'), ('_units_mass', 'amu')]
        if pressures is None:
            assert eos is not None and chempots is not None, \
                'Must provide chemical potentials and an equation of state to compute pressures'
            pressures = np.array([eos.calculate_pressure(temp, mu) for mu in chempots])
        if chempots is None:
            assert eos is not None, 'Must provide an equation of state to compute chemical potentials'
            chempots = np.array([eos.calculate_mu(temp, p) for p in pressures])
        chempots = np.asarray(chempots, dtype=np.float64)
        pressures = np.asarray(pressures, dtype=np.float64)
        fug = np.exp(self.fener.beta * chempots) / self.fener.beta / self.fener.wavelength ** 3
        absolute = self.return_loading(temp, chempots, excess=False)
        masses = getattr(getattr(self.host, 'mol', None), 'masses', None)
        host_mass = float(np.sum(masses)) if masses is not None else 1.0
        factor = convert_units(self.guest.mass, host_mass, self.host.cell.volume).conversion_factor('au/uc',
                                                                                                  loading_unit)
        absolute = absolute * factor
        if fn is None:
            fn = self.workdir / ('adsorption_%0.2fK.aif' % temp)
        cols = [['%.5E' % v for v in pressures / U.bar], ['%.5E' % v for v in fug / U.bar],
                ['%.5E' % v for v in chempots / U.kjmol], ['%.5E' % v for v in absolute]]
        write_aif(Path(fn), 'CmmDFT2aif', pairs, '_adsorp_',
                  ['pressure', 'fugacity', 'chemicalpotential', 'amount_absolute'], cols)

    # -- free energies -----------------------------------------------------
    def free_energy_contrib(self, temp, chempot, partname, over_loading=False, local=False, fn=None, rho=None):
        """One free-energy contribution of a stored (or given) density
        (calculator.py:343-406): 'fid' / 'fideal', or a part of the free
        energy by name."""
        if rho is None:
            rho = as_device(np.load(fn if fn is not None else self._rho_file(temp, chempot)))
        else:
            assert isinstance(rho, np.ndarray), 'The density must be a numpy array'
            rho = as_device(rho)
        N = self._integrate(rho) if over_loading else None
        if partname.lower() in ('fid', 'fideal'):
            kT = U.boltzmann * temp
            lam3 = self.fener.wavelength ** 3
            if local:
                c = torch_mod().clamp(rho, min=1e-20)
                out = (kT * c * (torch_mod().log(c * lam3) - 1.0)).cpu().numpy()
                return out / N if over_loading else out
            val = kT * self._integrate(rho, 'fid', lam3=lam3)
            return val / N if over_loading else val
        assert partname in self.fener.part_names, \
            '%s not found in %s. The provided partname must be present in the fener object, or the ideal gas ' \
            'contribution ("fid" or "fideal")' % (partname, self.fener.part_names)
        if self.fener.temperature != temp and partname in ('HardSphere', 'WDA-V'):
            self.fener.set_temperature(temp)
        part = self.fener.part_dict[partname]
        # by keyword: the reference's positional call hands ``local`` to the
        # WDA part's ``wd`` argument (functionals.py:1176), which then
        # evaluates f_ex of the constant False, i.e. a zero contribution
        val = part.value(rho.reshape(self.grid.npoints), local=local)
        if local and hasattr(val, 'cpu'):
            val = val.cpu().numpy()
        return val / N if over_loading else val

    def free_energy(self, temp, chempot, local=False):
        """F = F_id + sum of the parts (calculator.py:408-431; as there, the
        ideal-gas term is the integrated one also when ``local``)."""
        value = self.free_energy_contrib(temp, chempot, 'fid')
        for part in self.fener.parts:
            value = value + self.free_energy_contrib(temp, chempot, part.name, local=local)
        return value

    def excess_free_energy(self, temp, chempot, local=False, fn=None):
        """Sum of the excess parts (every part but the external potential;
        the reference's excess_table is not defined, calculator.py:433-461)."""
        value = 0.0
        for part in self.fener.parts:
            if part.name != 'ExtPot':
                value = value + self.free_energy_contrib(temp, chempot, part.name, local=local, fn=fn)
        return value

    def grand_potential(self, temp, chempot, local=False):
        """Omega = F - mu N (calculator.py:463-492)."""
        value = self.free_energy(temp, chempot, local=local)
        if local:
            return value - chempot * np.load(self._rho_file(temp, chempot))
        return value - chempot * self.loading(temp, chempot)
