"""Program: the caller of the hot path (cmmdft/program.py:21-404).

Keeps the reference's set-up sequence (set_system -> set_grid ->
init_free_energy -> add parts -> solve), its initial-density rules
(program.py:219-306), solver cascade (:315-329) and the rho_<suffix>.npy
solution cache (:378-404, file names from tools.get_file_suffix,
tools.py:64-79).
"""
from __future__ import annotations

import os
import warnings
from pathlib import Path

import numpy as np

from . import units as U
from ._device import torch_mod, as_device
from .functionals import FreeEnergy
from .system import System, Grid
from .solver import Solver, NoSolutionError

__all__ = ['Program', 'get_file_suffix']


def get_file_suffix(chempot, temp):
    """'<mu in kJ/mol>kJmol_<T>K' with '#7.5f' formatting (tools.py:64-79)."""
    if hasattr(chempot, '__iter__'):
        s = ''.join(f'{mu / U.kjmol:#7.5f}kJmol_' for mu in chempot)
        return s + f'{temp:#7.5f}K'
    return f'{chempot / U.kjmol:#7.5f}kJmol_{temp:#7.5f}K'


class Program:
    def __init__(self, prefix='', hostname='', guestname='', ff_suffix='', funct_suffix='', grid_suffix='',
                 suffix='', overwrite=False, write_files=True):
        self.name_dict = {'prefix': prefix, 'hostname': hostname, 'guestname': guestname, 'ff_suffix': ff_suffix,
                          'funct_suffix': funct_suffix, 'grid_suffix': grid_suffix, 'suffix': suffix}
        self.workdir = Path(prefix) / hostname / guestname / ff_suffix / funct_suffix / grid_suffix / suffix
        self.write_files = write_files
        if write_files and not self.workdir.is_dir():
            self.workdir.mkdir(parents=True, exist_ok=True)
        self.overwrite = overwrite
        self.rho_fn = None
        self.system = None
        self.grid = None
        self.fener = None

    def copy(self):
        """A Program sharing this one's system, grid, free energy and solver
        (program.py:72-96)."""
        new = Program(prefix=self.name_dict['prefix'], hostname=self.name_dict['hostname'],
                      guestname=self.name_dict['guestname'], ff_suffix=self.name_dict['ff_suffix'],
                      funct_suffix=self.name_dict['funct_suffix'], grid_suffix=self.name_dict['grid_suffix'],
                      suffix=self.name_dict['suffix'], overwrite=self.overwrite, write_files=self.write_files)
        new.workdir = self.workdir
        new.rho_fn = self.rho_fn
        new.pars_fn = getattr(self, 'pars_fn', None)
        new.system = self.system
        new.grid = self.grid
        new.fener = self.fener
        if hasattr(self, 'solver'):
            new.solver = self.solver
        return new

    def set_system(self, host, guest):
        self.system = System(host, guest)

    def calc_distance(self, rewrite=False):
        """Distance from every grid point to the nearest host atom
        (program.py:145-177; minimum image in fractional coordinates, as
        yaff's Cell.mic), on the device.  Kept as ``self.dis``; with file
        output on it is cached as <prefix>/<hostname>/<grid_suffix>/distances.npy."""
        dist_file = Path(self.name_dict['prefix']) / self.name_dict['hostname'] / self.name_dict['grid_suffix'] / \
            'distances.npy'
        if self.write_files:
            dist_file.parent.mkdir(parents=True, exist_ok=True)
            if rewrite and dist_file.is_file():
                dist_file.unlink()
            if dist_file.is_file():
                self.dis = np.load(dist_file)
                return
        torch = torch_mod()
        from . import _lib
        pos = as_device(np.asarray(self.system.host.mol.pos, dtype=np.float64).reshape(-1, 3))
        out = torch.empty(tuple(self.grid.npoints), dtype=torch.float64, device='cuda')
        _lib.min_distance(self.grid.npoints, np.asarray(self.grid.cell.rvecs, dtype=np.float64), pos, pos.shape[0],
                          out)
        self.dis = out.cpu().numpy()
        if self.write_files:
            np.save(dist_file, self.dis)

    def calc_regions(self, energy_cutoff=0.55, range_cutoff=3.4 * U.angstrom, mof_cutoff=5):
        """Three region masks of the framework (program.py:179-217): near the
        host and/or energetically favoured sites, the framework itself
        (V > mof_cutoff kT), and the empty pore space.  Returns
        (mask_site, mask_mof, mask_empty)."""
        self.calc_distance()
        range_mask = self.dis < range_cutoff
        ep = self._extpot()
        if ep is None:
            raise ValueError('The regions of a nanoporous material can only be calculated if an external potential '
                             'is defined')
        epot_data = np.asarray(ep.potential)
        crit = np.amin(epot_data) - energy_cutoff * np.amin(epot_data)
        energy_mask = epot_data < crit
        self.r_mask = range_mask
        self.e_mask = energy_mask
        self.mask_mof = epot_data > mof_cutoff * U.boltzmann * self.fener.temperature
        self.mask_site = (energy_mask + range_mask) * (~self.mask_mof)
        self.mask_empty = (~energy_mask) * (~range_mask) * (~self.mask_mof)
        return self.mask_site, self.mask_mof, self.mask_empty

    def set_grid(self, npoints=None, spacing=0.5 * U.angstrom):
        assert self.system is not None, "Host and guest must first be set using 'set_system'"
        self.grid = Grid(self.system.host.cell, npoints=npoints, spacing=spacing)

    def init_free_energy(self, temperature):
        assert self.system is not None and self.grid is not None
        self.fener = FreeEnergy(self.grid, self.system, temperature, workdir=self.workdir,
                                overwrite=self.overwrite, name_dict=self.name_dict, write_files=self.write_files)

    def set_temperature(self, temperature):
        self.fener.set_temperature(temperature)

    def _extpot(self):
        for name in self.fener.part_names:
            if 'ExtPot' in name:
                return self.fener.part_dict[name]
        return None

    def _set_initial_density(self, Ninit=None, chempot=None, rewrite=False, temperature=None, silent=False):
        """Initial density rules of program.py:219-306, on the device."""
        torch = torch_mod()
        if self.rho_fn is not None and os.path.isfile(self.rho_fn) and not self.overwrite and not rewrite:
            self.rho0 = as_device(np.load(self.rho_fn))
            return self.rho0
        ep = self._extpot()
        if Ninit is not None:
            if isinstance(Ninit, (str, Path)):
                if not Path(Ninit).is_file():
                    raise FileNotFoundError('File %s for setting initial density not found' % Ninit)
                self.rho0 = as_device(np.load(Ninit))
            elif isinstance(Ninit, float):
                if ep is not None:
                    V = ep.potential_dev
                    self.rho0 = Ninit * torch.exp(-torch.clamp(V, min=0.0) / U.boltzmann / temperature)
                else:
                    self.rho0 = torch.full(tuple(self.grid.npoints), Ninit, dtype=torch.float64, device='cuda')
            else:
                assert tuple(np.shape(Ninit)) == tuple(self.grid.npoints), 'Ninit must have the same shape as the grid'
                self.rho0 = as_device(Ninit)
        else:
            V = ep.potential_dev if ep is not None else torch.zeros(tuple(self.grid.npoints), dtype=torch.float64,
                                                                    device='cuda')
            self.rho0 = torch.exp(self.fener.beta * (chempot - V)) / self.fener.wavelength ** 3
        return self.rho0

    def set_solver(self, solver):
        assert isinstance(solver, Solver), 'solver is not an instance of Solver, aborting!'
        self.solver = solver

    def cascade_solver(self, solvers, chempot, **kwargs):
        """Try the solvers in turn until one converges (program.py:315-329).
        Returns that solve's (N, rho, converged); when none converges the
        reference only logs a warning: here a RuntimeWarning, and None."""
        for solver in solvers:
            self.set_solver(solver)
            try:
                N, rho, converged = self.solve(chempot, **kwargs)
                if converged:
                    return N, rho, converged
            except NoSolutionError:
                continue
        warnings.warn('All solvers failed, at %7.5fkJmol %7.5fK.' % (chempot / U.kjmol, self.fener.temperature),
                      RuntimeWarning)
        return None

    def solve(self, chempot, Ninit=None, rewrite=False, energy_tracking=True, silent=False, continue_solving=False):
        """Equilibrium density at ``chempot`` (program.py:331-404).  Returns
        (N, rho (numpy), converged).  With file output on, ``energy_tracking``
        writes the grand potential and its parts at every step to
        convergence_<suffix>.txt in the work directory."""
        self.file_suffix = get_file_suffix(chempot, self.fener.temperature)
        if energy_tracking and self.write_files:
            self.fener.init_tracking(os.path.join(self.workdir, 'convergence_%s.txt' % self.file_suffix))
        self.rho_fn = os.path.join(self.workdir, 'rho_%s.npy' % self.file_suffix) if self.write_files else None
        if self.rho_fn and os.path.isfile(self.rho_fn) and not self.overwrite and not rewrite and not continue_solving:
            rho = np.load(self.rho_fn)
            return self.grid.integrate(rho), rho, True
        rho0 = self._set_initial_density(Ninit=Ninit, chempot=chempot, rewrite=rewrite,
                                         temperature=self.fener.temperature, silent=silent)
        N, rho, converged = self.solver.solve(chempot, rho0.clone(), 2)
        rho_np = rho.cpu().numpy() if hasattr(rho, 'cpu') else rho
        if self.write_files:
            if self.solver.track_history:
                fn = self.workdir / ('solving_history_%s.csv' % self.file_suffix)
                np.savetxt(fn, self.solver.history[:self.solver.curr_step + 1, :], delimiter=',',
                           header=self.solver.history_header)
            np.save(self.rho_fn, rho_np)
        return N, rho_np, converged
