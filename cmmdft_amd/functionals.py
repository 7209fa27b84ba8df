"""Functionals of the grand potential with the reference's call signatures
(cmmdft/functionals.py), evaluated by the fused HIP pipeline.

Accepted ``krho`` forms (see ``density_from``):
  * numpy complex full spectrum in the reference convention
    ``Grid.fft(rho) = fftn(rho) P / N`` (functionals.py / system.py:310-311);
  * torch complex half spectrum (nx, ny, nz//2+1): the native convention, an
    unnormalised R2C spectrum without phase;
  * a real density (numpy or torch), the native solver's form.
Results come back as numpy when the input was numpy, as device tensors
otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import units as U
from ._device import torch_mod, as_device, like_input
from ._lib import CmmFunctional, S_PHI, S_FEX, S_N3MAX
from .system import NanoporousHost, SphericalLJGuest

__all__ = ['FreeEnergy', 'Functional', 'HardSphereFunctional', 'WDAVFunctional', 'ExternalPotential',
           'density_from', 'build_functional']

FMT_NAMES = ['FMT', 'aFMT', 'tFMT', 'atFMT', 'taFMT']
MFMT_NAMES = ['MFMT', 'aMFMT', 'tMFMT', 'atMFMT', 'taMFMT']
WBII_NAMES = ['WBII', 'aWBII', 'tWBII', 'atWBII', 'taWBII']


def hs_base_of(version):
    if version in MFMT_NAMES:
        return 1
    if version in FMT_NAMES:
        return 0
    if version in WBII_NAMES:
        return 2
    raise AssertionError('Unknown hard sphere functional version: %s' % version)


def build_functional(hs=None, wda=None, kT=None):
    """Pack hard-sphere and WDA parts into the device functional description."""
    f = CmmFunctional()
    f.hs_base = -1
    f.R = 1.0
    if hs is not None:
        f.hs_base = hs_base_of(hs.version)
        f.hs_antisym = 1 if 'a' in hs.version else 0
        f.hs_tensor = 1 if 't' in hs.version else 0
        f.R = hs.R
    if wda is not None:
        terms = wda.eos.device_terms()
        f.use_wda = 1
        f.n_eos = len(terms)
        for i, (kind, fac, par) in enumerate(terms):
            f.eos_kind[i] = kind
            f.eos_factor[i] = fac
            for k, v in enumerate(par):
                f.eos_par[i][k] = v
        f.D = wda.D
    f.kT = kT
    return f


def density_from(grid, krho):
    """Device real density from any accepted ``krho`` form."""
    torch = torch_mod()
    nzh = grid.npoints[2] // 2 + 1
    if isinstance(krho, torch.Tensor):
        if krho.is_complex():
            if tuple(krho.shape) == (grid.npoints[0], grid.npoints[1], nzh):
                return grid.irfft_native(krho.to(torch.complex128)) / grid.npts
            return grid.ifft(krho)
        return krho.to(device='cuda', dtype=torch.float64).reshape(grid.npoints).contiguous()
    a = np.asarray(krho)
    if np.iscomplexobj(a):
        return grid.ifft(as_device(a))
    return as_device(a).reshape(grid.npoints)


class Functional:
    def __init__(self):
        pass

    def copy(self, **kwargs):
        raise NotImplementedError

    def set_temperature(self, temperature, **kwargs):
        pass


class _PartEngine:
    """Runs the fused pipeline of one grid with a single part enabled."""

    def __init__(self, grid):
        self.grid = grid

    def run(self, func, rho, want_c1=True):
        torch = torch_mod()
        plan = self.grid.plan()
        plan.set_functional(func)
        plan.seed(rho)
        if want_c1:
            plan.derive()
            c1 = torch.empty_like(rho)
            plan.finish_c1(rho, None, 0, c1)
            return c1, plan.scalars()[0]
        plan.value()
        return None, plan.scalars()[0]


class HardSphereFunctional(Functional):
    """Hard-sphere FMT functional (functionals.py:379-618): Rosenfeld 'FMT',
    'MFMT', White-Bear II 'WBII', each with the optional 'a' prefix."""
    name = 'HardSphere'

    def __init__(self, grid, Rhs, version='atWBII'):
        assert version in FMT_NAMES + MFMT_NAMES + WBII_NAMES, 'Unknown hard sphere functional version: %s' % version
        self.temperature = None
        self.beta = None
        self.grid = grid
        self.R = Rhs
        self.version = version
        self._engine = _PartEngine(grid)

    def copy(self, grid=None):
        return type(self)(grid if grid is not None else self.grid.copy(), self.R, self.version)

    def set_temperature(self, temperature, Rhs, **kwargs):
        self.temperature = temperature
        self.beta = 1 / (U.boltzmann * temperature)
        self.R = Rhs
        self.krho = None

    def _func(self):
        return build_functional(hs=self, kT=1.0 / self.beta)

    def derive(self, krho):
        rho = density_from(self.grid, krho)
        c1, _ = self._engine.run(self._func(), rho)
        return like_input(c1.reshape(self.grid.npoints), krho)

    def value(self, krho, local=False):
        if local:
            raise NotImplementedError('local free-energy densities are not produced by the device path')
        rho = density_from(self.grid, krho)
        _, s = self._engine.run(self._func(), rho, want_c1=False)
        return s[S_PHI] * self.grid.dr / self.beta

    def weighted_densities_dev(self, krho):
        """n2, n3 (unclipped), nv2 (3), the WDA density and, for the tensor
        variant, n_t (xx, xy, xz, yy, yz, zz) as a (6 or 12, N) tensor."""
        torch = torch_mod()
        rho = density_from(self.grid, krho)
        plan = self.grid.plan()
        plan.set_functional(self._func())
        plan.seed(rho)
        out = torch.empty((plan.nfields,) + tuple(self.grid.npoints), dtype=torch.float64, device=rho.device)
        plan.weighted_densities(out)
        return out

    def get_n3(self, krho):                           # functionals.py:554-556
        n = self.weighted_densities_dev(krho)
        return like_input(n[1], krho)

    def set_density(self, krho):
        self.krho = krho
        n = self.weighted_densities_dev(krho)
        self.weighted_densities = n


class WDAVFunctional(Functional):
    """Weighted-density approximation with the excess free energy per volume
    of an EOS; weight = normalised ball of diameter D = 2 R
    (functionals.py:1116-1184)."""
    name = 'WDA-V'

    def __init__(self, grid, Rhs, eos):
        self.temperature = None
        self.grid = grid
        self.eos = eos
        self.R = Rhs
        self._engine = _PartEngine(grid)

    def copy(self, grid=None):
        return type(self)(grid if grid is not None else self.grid.copy(), self.R, self.eos)

    def set_temperature(self, temperature, Rhs, **kwargs):
        self.temperature = temperature
        self.eos.set_temperature(temperature, **kwargs)
        self.R = Rhs
        self.D = 2 * self.R

    def _func(self):
        return build_functional(wda=self, kT=U.boltzmann * self.temperature)

    def _get_weighted_density(self, krho):
        torch = torch_mod()
        rho = density_from(self.grid, krho)
        plan = self.grid.plan()
        f = build_functional(wda=self, kT=U.boltzmann * self.temperature)
        f.R = self.R
        plan.set_functional(f)
        plan.seed(rho)
        out = torch.empty((plan.nfields,) + tuple(self.grid.npoints), dtype=torch.float64, device=rho.device)
        plan.weighted_densities(out)
        return like_input(out[5], krho)

    def derive(self, krho, wd=None):
        if wd is not None:
            raise NotImplementedError('derive(krho, wd) with a precomputed weighted density')
        rho = density_from(self.grid, krho)
        c1, _ = self._engine.run(self._func(), rho)
        return like_input(c1.reshape(self.grid.npoints), krho)

    def value(self, krho, wd=None, local=False):
        if local or wd is not None:
            raise NotImplementedError('local / precomputed-wd WDA values')
        rho = density_from(self.grid, krho)
        _, s = self._engine.run(self._func(), rho, want_c1=False)
        return s[S_FEX] * self.grid.dr


class ExternalPotential(Functional):
    """External potential on the grid (functionals.py:941-1083), spherical LJ
    guests; the grid sum runs on the device (see extpot.get_external_potential)."""
    name = 'ExtPot'

    def __init__(self, grid, system, epot_dr='.', limit_potential=1e4 * U.kjmol, degree=9, interpolate=False,
                 cutoff=12 * U.angstrom):
        self.grid = grid
        self.potential = None
        self.kpotential = None
        self.system = system
        self.host = system.host
        self.guest = system.guest
        self.interpolate = interpolate
        self.cutoff = cutoff
        self.epot_dr = epot_dr
        self.limit_potential = limit_potential
        self.degree = degree
        self._dev = None
        self.generated = None     # {'images': bool} when generated from the host

    def copy(self, grid=None):
        e = type(self)(grid if grid is not None else self.grid.copy(), self.system, self.epot_dr,
                       self.limit_potential, self.degree, self.interpolate, self.cutoff)
        if self.potential is not None:
            e.potential = self.potential.copy()
        return e

    def generate_potential(self, temperature=None, rewrite=False, images=False):
        from .extpot import get_external_potential
        if not isinstance(self.guest, SphericalLJGuest):
            raise NotImplementedError('only spherical LJ guests are supported')
        V = get_external_potential(self.grid, self.host.ff, self.guest.sigma, self.guest.epsilon, self.host.mol,
                                   cutoff=self.cutoff, images=images, as_tensor=True)
        self.generated = {'images': bool(images)}
        self._dev = V.reshape(self.grid.npoints)
        self.potential = self._dev.cpu().numpy()

    def clamp(self):
        """Non-finite and larger-than-limit values -> limit
        (FreeEnergy.add_external_potential, functionals.py:247-251)."""
        V = np.array(self.potential, dtype=np.float64)
        V[~np.isfinite(V)] = self.limit_potential
        V[V > self.limit_potential] = self.limit_potential
        self.potential = V
        self._dev = None

    @property
    def potential_dev(self):
        if self._dev is None:
            self._dev = as_device(self.potential).reshape(self.grid.npoints)
        return self._dev

    def load_potential(self, fn):                      # functionals.py:1009-1018
        potential = np.load(fn)
        diff = np.array([self.grid.npoints[i] - potential.shape[i] for i in range(3)])
        if np.allclose(diff, -1):
            self.potential = potential[:-1, :-1, :-1]
        elif np.allclose(diff, 0):
            self.potential = potential
        else:
            raise ValueError('Grid shape %s does not match potential shape %s' % (self.grid.npoints, potential.shape))
        self.generated = None
        self._dev = None

    def dump_potential(self, fn):
        assert self.potential is not None
        np.save(fn, self.potential)

    def derive(self, krho):
        return self.potential if isinstance(krho, np.ndarray) else self.potential_dev

    def value(self, krho, local=False):
        rho = density_from(self.grid, krho)
        if local:
            return like_input(rho * self.potential_dev, krho)
        return float((rho * self.potential_dev).sum()) * self.grid.dr


class FreeEnergy:
    """Container of functional parts (functionals.py:25-362)."""

    def __init__(self, grid, system, temperature, workdir='.', name_dict=None, overwrite=False):
        self.grid = grid
        self.system = system
        self.temperature = None
        self.beta = None
        self.wavelength = None
        self.parts = []
        self.part_names = []
        self.part_dict = {}
        self.workdir = Path(workdir)
        self.name_dict = name_dict or {}
        self.overwrite = overwrite
        self.fn_tracking = None
        self.set_temperature(temperature)

    def set_temperature(self, temperature, **kwargs):
        self.temperature = temperature
        self.beta = 1.0 / (U.boltzmann * temperature)
        self.wavelength = self.system.guest.wavelength(temperature)
        self.system.guest.compute_hardsphere_radius(temperature, **kwargs)
        for part in self.parts:
            part.set_temperature(temperature, Rhs=self.system.guest.Rhs, **kwargs)

    def add_part(self, part):
        self.parts.append(part)
        self.part_names.append(part.name)
        self.part_dict[part.name] = part

    def remove_part(self, part_name):
        i = self.part_names.index(part_name)
        self.parts.pop(i)
        self.part_names.pop(i)
        self.part_dict.pop(part_name, None)

    def add_hard_sphere(self, version='atWBII'):
        hs = HardSphereFunctional(self.grid, self.system.guest.Rhs, version=version)
        hs.set_temperature(self.temperature, self.system.guest.Rhs)
        self.add_part(hs)

    def add_wdav(self, eos, **kwargs):
        wda = WDAVFunctional(self.grid, self.system.guest.Rhs, eos)
        wda.set_temperature(self.temperature, self.system.guest.Rhs)
        self.add_part(wda)

    def add_external_potential(self, temperature=None, cutoff=12 * U.angstrom, limit_potential=1e4 * U.kjmol,
                               degree=9, interpolate=False, rewrite=False, load_fn=None, save_fn=None,
                               potential=None, images=False, **kwargs):
        epot = ExternalPotential(self.grid, self.system, '.', limit_potential=limit_potential, degree=degree,
                                 interpolate=interpolate, cutoff=cutoff)
        if potential is not None:
            epot.potential = np.asarray(potential, dtype=np.float64).reshape(self.grid.npoints)
        elif load_fn is not None:
            assert str(load_fn).endswith('.npy') and os.path.isfile(load_fn)
            epot.load_potential(load_fn)
        else:
            epot.generate_potential(temperature, rewrite=rewrite, images=images)
            if save_fn is not None:
                epot.dump_potential(save_fn)
        epot.clamp()
        self.add_part(epot)
        return epot

    def track_omega(self, chempot, rho):
        """Grand potential of a density (FreeEnergy.track without file output,
        functionals.py:121-174)."""
        from .solver import grand_potential
        return grand_potential(self, chempot, rho)
