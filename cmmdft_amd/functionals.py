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
from ._lib import CmmFunctional, S_PHI, S_FEX, S_N3MAX, S_FMFA
from .system import NanoporousHost, SphericalLJGuest, NonSphericalGuest

__all__ = ['FreeEnergy', 'Functional', 'HardSphereFunctional', 'WDAVFunctional', 'MFAFunctional',
           'ExternalPotential', 'density_from', 'build_functional', 'device_parts', 'configure_plan']

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
    f.n3_series = 1e-8
    if hs is not None:
        f.hs_base = hs_base_of(hs.version)
        f.hs_antisym = 1 if 'a' in hs.version else 0
        f.hs_tensor = 1 if 't' in hs.version else 0
        f.R = hs.R
        f.n3_series = hs.n3_series
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
        # a WDA of mean-field EOS terms only is linear in rho: the device
        # applies it in k space unless the part asks for the pointwise route
        f.wda_pointwise = 0 if getattr(wda, 'kspace', True) else 1
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


def device_parts(parts):
    """(hard sphere, WDA, MFA, [external potentials]) of a list of
    functional parts; parts with no device implementation raise."""
    hs = wda = mfa = None
    eps = []
    for part in parts:
        if part.name == 'HardSphere':
            if hs is not None:
                raise NotImplementedError('more than one hard-sphere part')
            hs = part
        elif part.name == 'WDA-V':
            if wda is not None:
                raise NotImplementedError('more than one WDA part on the device')
            wda = part
        elif part.name == 'MFA':
            if mfa is not None:
                raise NotImplementedError('more than one mean-field part')
            mfa = part
        elif part.name == 'ExtPot':
            eps.append(part)
        else:
            raise NotImplementedError('functional part %r has no device implementation' % part.name)
    return hs, wda, mfa, eps


def configure_plan(plan, hs, wda, mfa, kT):
    """Set a plan's functional: hard sphere and WDA, then the mean-field
    kernel table (or none)."""
    plan.set_functional(build_functional(hs=hs, wda=wda, kT=kT))
    plan.set_mfa(mfa.table() if mfa is not None else None)


class _PartEngine:
    """Runs the fused pipeline of one grid with a single part enabled."""

    def __init__(self, grid):
        self.grid = grid

    def run(self, func, rho, want_c1=True, mfa=None):
        torch = torch_mod()
        plan = self.grid.plan()
        plan.set_functional(func)
        plan.set_mfa(mfa)
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

    def __init__(self, grid, Rhs, version='atWBII', n3_series=1e-8):
        """``n3_series``: the modified and White-Bear II forms of phi2, phi3
        and their derivatives switch to their second-order series below this
        packing fraction.  The reference's 1e-8 (functionals.py:726-767) is
        the default; above it the closed forms lose digits by cancellation
        (e.g. dphi3/dn3 keeps none for 3e-8 < n3 < 1e-6), which makes the
        solution a non-smooth function of the host parameters at that level.
        Up to 1e-5 the series is exact to rounding, and derivatives of the
        solution (finite differences, the adjoint) become smooth."""
        assert version in FMT_NAMES + MFMT_NAMES + WBII_NAMES, 'Unknown hard sphere functional version: %s' % version
        assert 0.0 <= n3_series <= 1e-3, 'n3_series must lie in [0, 1e-3]'
        self.temperature = None
        self.beta = None
        self.grid = grid
        self.R = Rhs
        self.version = version
        self.n3_series = float(n3_series)
        self._engine = _PartEngine(grid)

    def copy(self, grid=None):
        return type(self)(grid if grid is not None else self.grid.copy(), self.R, self.version, self.n3_series)

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
        """F_HS = int Phi / beta, or Phi / beta per point with ``local``
        (functionals.py:611-618)."""
        if local:
            torch = torch_mod()
            nf = self.weighted_densities_dev(krho)
            plan = self.grid.plan()
            phi = torch.empty(tuple(self.grid.npoints), dtype=torch.float64, device='cuda')
            plan.local_density(nfld=nf, phi=phi)
            return like_input(phi, krho)
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

    def get_n2_nv2(self, krho):                       # functionals.py:558-563
        """|n2| / n_v2 per vector component, shape (nx, ny, nz, 3)."""
        torch = torch_mod()
        n = self.weighted_densities_dev(krho)
        out = n[0].abs()[..., None] / torch.stack([n[2], n[3], n[4]], dim=-1)
        return like_input(out, krho)

    def _weight_table(self):
        torch = torch_mod()
        key = float(self.R)
        if getattr(self, '_wtab_key', None) != key:
            plan = self.grid.plan()
            t = torch.empty((2,) + tuple(self.grid.npoints), dtype=torch.float64, device='cuda')
            plan.hs_weight_table(self.R, t)
            self._wtab, self._wtab_key = t, key
        return self._wtab

    @property
    def scalar_weight_functions(self):
        """[w0, w1, w2, w3]: the reference's k-space weights on the full grid
        (complex numpy arrays, functionals.py:433-447), from the device."""
        t = self._weight_table()
        R = self.R
        w0 = t[0].cpu().numpy().astype(np.complex128)
        w3 = (4.0 * np.pi * R ** 3 / 3.0) * t[1].cpu().numpy().astype(np.complex128)
        return [w0, R * w0, 4.0 * np.pi * R ** 2 * w0, w3]

    @property
    def vector_weight_functions(self):
        """[w_v1, w_v2] with w_v2 = -i k w3 (zero where kR <= 1e-8) and
        w_v1 = w_v2 / (4 pi R) (functionals.py:449-451), shape (nx, ny, nz, 3)."""
        R = self.R
        w3 = self.scalar_weight_functions[3]
        k = [self.grid.k1d(a) for a in range(3)]
        kv = np.stack(np.meshgrid(*k, indexing='ij'), axis=-1)
        omega = np.sqrt(np.sum(kv * kv, axis=-1)) * R
        wv2 = -1j * w3[..., None] * kv
        wv2[omega <= 1e-8] = 0.0
        return [wv2 / (4 * np.pi * R), wv2]

    def set_density(self, krho):
        self.krho = krho
        n = self.weighted_densities_dev(krho)
        self.weighted_densities = n


class WDAVFunctional(Functional):
    """Weighted-density approximation with the excess free energy per volume
    of an EOS; weight = normalised ball of diameter D = 2 R
    (functionals.py:1116-1184)."""
    name = 'WDA-V'

    def __init__(self, grid, Rhs, eos, kspace=True):
        """``kspace``: when every EOS term is mean field (f_ex = A rho_bar^2,
        the MFA EOS) the functional is quadratic and its derivative
        w_D * 2A rho_bar is linear in rho; the device then applies it in k
        space (2A w_D^2 rho^) without forming rho_bar.  False keeps the
        pointwise route that every other EOS takes.  The two agree to
        rounding."""
        self.temperature = None
        self.grid = grid
        self.eos = eos
        self.R = Rhs
        self.kspace = bool(kspace)
        self._engine = _PartEngine(grid)

    def copy(self, grid=None):
        return type(self)(grid if grid is not None else self.grid.copy(), self.R, self.eos, self.kspace)

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

    def _eos_fields(self, wd, f=False, df=False):
        """f_ex(wd) and/or f_ex'(wd) per point on the device."""
        torch = torch_mod()
        rb = as_device(wd).reshape(tuple(self.grid.npoints)).contiguous()
        plan = self.grid.plan()
        plan.set_functional(self._func())
        fo = torch.empty_like(rb) if f else None
        dfo = torch.empty_like(rb) if df else None
        plan.local_density(rbar=rb, fex=fo, dfex=dfo)
        return fo, dfo

    def derive(self, krho, wd=None):
        """c_WDA = ifft(fft(f_ex'(wd)) w_D) with wd the weighted density of
        krho unless given (functionals.py:1160-1174)."""
        if wd is not None:
            torch = torch_mod()
            _, g = self._eos_fields(wd, df=True)
            plan = self.grid.plan()
            plan.set_functional(self._func())
            plan.seed(g.reshape(-1))
            out = torch.empty((plan.nfields,) + tuple(self.grid.npoints), dtype=torch.float64, device='cuda')
            plan.weighted_densities(out)
            return like_input(out[5].clone(), wd if isinstance(wd, np.ndarray) else krho)
        rho = density_from(self.grid, krho)
        c1, _ = self._engine.run(self._func(), rho)
        return like_input(c1.reshape(self.grid.npoints), krho)

    def value(self, krho, wd=None, local=False):
        """int f_ex(wd), or f_ex(wd) per point with ``local``; wd is the
        weighted density of krho unless given (functionals.py:1176-1184)."""
        if wd is not None or local:
            from ._lib import field_reduce
            ref = wd if (wd is not None and isinstance(wd, np.ndarray)) else krho
            if wd is None:
                wd = self._get_weighted_density(krho)
            f, _ = self._eos_fields(wd, f=True)
            if local:
                return like_input(f, ref)
            return field_reduce(f, 'sum') * self.grid.dr
        rho = density_from(self.grid, krho)
        _, s = self._engine.run(self._func(), rho, want_c1=False)
        return s[S_FEX] * self.grid.dr


class MFAFunctional(Functional):
    """Mean-field attraction between guests, c_MFA = U * rho
    (functionals.py:770-939).

    The convolution is the reference's ifft(krho kpotential) V, phase
    included (its phased transforms leave one factor exp(i pi scalprod) in
    the product); the device keeps its kernel table as a half spectrum (see
    cmm_mfa_table) and the derivative chain adds it to c_ex.  With tail
    corrections the pair potential lives on a supercell of ``repetitions``
    times the grid; the reference samples every repetitions-th supercell
    wave vector, which is the transform of the potential folded onto the
    grid: the device folds it in real space.  The potential is generated on
    the device; ``potential`` (the reference's array on ``self.grid``, the
    supercell with tail corrections) is materialised on first access.
    ``kpotential`` is not materialised (None).
    """
    name = 'MFA'

    def __init__(self, grid, tailcorrections=False, repetitions=[2, 2, 2]):
        self.tailcorrections = tailcorrections
        self.repetitions = repetitions
        if tailcorrections:
            self.small_grid = grid
            self.grid = grid.supercell(repetitions)
        else:
            self.grid = grid
        self.base_grid = grid
        self.reps = [int(r) for r in repetitions] if tailcorrections else [1, 1, 1]
        self.kpotential = None
        self._potential = None      # numpy, on self.grid
        self._lj = None             # parameters of a potential generated on the device
        self._fold = None           # device: potential folded onto the base grid
        self._kd = None             # device: kernel table (half spectrum)
        self._engine = _PartEngine(grid)

    def copy(self, grid=None):
        mfa = type(self)(grid if grid is not None else self.base_grid.copy(), self.tailcorrections, self.repetitions)
        mfa._potential = None if self._potential is None else self._potential.copy()
        mfa._lj = None if self._lj is None else dict(self._lj)
        mfa._fold, mfa._kd = self._fold, self._kd
        return mfa

    def set_temperature(self, temperature, **kwargs):
        pass

    # -- the pair potential ------------------------------------------------
    def _lj_field(self, reps):
        from ._lib import mfa_lj_potential
        lj = self._lj
        return mfa_lj_potential(self.grid.npoints, self.grid.cell.rvecs, reps, lj['sigma'], lj['epsilon'], lj['rmin'],
                                lj['limit'], lj['cutoff'])

    @property
    def potential(self):
        if self._potential is None and self._lj is not None:
            self._potential = self._lj_field([1, 1, 1]).cpu().numpy()
        return self._potential

    @potential.setter
    def potential(self, value):
        self._potential = value
        self._lj = None

    def generate_potential_lj(self, sigma, epsilon, rmin=None, limit_potential=0, cutoff=None, **kwargs):
        """U(r) on the (super)grid (functionals.py:878-910); the kernel table
        takes the Lanczos factor."""
        if rmin is None:
            rmin = sigma
        self._lj = dict(sigma=float(sigma), epsilon=float(epsilon), rmin=float(rmin), limit=float(limit_potential),
                        cutoff=None if cutoff is None else float(cutoff))
        self._potential = None
        self._fold = self._lj_field(self.reps)
        self._build_table(lanczos=True)

    def generate_potential(self, *args, **kwargs):
        raise NotImplementedError('MFA potentials from a force field need yaff; use generate_potential_lj')

    def load_potential(self, fn):
        """(functionals.py:809-812): the kernel table of a loaded potential
        has no Lanczos factor."""
        from ._lib import fold
        pot = np.load(fn)
        assert tuple(self.grid.npoints) == pot.shape
        self.potential = pot
        dev = as_device(pot)
        self._fold = dev if self.reps == [1, 1, 1] else fold(dev, self.reps)
        self._build_table(lanczos=False)

    def dump_potential(self, fn):
        assert self.potential is not None
        dn = os.path.dirname(fn)
        if dn and not os.path.exists(dn):
            os.makedirs(dn)
        np.save(fn, self.potential)

    def compute_vdw_a(self):
        """a = (1/2) int U (functionals.py:814-822), over the supercell with
        tail corrections (the sum of the folded potential)."""
        if self._fold is None:
            raise AssertionError('no MFA potential')
        self.a = 0.5 * float(self._fold.sum()) * self.base_grid.dr
        return self.a

    def _build_table(self, lanczos):
        torch = torch_mod()
        g = self.base_grid
        nx, ny, nz = g.npoints
        plan = g.plan()
        uhat = torch.empty((nx, ny, nz // 2 + 1), dtype=torch.complex128, device='cuda')
        plan.fft_r2c(self._fold.contiguous(), torch.view_as_real(uhat))
        kd = torch.empty_like(uhat)
        plan.mfa_table(torch.view_as_real(uhat), self.reps, self.grid.cell.rvecs, lanczos, g.dr / g.npts,
                       torch.view_as_real(kd))
        self._kd = torch.view_as_real(kd)

    def table(self):
        """Device kernel table (real view of the complex half spectrum)."""
        if self._kd is None:
            raise AssertionError('the MFA potential has not been generated or loaded')
        return self._kd

    # -- functional --------------------------------------------------------
    def _func(self):
        # no hard-sphere or WDA part: only the mean-field table acts (kT unused)
        return build_functional(kT=1.0)

    def derive(self, krho):
        """c_MFA (functionals.py:912-926)."""
        rho = density_from(self.base_grid, krho)
        c1, _ = self._engine.run(self._func(), rho, mfa=self.table())
        return like_input(c1.reshape(self.base_grid.npoints), krho)

    def value(self, krho, local=False):
        """(1/2) int rho c_MFA (functionals.py:928-939)."""
        rho = density_from(self.base_grid, krho)
        if local:
            c1, _ = self._engine.run(self._func(), rho, mfa=self.table())
            return like_input(0.5 * rho * c1.reshape(rho.shape), krho)
        _, s = self._engine.run(self._func(), rho, want_c1=False, mfa=self.table())
        return s[S_FMFA] * self.base_grid.dr


class ExternalPotential(Functional):
    """External potential on the grid (functionals.py:941-1083), spherical LJ
    guests; the grid sum runs on the device (see extpot.get_external_potential)."""
    name = 'ExtPot'

    def __init__(self, grid, system, epot_dr='.', limit_potential=1e4 * U.kjmol, degree=9, interpolate=False,
                 cutoff=12 * U.angstrom):
        self.grid = grid
        self._np = None           # numpy copy of the potential (made on first access)
        self._dev = None          # device potential
        self.kpotential = None
        self.system = system
        self.host = system.host
        self.guest = system.guest
        self.interpolate = interpolate
        self.cutoff = cutoff
        self.epot_dr = epot_dr
        self.limit_potential = limit_potential
        self.degree = degree
        self.generated = None     # {'images': bool} when generated from the host

    def copy(self, grid=None):
        e = type(self)(grid if grid is not None else self.grid.copy(), self.system, self.epot_dr,
                       self.limit_potential, self.degree, self.interpolate, self.cutoff)
        if self._dev is not None:
            e._dev = self._dev.clone()
        elif self._np is not None:
            e._np = self._np.copy()
        return e

    @property
    def potential(self):
        """The potential on the grid as a numpy array (functionals.py:941-1083);
        the device copy is authoritative and the numpy one is made on first
        access."""
        if self._np is None and self._dev is not None:
            self._np = self._dev.reshape(self.grid.npoints).cpu().numpy()
        return self._np

    @potential.setter
    def potential(self, value):
        self._np = None if value is None else np.asarray(value, dtype=np.float64)
        self._dev = None

    def generate_potential(self, temperature=None, rewrite=False, images=False):
        """The host's potential on the grid (functionals.py:1030-1066): the LJ
        grid sum for a spherical guest; for a non-spherical guest
        (``NonSphericalGuest``, ``DualModelGuest``) the orientational average
        of ``generate_effective_potential`` at ``temperature`` with
        ``degree`` Lebedev directions, on the device."""
        from .extpot import get_external_potential, get_external_potential_dict, generate_effective_potential
        if isinstance(self.guest, NonSphericalGuest):
            assert temperature is not None, 'the effective potential of a non-spherical guest needs a temperature'
            beta = 1 / temperature / U.boltzmann
            if self.interpolate:
                gens = self._interpolators(images)
            else:
                gens = get_external_potential_dict(self.host.ff, self.guest.ff, self.host.mol, cutoff=self.cutoff,
                                                   images=images)
            V = generate_effective_potential(self.guest.mol, self.grid, beta, gens, degree=self.degree,
                                             as_tensor=True)
            self.generated = {'images': bool(images), 'temperature': temperature}
        elif isinstance(self.guest, SphericalLJGuest):
            V = get_external_potential(self.grid, self.host.ff, self.guest.sigma, self.guest.epsilon, self.host.mol,
                                       cutoff=self.cutoff, images=images, as_tensor=True)
            self.generated = {'images': bool(images)}
        else:
            raise NotImplementedError('no external potential for guests of type %s' % type(self.guest).__name__)
        self._dev = V.reshape(self.grid.npoints)
        self._np = None

    def _interpolators(self, images=False):
        """The reference's interpolate=True set-up (functionals.py:1035-1049):
        per guest atom type, the LJ field and its derivatives on a grid of
        the host cell at 0.15 A spacing (npoints = ceil(L / 0.15 A), so the
        actual spacing is L / npoints), interpolated tricubically with origin
        the first point of this grid and a spacing of exactly 0.15 A, as the
        reference does."""
        from .extpot import get_external_potential_derivatives, Interpolator
        from .system import Grid
        h = 0.15 * U.angstrom
        epot_grid = Grid(self.host.mol.cell, spacing=h)
        origin = self.grid.points[0, 0, 0, :3]
        gens = {}
        for atom, (sigmaff, epsilonff) in self.guest.ff.items():
            vals = get_external_potential_derivatives(epot_grid, self.host.ff, sigmaff, epsilonff, self.host.mol,
                                                      epot_grid.spacings, cutoff=self.cutoff, images=images,
                                                      as_tensor=True)
            gens[atom] = Interpolator(vals, origin, [h, h, h]).tricubic
        return gens

    def set_temperature(self, temperature, **kwargs):
        """A non-spherical guest's potential depends on the temperature
        (functionals.py:1020-1026): regenerate it when it was generated at
        another one."""
        if isinstance(self.guest, NonSphericalGuest) and self.generated and \
                self.generated.get('temperature') not in (None, temperature):
            self.generate_potential(temperature, images=self.generated.get('images', False))
            self.clamp()

    def clamp(self):
        """Non-finite and larger-than-limit values -> limit
        (FreeEnergy.add_external_potential, functionals.py:247-251), on the
        device."""
        torch = torch_mod()
        V = self.potential_dev
        lim = torch.full((), self.limit_potential, dtype=V.dtype, device=V.device)
        self._dev = torch.where(torch.isfinite(V) & (V <= lim), V, lim)
        self._np = None

    @property
    def potential_dev(self):
        if self._dev is None:
            self._dev = as_device(self._np).reshape(self.grid.npoints)
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

    def __init__(self, grid, system, temperature, workdir='.', name_dict=None, overwrite=False, write_files=True):
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
        self.write_files = write_files
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

    def add_hard_sphere(self, version='atWBII', n3_series=1e-8):
        hs = HardSphereFunctional(self.grid, self.system.guest.Rhs, version=version, n3_series=n3_series)
        hs.set_temperature(self.temperature, self.system.guest.Rhs)
        self.add_part(hs)

    def add_wdav(self, eos, kspace=True, **kwargs):
        wda = WDAVFunctional(self.grid, self.system.guest.Rhs, eos, kspace=kspace)
        wda.set_temperature(self.temperature, self.system.guest.Rhs)
        self.add_part(wda)

    def _epot_cache_file(self, temperature):
        """File of the cached external potential and its directory
        (functionals.py:218-235): <prefix>/<hostname>/<guestname>/<ff_suffix>/
        <grid_suffix>/<suffix>/epot.npy, or eff_epot_<T>K.npy for a
        many-atom non-spherical guest; the work directory gets an 'ExtPots'
        link to that directory."""
        nd = self.name_dict
        epot_dr = Path(nd['prefix']) / nd['hostname'] / nd['guestname'] / nd['ff_suffix'] / nd['grid_suffix'] / \
            nd['suffix']
        epot_dr.mkdir(parents=True, exist_ok=True)
        guest = self.system.guest
        if isinstance(guest, NonSphericalGuest) and guest.mol.natom != 1:
            assert temperature is not None, 'Temperature must be provided for non-spherical particles'
            fn = epot_dr / f'eff_epot_{temperature:#3.2f}K.npy'
        else:
            fn = epot_dr / 'epot.npy'
        sym = self.workdir / 'ExtPots'
        if not sym.is_symlink():
            self.workdir.mkdir(parents=True, exist_ok=True)
            sym.symlink_to(epot_dr.absolute())
        return fn

    def add_external_potential(self, temperature=None, cutoff=12 * U.angstrom, limit_potential=1e4 * U.kjmol,
                               degree=9, interpolate=False, rewrite=False, load_fn=None, save_fn=None,
                               potential=None, images=False, **kwargs):
        """The host's external potential as a part (functionals.py:176-251):
        loaded from ``load_fn``; or generated on the device and written to
        ``save_fn``; or (with file output on and a program's name_dict)
        taken from the epot.npy cache, generated and written there when the
        file is missing or ``rewrite``/``overwrite`` is set.  ``potential``
        gives the array directly.  Non-finite and larger values are clamped
        to ``limit_potential``."""
        epot = ExternalPotential(self.grid, self.system, '.', limit_potential=limit_potential, degree=degree,
                                 interpolate=interpolate, cutoff=cutoff)
        if potential is not None:
            epot.potential = np.asarray(potential, dtype=np.float64).reshape(self.grid.npoints)
        elif load_fn is not None:
            assert str(load_fn).endswith('.npy'), 'fn must be a filename of an external potential'
            assert os.path.isfile(load_fn), f'fn must be a filename of an external potential, {load_fn}'
            epot.epot_dr = Path(load_fn).parent
            epot.load_potential(load_fn)
        else:
            fn = None
            if save_fn is not None:
                fn = Path(save_fn)
            elif self.write_files and 'prefix' in self.name_dict:
                fn = self._epot_cache_file(temperature)
            if fn is not None:
                epot.epot_dr = fn.parent
            if fn is None or not os.path.isfile(fn) or self.overwrite or rewrite:
                epot.generate_potential(temperature, rewrite=rewrite, images=images)
                if fn is not None:
                    epot.dump_potential(str(fn))
            else:
                epot.load_potential(fn)
        epot.clamp()
        self.add_part(epot)
        return epot

    def add_mean_field(self, tailcorrections=False, repetitions=[2, 2, 2], cutoff=12 * U.angstrom, **kwargs):
        """Mean-field attraction of spherical LJ guests
        (functionals.py:301-337): the pair potential is generated, or
        loaded from ``workdir/mfa.npy`` when that exists and neither
        ``overwrite`` nor ``rewrite`` is set, and a generated one is saved
        there; a free energy made with ``write_files=False`` does no file
        I/O and always generates."""
        fn = self.workdir / 'mfa.npy'
        # as in the reference (functionals.py:322-325), ``repetitions`` is
        # passed on only when it also appears in kwargs, which a named
        # parameter never does: the supercell is the default 2 x 2 x 2
        if 'repetitions' in kwargs:
            mfa = MFAFunctional(self.grid, tailcorrections=tailcorrections, repetitions=repetitions)
        else:
            mfa = MFAFunctional(self.grid, tailcorrections=tailcorrections)
        if not self.write_files or not os.path.isfile(fn) or self.overwrite or kwargs.get('rewrite', False):
            if not isinstance(self.system.guest, SphericalLJGuest):
                raise NotImplementedError('MFA potential generation only implemented for spherical LJ particles')
            gen = {k: v for k, v in kwargs.items() if k in ('rmin', 'limit_potential')}
            mfa.generate_potential_lj(self.system.guest.sigma, self.system.guest.epsilon, cutoff=cutoff, **gen)
            if self.write_files:
                mfa.dump_potential(str(fn))
        else:
            mfa.load_potential(fn)
        self.add_part(mfa)
        return mfa

    def add_correlation_wda_lj(self, **kwargs):
        """Correlation WDA: MBWR - CS - MFA (functionals.py:339-362), with the
        mean-field a of the MFA part when there is one."""
        from .eos import ModifiedBenedictWebbRubinEOS, CarnahanStarlingEOS, MFAEOS, SumOfEOS
        g = self.system.guest
        a = self.part_dict['MFA'].compute_vdw_a() if 'MFA' in self.part_names else None
        mbwr = ModifiedBenedictWebbRubinEOS(g.mass, g.sigma, g.epsilon)
        cs = CarnahanStarlingEOS(g.mass, g.Rhs)
        mfa = MFAEOS(g.mass, g.sigma, g.epsilon, a=a)
        corr = WDAVFunctional(self.grid, g.Rhs, SumOfEOS(g.mass, [mbwr, cs, mfa], factors=[1, -1, -1]))
        corr.set_temperature(self.temperature, g.Rhs)
        self.add_part(corr)
        return corr

    def init_tracking(self, fn, rewrite=False):
        """Start the energy-tracking file (functionals.py:90-110): a header of
        the columns, or continue an existing file's step count."""
        self.fn_tracking = fn
        if not os.path.isfile(fn) or self.overwrite or rewrite:
            with open(self.fn_tracking, 'w') as f:
                print('#phase\tstep\t     loading\t         -\u00b5N\t        IdGas\t'
                      + '\t'.join(['%s%s' % (' ' * (13 - len(part.name)), part.name) for part in self.parts])
                      + '\t        Total', file=f)
            self.tracking_step = 0
        else:
            with open(self.fn_tracking, 'r') as f:
                lines = f.readlines()
            self.tracking_step = 0 if len(lines) <= 1 else int(lines[-1].split()[1]) + 1

    def track(self, chempot, rho, krho=None, iphase=0, write=True, print_out=False, fn=None, unit=1):
        """Grand potential of a density and, with ``write``, its line in the
        tracking file (functionals.py:112-174).  The ideal-gas term takes
        rho <= 1e-8 as 1e-30, as the reference's does here."""
        from ._lib import field_reduce
        r = as_device(rho).reshape(tuple(self.grid.npoints))
        N = field_reduce(r, 'sum') * self.grid.dr
        Fid = field_reduce(r, 'fid_track', lam3=self.wavelength ** 3) * self.grid.dr / self.beta
        G = Fid - chempot * N
        line = '%6i\t%4i\t%.6e\t%.6e\t% .6e' % (iphase, getattr(self, 'tracking_step', 0), N, (-chempot * N / unit),
                                                   Fid / unit)
        k = r if krho is None else krho
        for part in self.parts:
            Fpart = part.value(k)
            if print_out:
                print(part.name, round(Fpart / U.kjmol, 2))
            G += Fpart
            line += '\t% .6e' % (Fpart / unit)
        line += '\t% .6e' % (G / unit)
        if write:
            with open(fn if fn is not None else self.fn_tracking, 'a') as f:
                print(line, file=f)
            self.tracking_step = getattr(self, 'tracking_step', 0) + 1
        return G

    def track_omega(self, chempot, rho):
        """Grand potential of a density (the solver's form,
        solver.py:103-119), without file output."""
        from .solver import grand_potential
        return grand_potential(self, chempot, rho)
