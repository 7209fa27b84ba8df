"""Grid, guest and host objects with the reference's interface
(cmmdft/system.py), backed by the HIP plan for transforms.

``Grid`` keeps the reference attributes (npoints, spacings, dr, points,
kpoints, scalprod, sigma_lanczos; system.py:231-295) -- the big (N,N,N,4)
arrays are built on first access -- and its transforms follow the reference
convention fft = fftn(x) P / N, ifft = Re ifftn(X conj(P)) N with
P = exp(i pi scalprod) (system.py:310-332), computed by the HIP R2C/C2R.
"""
from __future__ import annotations

import os

import numpy as np

from . import units as U
from ._device import torch_mod, as_device

__all__ = ['System', 'Host', 'EmptyHost', 'NanoporousHost', 'Guest', 'SphericalLJGuest', 'Grid', 'Cell']


class Cell:
    """Minimal stand-in for yaff's Cell: ``rvecs`` rows are the cell vectors."""

    def __init__(self, rvecs):
        r = np.asarray(rvecs, dtype=np.float64)
        # a molecule's checkpoint has no cell vectors (kind=fltar 0,3)
        self.rvecs = r.reshape(0, 3) if r.size == 0 else r.reshape(3, 3)
        self.nvec = len(self.rvecs)

    @property
    def volume(self):
        return abs(float(np.linalg.det(self.rvecs)))

    @property
    def parameters(self):
        lengths = np.linalg.norm(self.rvecs, axis=1)
        return lengths, None


class System:
    def __init__(self, host, guest):
        self.host = host
        self.guest = guest

    def copy(self):
        return System(self.host, self.guest.copy())


class Host:
    def __init__(self, name, cell):
        self.name = name
        self.cell = cell if hasattr(cell, 'rvecs') else Cell(cell)


class EmptyHost(Host):
    """Host without atoms (system.py:79-96)."""

    def __init__(self, name, cell=None, volume=None):
        if cell is None:
            assert volume is not None, 'Either cell or volume keyword argument must be defined'
            cell = Cell(np.eye(3) * volume ** (1.0 / 3.0))
        super().__init__(name, cell)


def read_chk(fn):
    """Fields of a yaff checkpoint text file: ``name kind=KIND [shape]``
    header lines followed by the values of array kinds (intar, fltar, strar,
    blnar), scalars (int, flt, str, bln) on the header line, ``none``."""
    conv = {'int': int, 'flt': float, 'str': str, 'bln': lambda v: v.lower() in ('true', '1')}
    out = {}
    with open(fn) as fh:
        tokens = None
        key = kind = None
        shape = None
        for line in fh:
            words = line.split()
            if len(words) >= 2 and words[1].startswith('kind='):
                if tokens is not None:
                    raise ValueError('%s: field %s is short of values' % (fn, key))
                key, kind = words[0], words[1][5:]
                if kind == 'none':
                    out[key] = None
                elif kind.endswith('ar'):
                    shape = tuple(int(v) for v in words[2].split(','))
                    tokens = []
                    if int(np.prod(shape)) == 0:
                        out[key] = np.zeros(shape)
                        tokens = None
                else:
                    out[key] = conv[kind](' '.join(words[2:]) if kind == 'str' else words[2])
                continue
            if tokens is None:
                continue
            tokens.extend(words)
            n = int(np.prod(shape))
            if len(tokens) >= n:
                base = kind[:-2]
                if base == 'str':
                    out[key] = list(tokens[:n]) if len(shape) == 1 else np.array(tokens[:n]).reshape(shape)
                else:
                    dtype = {'int': np.int64, 'flt': np.float64, 'bln': bool}[base]
                    vals = tokens[:n] if base != 'bln' else [v.lower() in ('true', '1') for v in tokens[:n]]
                    out[key] = np.array(vals, dtype=dtype).reshape(shape)
                tokens = None
        if tokens is not None:
            raise ValueError('%s: field %s is short of values' % (fn, key))
    return out


class ChkSystem:
    """The parts of a yaff System the hot path reads (positions, cell,
    force-field atom types, masses), from a checkpoint file."""

    def __init__(self, pos, rvecs, ffatypes, ffatype_ids, masses=None, numbers=None, charges=None):
        self.pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        self.cell = Cell(rvecs)
        self.ffatypes = list(ffatypes)
        self.ffatype_ids = np.asarray(ffatype_ids, dtype=np.int64)
        self.masses = None if masses is None else np.asarray(masses, dtype=np.float64)
        self.numbers = None if numbers is None else np.asarray(numbers)
        self.charges = None if charges is None else np.asarray(charges, dtype=np.float64)
        self.natom = len(self.pos)

    @classmethod
    def from_file(cls, fn):
        d = read_chk(fn)
        for k in ('pos', 'ffatypes', 'ffatype_ids'):
            if d.get(k) is None:
                raise ValueError('%s has no %s' % (fn, k))
        # a molecule may have no cell vectors at all
        rvecs = d['rvecs'] if d.get('rvecs') is not None else np.zeros((0, 3))
        return cls(d['pos'], rvecs, d['ffatypes'], d['ffatype_ids'], d.get('masses'), d.get('numbers'),
                   d.get('charges'))


class NanoporousHost(Host):
    """Host framework (system.py:53-76): ``chk`` is a yaff checkpoint file
    (or an object with ``pos``, ``ffatype_ids``, ``ffatypes`` and
    ``cell.rvecs``), ``par`` a yaff LJ parameter file (or a dict
    {ffatype: (sigma, epsilon)}).  Positions are centred on their mean
    (system.py:68-70)."""

    def __init__(self, name, chk, par):
        import copy
        from .extpot import read_pars_file
        if isinstance(chk, (str, os.PathLike)):
            self.chk = chk
            mol = ChkSystem.from_file(chk)
        else:
            self.chk = None
            mol = copy.copy(chk)
        if isinstance(par, (str, os.PathLike)):
            self.par = par
            ff = read_pars_file(par)
        else:
            self.par = None
            ff = par
        self.mol = mol
        pos = np.asarray(mol.pos, dtype=np.float64)
        self.mol.pos = pos - pos.sum(axis=0) / len(pos)
        self.ff = dict(ff)
        used = {self.mol.ffatypes[int(i)] for i in np.unique(np.asarray(self.mol.ffatype_ids))}
        missing = sorted(used - set(self.ff))
        if missing:
            raise ValueError('no LJ parameters for host atom types %s' % missing)
        super().__init__(name, mol.cell)

    def copy(self):
        return NanoporousHost(self.name, self.chk if self.chk is not None else self.mol,
                              self.par if self.par is not None else self.ff)


class Guest:
    def __init__(self, name, mass):
        self.name = name
        self.mass = mass
        self.preset_Rhs = None
        self.preset_Rhs_zero = None
        self.Rhs = None
        self.Rhs_zero = None

    def copy(self):
        return type(self)(self.name, self.mass)

    def wavelength(self, temperature):                 # system.py:115-117
        kT = U.boltzmann * temperature
        return U.planck / np.sqrt(2 * np.pi * self.mass * kT)

    def set_fixed_rhs(self, Rhs, Rhs_zero):
        self.preset_Rhs = Rhs
        self.preset_Rhs_zero = Rhs_zero

    def _calculate_hardsphere_radius(self, temperature, **kwargs):
        raise NotImplementedError

    def compute_hardsphere_radius(self, temperature, **kwargs):
        if self.preset_Rhs_zero is not None:
            self.Rhs, self.Rhs_zero = self.preset_Rhs, self.preset_Rhs_zero
        else:
            self.Rhs, self.Rhs_zero = self._calculate_hardsphere_radius(temperature, **kwargs)


class SphericalLJGuest(Guest):
    """Spherical LJ guest; hard-sphere radius from the Barker-Henderson fit
    R = sigma (1 + 0.2977 T*) / (1 + 0.33163 T* + 0.0010477 T*^2) / 2 with
    T* = kT / epsilon (system.py:141-158)."""

    def __init__(self, name, mass, sigma, epsilon):
        Guest.__init__(self, name, mass)
        self.sigma = sigma
        self.epsilon = epsilon
        self.natom = 1

    def copy(self):
        return type(self)(self.name, self.mass, self.sigma, self.epsilon)

    def _calculate_hardsphere_radius(self, temperature, **kwargs):
        Tt = U.boltzmann * temperature / self.epsilon
        Rhs = self.sigma * (1 + 0.2977 * Tt) / (1 + 0.33163 * Tt + 0.0010477 * Tt ** 2) / 2
        return Rhs, self.sigma


class NonSphericalGuest(Guest):
    """Rigid multi-site guest (system.py:162-212): atoms, masses and force-field
    types from a checkpoint file ``chk`` (or a system object), LJ parameters
    of its atom types from the yaff parameter file ``par`` (or a dict).  The
    mass is the sum of the atomic masses."""

    def __init__(self, name, chk, par):
        from .extpot import read_pars_file
        if isinstance(chk, (str, os.PathLike)):
            self.chk = chk
            self.mol = ChkSystem.from_file(chk)
        else:
            self.chk = chk
            self.mol = chk
        self.par = par
        self.ff = read_pars_file(par) if isinstance(par, (str, os.PathLike)) else dict(par)
        self.natom = self.mol.natom
        if self.mol.masses is None:
            raise ValueError('Masses not defined in guest chk file, cannot compute total mass for NonSphericalGuest')
        Guest.__init__(self, name, float(np.sum(self.mol.masses)))

    def copy(self):
        return type(self)(self.name, self.chk, self.par)

    def _calculate_hardsphere_radius(self, temperature, **kwargs):
        raise NotImplementedError('Using generic NonSphericalGuest class, cannot compute hardsphere radius, '
                                  'use DualModelGuest class')


class DualModelGuest(SphericalLJGuest, NonSphericalGuest):
    """Guest with a spherical LJ model (hard-sphere radius, mean field) and a
    multi-site model (external potential), system.py:214-229."""

    def __init__(self, name, mass, sigma, epsilon, chk, par):
        SphericalLJGuest.__init__(self, name, mass, sigma, epsilon)
        NonSphericalGuest.__init__(self, name, chk, par)

    def copy(self):
        return type(self)(self.name, self.mass, self.sigma, self.epsilon, self.chk, self.par)

    def _calculate_hardsphere_radius(self, temperature, **kwargs):
        return SphericalLJGuest._calculate_hardsphere_radius(self, temperature, **kwargs)


class Grid:
    """Periodic real-space grid centred at the origin (system.py:231-332)."""

    def __init__(self, cell, npoints=None, spacing=0.25 * U.angstrom, silent=False):
        self.cell = cell if hasattr(cell, 'rvecs') else Cell(cell)
        rv = np.asarray(self.cell.rvecs, dtype=np.float64)
        if npoints is None:
            lengths = np.linalg.norm(rv, axis=1)
            self.npoints = [int(np.ceil(l / spacing)) for l in lengths]
        elif isinstance(npoints, (int, np.integer)):
            self.npoints = [int(npoints)] * 3
        else:
            self.npoints = [int(n) for n in npoints]
        self.suffix = '_'.join('%d' % n for n in self.npoints)
        self.spacings = [np.linalg.norm(rv[:, a]) / self.npoints[a] for a in range(3)]
        self.dr = self.cell.volume / np.prod(self.npoints)
        self.dk = 1.0 / self.dr
        self.npts = int(np.prod(self.npoints))
        self._plans = {}
        self._cache = {}

    # -- reference attributes, built on demand ---------------------------------
    def _frac(self):
        return [np.linspace(-0.5, 0.5, num=n, endpoint=False) for n in self.npoints]

    @property
    def points(self):
        if 'points' not in self._cache:
            fg = np.asarray(np.meshgrid(*self._frac(), indexing='ij'))
            pts = np.zeros(self.npoints + [4])
            pts[..., :3] = np.einsum('ab,aijk->ijkb', self.cell.rvecs, fg)
            pts[..., 3] = np.sqrt(pts[..., 0] ** 2 + pts[..., 1] ** 2 + pts[..., 2] ** 2)
            self._cache['points'] = pts
        return self._cache['points']

    def k1d(self, axis):
        return 2 * np.pi * np.fft.fftfreq(self.npoints[axis], d=self.spacings[axis])

    @property
    def kpoints(self):
        if 'kpoints' not in self._cache:
            kg = np.meshgrid(*[self.k1d(a) for a in range(3)], indexing='ij')
            kp = np.zeros(self.npoints + [4])
            for a in range(3):
                kp[..., a] = kg[a]
            kp[..., 3] = np.sqrt(kp[..., 0] ** 2 + kp[..., 1] ** 2 + kp[..., 2] ** 2)
            self._cache['kpoints'] = kp
        return self._cache['kpoints']

    @property
    def scalprod(self):
        kp = self.kpoints
        return sum(kp[..., a] * self.spacings[a] * self.npoints[a] for a in range(3))

    @property
    def sigma_lanczos(self):
        kcut = 2 * np.pi / np.array(self.spacings)
        kp = self.kpoints
        return np.sinc(kp[..., 0] / kcut[0]) * np.sinc(kp[..., 1] / kcut[1]) * np.sinc(kp[..., 2] / kcut[2])

    def supercell(self, supercell):
        supercell = np.asarray(supercell)
        return Grid(Cell(self.cell.rvecs * supercell), npoints=list(np.asarray(self.npoints) * supercell))

    def copy(self):
        return Grid(self.cell, npoints=self.npoints)

    # -- device plan ------------------------------------------------------------
    def plan(self, nbatch=1):
        from ._lib import Plan
        key = int(nbatch)
        if key not in self._plans:
            self._plans[key] = Plan(self.npoints, self.cell.rvecs, nbatch=key)
        p = self._plans[key]
        p.bind_stream()
        return p

    def phase_half(self):
        """P = exp(i pi scalprod) on the half spectrum, as a device tensor."""
        torch = torch_mod()
        if 'phase_half' not in self._cache:
            nzh = self.npoints[2] // 2 + 1
            sp = [self.k1d(a) * self.spacings[a] * self.npoints[a] for a in range(3)]
            s = sp[0][:, None, None] + sp[1][None, :, None] + sp[2][None, None, :nzh]
            self._cache['phase_half'] = torch.from_numpy(np.exp(1j * np.pi * s)).to('cuda')
        return self._cache['phase_half']

    # -- transforms (reference convention) --------------------------------------
    def integrate(self, data):                          # system.py:307-308
        if hasattr(data, 'is_cuda'):
            return float(data.sum()) * self.dr
        return np.sum(data) * self.dr

    def rfft_native(self, rho):
        """Unnormalised R2C half spectrum of a device density (complex128)."""
        torch = torch_mod()
        nzh = self.npoints[2] // 2 + 1
        x = as_device(rho).reshape(self.npoints).contiguous()
        X = torch.empty(tuple(self.npoints[:2]) + (nzh,), dtype=torch.complex128, device=x.device)
        self.plan().fft_r2c(x, torch.view_as_real(X))
        return X

    def irfft_native(self, X):
        """Unnormalised C2R of a device half spectrum (real part semantics)."""
        torch = torch_mod()
        X = X.contiguous()
        out = torch.empty(tuple(self.npoints), dtype=torch.float64, device=X.device)
        self.plan().fft_c2r(torch.view_as_real(X), out)
        return out

    def fft(self, rdata):                               # system.py:310-311
        torch = torch_mod()
        X = self.rfft_native(rdata) * self.phase_half() / self.npts
        full = _hermitian_full(X, self.npoints, self)
        return full.cpu().numpy() if isinstance(rdata, np.ndarray) else full

    def ifft(self, fdata):                              # system.py:322-323
        torch = torch_mod()
        nzh = self.npoints[2] // 2 + 1
        F = as_device(fdata)[..., :nzh].to(torch.complex128)
        rho = self.irfft_native((F * torch.conj(self.phase_half())).contiguous())
        return rho.cpu().numpy() if isinstance(fdata, np.ndarray) else rho

    def fftn(self, rdata):                              # system.py:313-320
        comps = [self.fft(rdata[..., i]) for i in range(rdata.shape[-1])]
        return np.stack(comps, axis=-1) if isinstance(rdata, np.ndarray) else torch_mod().stack(comps, dim=-1)

    def ifftn(self, fdata):                             # system.py:325-332
        comps = [self.ifft(fdata[..., i]) for i in range(fdata.shape[-1])]
        return np.stack(comps, axis=-1) if isinstance(fdata, np.ndarray) else torch_mod().stack(comps, dim=-1)


def _hermitian_full(Xh, npoints, grid):
    """Full spectrum from the half spectrum of a real field (device).  The
    phase-weighted half spectrum is extended by X(-k) = conj(X(k)) P(k)^2 ...
    computed here directly: full(k) = conj(Xu(-k)) P(k) / N for kz beyond the
    half, Xu being the unphased spectrum."""
    torch = torch_mod()
    nx, ny, nz = npoints
    nzh = nz // 2 + 1
    P = grid.phase_half()
    Xu = Xh / P * grid.npts                  # unphased, unnormalised half spectrum
    full = torch.empty((nx, ny, nz), dtype=torch.complex128, device=Xh.device)
    full[..., :nzh] = Xh
    if nz - nzh > 0:
        kz = torch.arange(nzh, nz, device=Xh.device)
        mz = (nz - kz) % nz
        ix = (-torch.arange(nx, device=Xh.device)) % nx
        iy = (-torch.arange(ny, device=Xh.device)) % ny
        Xm = torch.conj(Xu[ix][:, iy][:, :, mz])
        sp = [grid.k1d(a) * grid.spacings[a] * grid.npoints[a] for a in range(3)]
        s = sp[0][:, None, None] + sp[1][None, :, None] + sp[2][None, None, nzh:]
        Pf = torch.from_numpy(np.exp(1j * np.pi * s)).to(Xh.device)
        full[..., nzh:] = Xm * Pf / grid.npts
    return full
