"""Lennard-Jones host-guest external potential (cmmdft/extpot_calculator.py),
with the grid sum on the device (cmm_lj_grid / cmm_lj_points).

Semantics of the reference (extpot_calculator.py:156-222): for each host atom
Lorentz-Berthelot mixing sigma_ij = (sigma_i + sigma_f)/2,
eps_ij = sqrt(eps_i eps_f); minimum image per Cartesian component with the
row norms of the cell vectors; R = sqrt(r^2 + 1e-16); truncated-and-shifted
LJ inside the cutoff (strict R < r_c), zero beyond.  ``images=True`` sums
every periodic image within the cutoff instead (cells with L/2 < r_c).
"""
from __future__ import annotations

import re

import numpy as np

from . import units as U
from ._device import torch_mod, as_device
from . import _lib

__all__ = ['lennard_jones', 'get_external_potential', 'read_pars_file', 'host_atom_table',
           'ExternalPotentialGenerator', 'get_external_potential_dict', 'effective_potential_vectorized',
           'generate_effective_potential', 'get_external_potential_derivatives', 'Interpolator',
           'get_interpolator_dict']


def lennard_jones(r, sigma, epsilon, derivative=False, cutoff=12 * U.angstrom):
    """Truncated and shifted LJ on an array of distances
    (extpot_calculator.py:156-183); a pointwise helper."""
    r = np.asarray(r, dtype=np.float64)
    inside = r < cutoff
    V = np.zeros_like(r)
    s6 = (sigma / r[inside]) ** 6
    s12 = s6 * s6
    rc6 = (sigma / cutoff) ** 6
    V[inside] = 4 * epsilon * (s12 - s6) - 4 * epsilon * (rc6 ** 2 - rc6)
    if derivative:
        dV, ddV, dddV = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
        ri = r[inside]
        dV[inside] = 24 * epsilon * (s6 - 2 * s12) / ri ** 2
        ddV[inside] = 96 * epsilon * (7 * s12 - 2 * s6) / ri ** 4
        dddV[inside] = 384 * epsilon * (5 * s6 - 28 * s12) / ri ** 6
        return V, dV, ddV, dddV
    return V


def host_atom_table(FF_dict, host_syst):
    """(natom, 5) array x, y, z, sigma_i, epsilon_i from a host system with
    ``pos``, ``ffatype_ids``, ``ffatypes`` (FF_dict lookups as at
    extpot_calculator.py:204-206)."""
    pos = np.asarray(host_syst.pos, dtype=np.float64).reshape(-1, 3)
    ids = np.asarray(host_syst.ffatype_ids)
    tab = np.zeros((len(pos), 5))
    tab[:, :3] = pos
    for i, t in enumerate(ids):
        s, e = FF_dict[host_syst.ffatypes[t]]
        tab[i, 3] = s
        tab[i, 4] = e
    return tab


def get_external_potential(points, FF_dict, sigmaff, epsilonff, host_syst, cutoff=12 * U.angstrom, images=False,
                           limit=0.0, as_tensor=False):
    """External potential at ``points``: a ``Grid`` (regular-grid kernel) or an
    (M, 3) array / tensor of Cartesian points (explicit-point kernel).
    Returns (M,) float64 (numpy unless ``as_tensor`` or a tensor was given)."""
    torch = torch_mod()
    atoms = as_device(host_atom_table(FF_dict, host_syst))
    rvecs = np.asarray(host_syst.cell.rvecs, dtype=np.float64)
    natom = atoms.shape[0]
    if hasattr(points, 'npoints'):
        grid = points
        out = torch.empty(grid.npts, dtype=torch.float64, device='cuda')
        _lib.lj_grid(grid.npoints, np.asarray(grid.cell.rvecs, dtype=np.float64), atoms, natom, sigmaff, epsilonff,
                     cutoff, int(images), limit, out)
        was_tensor = True
    else:
        was_tensor = isinstance(points, torch.Tensor)
        P = as_device(points).reshape(-1, 3).contiguous()
        out = torch.empty(P.shape[0], dtype=torch.float64, device='cuda')
        _lib.lj_points(P, P.shape[0], rvecs, atoms, natom, sigmaff, epsilonff, cutoff, int(images), limit, out)
    if as_tensor or (was_tensor and not hasattr(points, 'npoints')):
        return out
    return out.cpu().numpy()


def read_pars_file(pars_file):
    """LJ parameters {ffatype: (sigma, epsilon)} from a yaff-format parameter
    text file: the 'LJ:UNIT SIGMA/EPSILON <unit>' and 'LJ:PARS <type> <sigma>
    <epsilon>' lines that extpot_calculator.py:705-716 reads through yaff."""
    units = {}
    ff = {}
    with open(pars_file) as fh:
        for raw in fh:
            line = raw.split('#', 1)[0].strip()
            if not line.startswith('LJ:'):
                continue
            words = line.split()
            key = words[0]
            if key == 'LJ:UNIT':
                units[words[1].upper()] = U.parse_unit(words[2])
            elif key == 'LJ:PARS':
                ff[words[1]] = (float(words[2]) * units['SIGMA'], float(words[3]) * units['EPSILON'])
    return ff


# ---------------------------------------------------------------------------
# Non-spherical guests: orientationally averaged potential
# (extpot_calculator.py:572-646, 678-701)
# ---------------------------------------------------------------------------


class ExternalPotentialGenerator:
    """The per-guest-atom-type generator of the reference's
    ``get_external_potential_dict``: ``partial(get_external_potential,
    FF_dict=..., sigmaff=..., epsilonff=..., host_syst=..., cutoff=...)``
    (extpot_calculator.py:696-698).  Calling it evaluates the LJ field at an
    (M, 3) array of points on the device; ``generate_effective_potential``
    reads its parameters and runs the whole orientational average in one
    kernel instead."""

    def __init__(self, FF_dict, sigmaff, epsilonff, host_syst, cutoff=12 * U.angstrom, images=False):
        self.FF_dict = FF_dict
        self.sigmaff = float(sigmaff)
        self.epsilonff = float(epsilonff)
        self.host_syst = host_syst
        self.cutoff = float(cutoff)
        self.images = bool(images)

    def __call__(self, coords):
        return get_external_potential(coords, self.FF_dict, self.sigmaff, self.epsilonff, self.host_syst,
                                      cutoff=self.cutoff, images=self.images)


def get_external_potential_dict(pars_file_host, pars_file_guest, chk_host, cutoff=12 * U.angstrom, images=False):
    """{guest ffatype: generator} (extpot_calculator.py:678-701).  The host is
    read from its checkpoint file (or taken as a system object) and centred
    on its mean position; parameter files may also be given as dicts."""
    import copy
    from .system import ChkSystem
    FF_dict = read_pars_file(pars_file_host) if not isinstance(pars_file_host, dict) else dict(pars_file_host)
    FF_guest = read_pars_file(pars_file_guest) if not isinstance(pars_file_guest, dict) else dict(pars_file_guest)
    host = ChkSystem.from_file(chk_host) if isinstance(chk_host, (str, bytes)) or hasattr(chk_host, '__fspath__') \
        else copy.copy(chk_host)
    pos = np.asarray(host.pos, dtype=np.float64)
    host.pos = pos - pos.sum(axis=0) / len(pos)
    return {key: ExternalPotentialGenerator(FF_dict, sig, eps, host, cutoff=cutoff, images=images)
            for key, (sig, eps) in FF_guest.items()}


def _guest_system(guest):
    from .system import ChkSystem
    if isinstance(guest, (str, bytes)) or hasattr(guest, '__fspath__'):
        return ChkSystem.from_file(guest)
    return getattr(guest, 'mol', guest)


def _merge_orientations(off, weights, tol=1e-12):
    """Orientations whose site offsets coincide (to ``tol`` times the largest
    offset, in bohr) give the same U_r: keep one of each with the summed
    weight, in first-occurrence order.  For a linear guest along its z axis
    every rotation about z is such a copy, so the 'degree' angles collapse
    to one; sum_r w_r exp(-beta U_r) is unchanged up to rounding."""
    off = np.asarray(off, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64).reshape(-1)
    scale = tol * max(1.0, float(np.abs(off).max()))
    key = np.round(off.reshape(len(off), -1) / scale).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first, kind='stable')             # groups in first-occurrence order
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    wsum = np.zeros(len(order))
    np.add.at(wsum, rank[inv], w)
    return np.ascontiguousarray(off[np.sort(first)]), wsum


def _effpot_setup(guest, epot_generator_dict, rotations):
    """Device inputs of cmm_effpot: host table, guest types and parameters,
    the (nrot, ng, 3) offsets COM + R_r (x_i - COM) and their largest norm."""
    gens = [epot_generator_dict[guest.ffatypes[int(t)]] for t in np.asarray(guest.ffatype_ids)]
    interp = all(isinstance(g, InterpolatorMethod) for g in gens)
    if not interp and not all(isinstance(g, ExternalPotentialGenerator) for g in gens):
        raise TypeError('the device orientational average needs the generators of get_external_potential_dict '
                        'or the interpolators of get_interpolator_dict')
    g0 = gens[0]
    if not interp:
        for g in gens[1:]:
            if (g.host_syst is not g0.host_syst or g.FF_dict is not g0.FF_dict or g.cutoff != g0.cutoff
                    or g.images != g0.images):
                raise ValueError('all guest-atom generators must share one host, parameter set and cut-off')
    ids = [int(t) for t in np.asarray(guest.ffatype_ids)]
    used = sorted(set(ids))
    tmap = {t: k for k, t in enumerate(used)}
    gtype = [tmap[t] for t in ids]
    if interp:
        gsig = [epot_generator_dict[guest.ffatypes[t]] for t in used]      # the interpolators, per type
        geps = None
    else:
        gsig = [epot_generator_dict[guest.ffatypes[t]].sigmaff for t in used]
        geps = [epot_generator_dict[guest.ffatypes[t]].epsilonff for t in used]
    if guest.masses is None:
        raise ValueError('the guest needs atomic masses for its centre of mass')
    pos = np.asarray(guest.pos, dtype=np.float64).reshape(-1, 3)
    masses = np.asarray(guest.masses, dtype=np.float64).reshape(-1, 1)
    com = np.sum(pos * masses, axis=0) / np.sum(masses)
    rel = pos - com
    off = np.einsum('rij,nj->rni', np.asarray(rotations, dtype=np.float64), rel) + com     # (nrot, ng, 3)
    dmax = float(np.sqrt((off ** 2).sum(axis=-1)).max()) * (1 + 1e-12)
    return g0, gtype, gsig, geps, np.ascontiguousarray(off), dmax


def _brick_order(shape, bx=4, by=8, bz=8):
    """Permutation that walks an (nx, ny, nz) point array brick by brick, so
    that 256 consecutive points of the list are a compact 4 x 8 x 8 brick."""
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    nby, nbz = -(-ny // by), -(-nz // bz)
    key = (((ix // bx) * nby + iy // by) * nbz + iz // bz) * (bx * by * bz) + \
        ((ix % bx) * by + iy % by) * bz + iz % bz
    return np.argsort(key.ravel(), kind='stable')


def effective_potential_vectorized(guest, position_shifts, epot_generator_dict, beta, rotations, weights,
                                   limit_potential=1e4 * U.kjmol):
    """-(1/beta) log sum_r w_r exp(-beta sum_i V_i(p + COM + R_r (x_i - COM)))
    at the (m, 3) points ``position_shifts`` (extpot_calculator.py:572-613);
    NaN -> ``limit_potential``.  One device kernel (cmm_effpot), FP64
    throughout (the reference holds the per-atom values in float32)."""
    torch = torch_mod()
    guest = _guest_system(guest)
    g0, gtype, gsig, geps, off, dmax = _effpot_setup(guest, epot_generator_dict, rotations)
    off, w = _merge_orientations(off, weights)
    was_tensor = isinstance(position_shifts, torch.Tensor)
    P = as_device(position_shifts).reshape(-1, 3).contiguous()
    M = P.shape[0]
    out = torch.empty(M, dtype=torch.float64, device='cuda')
    wdev = as_device(w)
    _launch_effpot(None, P, M, g0, gtype, gsig, geps, as_device(off), wdev, len(w), dmax, beta,
                   limit_potential, out)
    return out if was_tensor else out.cpu().numpy()


def _launch_effpot(npoints, P, M, g0, gtype, gsig, geps, offdev, wdev, nrot, dmax, beta, limit, out, rvecs=None):
    """One launch of the direct (LJ generators) or interpolated (interpolator
    methods: ``gsig`` holds them, ``geps`` is None) orientational average."""
    if geps is None:
        tabs = [m.table() for m in gsig]
        _lib.effpot_interp(npoints, P, M, rvecs, gtype, tabs, [m.interp.n for m in gsig],
                           [m.interp.origin for m in gsig], [m.interp.spacing for m in gsig],
                           [m.mode for m in gsig], offdev, wdev, nrot, beta, limit, out)
        return
    atoms = as_device(host_atom_table(g0.FF_dict, g0.host_syst))
    cell = rvecs if rvecs is not None else np.asarray(g0.host_syst.cell.rvecs, dtype=np.float64)
    _lib.effpot(npoints, P, M, cell, atoms, atoms.shape[0], gtype, gsig, geps, offdev, wdev, nrot, dmax, beta,
                g0.cutoff, g0.images, limit, out)


def generate_effective_potential(guest_chk_fn, points, beta, epot_generator_dict, degree=11, max_size=None,
                                 limit_potential=1e4 * U.kjmol, as_tensor=False):
    """Orientationally averaged external potential of a non-spherical guest
    (extpot_calculator.py:615-646) with ``degree`` Lebedev directions x
    ``degree`` angles about the guest's z axis.  ``points``: a ``Grid``
    (grid kernel), or an (nx, ny, nz, 3) / (M, 3) array or tensor of points;
    the result has the points' shape without the last axis.  ``max_size``
    (the reference's host-memory batching) does not change the result and is
    ignored: the kernel keeps no per-orientation arrays."""
    from .rotations import combined_rotations
    torch = torch_mod()
    guest = _guest_system(guest_chk_fn)
    rot, w = combined_rotations(degree)
    g0, gtype, gsig, geps, off, dmax = _effpot_setup(guest, epot_generator_dict, rot)
    off, w = _merge_orientations(off, w)
    wdev = as_device(w)
    offdev = as_device(off)
    if hasattr(points, 'npoints'):
        grid = points
        out = torch.empty(grid.npts, dtype=torch.float64, device='cuda')
        _launch_effpot(grid.npoints, None, 0, g0, gtype, gsig, geps, offdev, wdev, len(w), dmax, beta,
                       limit_potential, out, rvecs=np.asarray(grid.cell.rvecs, dtype=np.float64))
        out = out.reshape(tuple(grid.npoints))
        return out if as_tensor else out.cpu().numpy()
    was_tensor = isinstance(points, torch.Tensor)
    shape = tuple(points.shape[:-1])
    P = as_device(points).reshape(-1, 3)
    perm = None
    if len(shape) == 3:
        perm = torch.from_numpy(_brick_order(shape)).to(P.device)
        P = P[perm]
    P = P.contiguous()
    M = P.shape[0]
    val = torch.empty(M, dtype=torch.float64, device='cuda')
    _launch_effpot(None, P, M, g0, gtype, gsig, geps, offdev, wdev, len(w), dmax, beta, limit_potential, val)
    if perm is not None:
        out = torch.empty_like(val)
        out[perm] = val
    else:
        out = val
    out = out.reshape(shape)
    return out if (as_tensor or was_tensor) else out.cpu().numpy()


# ---------------------------------------------------------------------------
# Interpolated site potentials (extpot_calculator.py:264-346, 391-548, 649-675)
# ---------------------------------------------------------------------------


def get_external_potential_derivatives(points, FF_dict, sigmaff, epsilonff, host_syst, spacings,
                                       cutoff=12 * U.angstrom, images=False, as_tensor=False):
    """(8, M) field and derivatives of the LJ sum (extpot_calculator.py:264-346):
    V, dV/dx dx, dV/dy dy, dV/dz dz, d2V/dxdy dx dy, d2V/dxdz dx dz,
    d2V/dydz dy dz, d3V/dxdydz dx dy dz (``spacings`` = dx, dy, dz, the
    "unit cube" scaling); V and the first derivatives clipped to
    +-1e6 kJ/mol, the higher ones zeroed where V exceeds it.  ``points``: a
    ``Grid`` (grid kernel; result (8, nx, ny, nz)) or (M, 3) points."""
    torch = torch_mod()
    atoms = as_device(host_atom_table(FF_dict, host_syst))
    vmax = 1e6 * U.kjmol
    if hasattr(points, 'npoints'):
        grid = points
        out = torch.empty((8,) + tuple(grid.npoints), dtype=torch.float64, device='cuda')
        _lib.lj_derivs(grid.npoints, None, 0, np.asarray(grid.cell.rvecs, dtype=np.float64), atoms, atoms.shape[0],
                       sigmaff, epsilonff, cutoff, int(images), spacings, vmax, out)
        return out if as_tensor else out.cpu().numpy()
    was_tensor = isinstance(points, torch.Tensor)
    P = as_device(points).reshape(-1, 3).contiguous()
    out = torch.empty((8, P.shape[0]), dtype=torch.float64, device='cuda')
    _lib.lj_derivs(None, P, P.shape[0], np.asarray(host_syst.cell.rvecs, dtype=np.float64), atoms, atoms.shape[0],
                   sigmaff, epsilonff, cutoff, int(images), spacings, vmax, out)
    return out if (as_tensor or was_tensor) else out.cpu().numpy()


class InterpolatorMethod:
    """A bound interpolation method of an ``Interpolator`` (what the
    reference stores in its interpolator dictionary, ``getattr(interpolator,
    int_method)``): callable on (N, 3) points, and readable by the device
    orientational average."""

    def __init__(self, interp, mode, estimate):
        self.interp = interp
        self.mode = mode            # 0 tricubic, 1 trilinear
        self.estimate = estimate    # derivative channels by central differences

    def table(self):
        return self.interp._table(self.estimate)

    def __call__(self, positions):
        torch = torch_mod()
        was_tensor = isinstance(positions, torch.Tensor)
        P = as_device(positions).reshape(-1, 3).contiguous()
        out = torch.empty(P.shape[0], dtype=torch.float64, device='cuda')
        _lib.interp_points(self.table(), self.interp.n, self.interp.origin, self.interp.spacing, self.mode, P,
                           P.shape[0], out)
        if was_tensor:
            return out
        res = out.cpu().numpy()
        return res if res.shape[0] > 1 else res[0]


class Interpolator:
    """Periodic interpolation of a potential grid (extpot_calculator.py:391-548).

    ``grid_values``: (8, Nx, Ny, Nz) -- V and its derivatives dx, dy, dz, dxdy,
    dxdz, dydz, dxdydz in unit-cube form -- or (Nx, Ny, Nz) V only;
    ``grid_origin`` and ``grid_spacing`` map a position x to the fractional
    index (x - origin) / spacing.  ``tricubic`` (exact derivative channels;
    central differences when only V is given), ``tricubic_estimated``
    (central differences of V always) and ``trilinear`` are the reference's
    three methods; tables live on the device in channel-fastest layout."""

    def __init__(self, grid_values, grid_origin, grid_spacing):
        torch = torch_mod()
        vals = grid_values if isinstance(grid_values, torch.Tensor) else np.asarray(grid_values, dtype=np.float64)
        if vals.ndim not in (3, 4) or (vals.ndim == 4 and vals.shape[0] != 8):
            raise ValueError('grid values must be (8, Nx, Ny, Nz) or (Nx, Ny, Nz), got %s' % (tuple(vals.shape),))
        self.grid_values = vals
        self.origin = np.array(grid_origin, dtype=np.float64).reshape(3)
        self.spacing = np.array(grid_spacing, dtype=np.float64).reshape(3)
        self.n = [int(v) for v in vals.shape[-3:]]
        self._tabs = {}

    def _table(self, estimate):
        torch = torch_mod()
        est = bool(estimate or self.grid_values.ndim == 3)
        if est not in self._tabs:
            v = as_device(self.grid_values).contiguous()
            if v.dim() == 4 and est:
                v = v[0].contiguous()
            tab = torch.empty(tuple(self.n) + (8,), dtype=torch.float64, device='cuda')
            _lib.interp_table(v, self.n, int(est), tab)
            self._tabs[est] = tab
        return self._tabs[est]

    @property
    def tricubic(self):
        return InterpolatorMethod(self, 0, False)

    @property
    def tricubic_estimated(self):
        return InterpolatorMethod(self, 0, True)

    @property
    def trilinear(self):
        return InterpolatorMethod(self, 1, False)

    def tricubic_interpolation(self, positions, estimate_derivatives=False):
        return InterpolatorMethod(self, 0, estimate_derivatives)(positions)

    def trilinear_interpolation(self, positions):
        return InterpolatorMethod(self, 1, False)(positions)


def get_interpolator_dict(grid_values_fn_dict, grid_origin, grid_spacing, int_method='tricubic'):
    """{key: interpolation method} (extpot_calculator.py:649-675); the values
    of ``grid_values_fn_dict`` are .npy file names (or arrays)."""
    out = {}
    for key, src in grid_values_fn_dict.items():
        vals = np.load(src, allow_pickle=False) if isinstance(src, (str, bytes)) or hasattr(src, '__fspath__') \
            else src
        out[key] = getattr(Interpolator(vals, grid_origin, grid_spacing), int_method)
    return out
