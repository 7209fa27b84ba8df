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

__all__ = ['lennard_jones', 'get_external_potential', 'read_pars_file', 'host_atom_table']


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
