"""Seeded synthetic workloads: the guest, hosts, grids and state points of the
five benchmark cases (methane-like united-atom LJ guest in LJ host frameworks).

Everything is in atomic units (see :mod:`cmmdft_amd.units`).  Seeds are fixed
per case: ``numpy.random.default_rng(1001 + case_index)`` for cases 0-4 and
``1100 + k`` for framework ``k`` of the uneven batch.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import units as U

# Guest: united-atom methane (reference notebook examples/methane_ZIF8.ipynb)
GUEST_SIGMA = 3.73 * U.angstrom
GUEST_EPSILON = 148.0 * U.boltzmann
GUEST_MASS = 16.04 * U.amu
TEMPERATURE = 300.0
CUTOFF = 12.0 * U.angstrom


@dataclass
class SyntheticHost:
    """Duck-types the attributes of a yaff ``System`` that the external
    potential generator reads (cmmdft/extpot_calculator.py:186-222):
    ``pos``, ``ffatype_ids``, ``ffatypes``, ``cell.rvecs``; plus the force-field
    dictionary ``ff`` {ffatype: (sigma, epsilon)}."""
    rvecs: np.ndarray
    pos: np.ndarray
    ffatype_ids: np.ndarray
    ffatypes: list
    ff: dict
    name: str = 'synthetic'
    seed: int = 0
    notes: dict = field(default_factory=dict)

    @property
    def cell(self):
        return _Cell(self.rvecs)

    @property
    def natom(self):
        return len(self.pos)

    def atom_params(self):
        sig = np.array([self.ff[self.ffatypes[t]][0] for t in self.ffatype_ids])
        eps = np.array([self.ff[self.ffatypes[t]][1] for t in self.ffatype_ids])
        return sig, eps


class _Cell:
    def __init__(self, rvecs):
        self.rvecs = np.asarray(rvecs, dtype=np.float64)
        self.nvec = 3

    @property
    def volume(self):
        return abs(float(np.linalg.det(self.rvecs)))

    @property
    def parameters(self):
        lengths = np.linalg.norm(self.rvecs, axis=1)
        return lengths, None


def _rsa_positions(rng, lengths, n, dmin, max_tries=200000):
    """Random sequential placement in an orthorhombic box with a minimum
    minimum-image separation ``dmin``; returns as many as fit within
    ``max_tries`` attempts."""
    L = np.asarray(lengths)
    pos = np.empty((0, 3))
    tries = 0
    while len(pos) < n and tries < max_tries:
        tries += 1
        p = (rng.random(3) - 0.5) * L
        if len(pos):
            d = pos - p
            d -= L * np.round(d / L)
            if np.min(np.sum(d * d, axis=1)) < dmin * dmin:
                continue
        pos = np.vstack([pos, p])
    return pos


def _centre(pos):
    # hosts are centred on the mean atom position (cmmdft/system.py:68-70)
    return pos - pos.sum(axis=0) / len(pos)


def host_case1(seed=1001):
    """64 atoms, cubic a = 24 A, per-atom sigma ~ U(3.0, 3.5) A, eps/kB ~ U(20, 60) K,
    min separation 3.0 A."""
    rng = np.random.default_rng(seed)
    a = 24.0 * U.angstrom
    pos = _rsa_positions(rng, [a] * 3, 64, 3.0 * U.angstrom)
    n = len(pos)
    sig = rng.uniform(3.0, 3.5, n) * U.angstrom
    eps = rng.uniform(20.0, 60.0, n) * U.boltzmann
    types = ['A%d' % i for i in range(n)]
    ff = {t: (s, e) for t, s, e in zip(types, sig, eps)}
    return SyntheticHost(np.eye(3) * a, _centre(pos), np.arange(n), types, ff, 'case1', seed)


TYPE_SIGMA = [2.5, 3.3, 3.4, 2.5]
TYPE_EPS_K = [22.0, 52.0, 35.0, 30.0]


def host_case2(seed=1002, a_angstrom=25.6, natom=276, dmin=2.8):
    """276 atoms of 4 types (t = i mod 4), cubic a = 25.6 A, min separation 2.8 A."""
    rng = np.random.default_rng(seed)
    a = a_angstrom * U.angstrom
    pos = _rsa_positions(rng, [a] * 3, natom, dmin * U.angstrom)
    n = len(pos)
    types = ['T0', 'T1', 'T2', 'T3']
    ff = {t: (s * U.angstrom, e * U.boltzmann) for t, s, e in zip(types, TYPE_SIGMA, TYPE_EPS_K)}
    return SyntheticHost(np.eye(3) * a, _centre(pos), np.arange(n) % 4, types, ff, 'case2', seed)


def host_case4(seed=1004):
    """Case-1 cell and positions, atoms assigned to 4 types round-robin; per-type
    (sigma, eps) drawn from the case-1 ranges with this seed."""
    h1 = host_case1()
    rng = np.random.default_rng(seed)
    sig = rng.uniform(3.0, 3.5, 4) * U.angstrom
    eps = rng.uniform(20.0, 60.0, 4) * U.boltzmann
    types = ['T0', 'T1', 'T2', 'T3']
    ff = {t: (s, e) for t, s, e in zip(types, sig, eps)}
    return SyntheticHost(h1.rvecs.copy(), h1.pos.copy(), np.arange(h1.natom) % 4, types, ff, 'case4', seed)


def smooth5(n):
    """Nearest integer >= 8 whose only prime factors are 2, 3 and 5."""
    def ok(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    best = None
    for d in range(0, 64):
        for c in (n - d, n + d):
            if c >= 8 and ok(c):
                best = c if best is None else best
        if best is not None:
            return best
    raise ValueError(n)


def host_case5(k):
    """Framework k of the uneven batch: orthorhombic edges U(18, 40) A, 4 types
    with case-1 ranges, n ~ U{100..600} capped at 80 % of the RSA capacity."""
    seed = 1100 + k
    rng = np.random.default_rng(seed)
    edges = rng.uniform(18.0, 40.0, 3)
    want = int(rng.integers(100, 601))
    sig = rng.uniform(3.0, 3.5, 4) * U.angstrom
    eps = rng.uniform(20.0, 60.0, 4) * U.boltzmann
    L = edges * U.angstrom
    cap = len(_rsa_positions(np.random.default_rng(seed + 50000), L, 10 ** 6, 2.8 * U.angstrom, max_tries=20000))
    n = min(want, int(0.8 * cap))
    pos = _rsa_positions(rng, L, n, 2.8 * U.angstrom)
    types = ['T0', 'T1', 'T2', 'T3']
    ff = {t: (s, e) for t, s, e in zip(types, sig, eps)}
    npts = [smooth5(int(round(e / 0.25))) for e in edges]
    h = SyntheticHost(np.diag(L), _centre(pos), np.arange(len(pos)) % 4, types, ff, 'case5_%d' % k, seed,
                      notes={'requested_atoms': want, 'rsa_capacity': cap, 'npoints': npts})
    return h


def guest_co2():
    """Rigid three-site CO2 (TraPPE: C-O 1.16 A along z, C 2.80 A / 27.0 K,
    O 3.05 A / 79.0 K, masses 12.011 and 15.999 amu) as a checkpoint-like
    system and its LJ parameters {ffatype: (sigma, epsilon)}."""
    from .system import ChkSystem
    d = 1.16 * U.angstrom
    pos = np.array([[0.0, 0.0, d], [0.0, 0.0, 0.0], [0.0, 0.0, -d]])
    mol = ChkSystem(pos, np.zeros((0, 3)), ['O_CD', 'C_CD'], [0, 1, 0],
                    masses=np.array([15.999, 12.011, 15.999]) * U.amu)
    ff = {'O_CD': (3.05 * U.angstrom, 79.0 * U.boltzmann), 'C_CD': (2.80 * U.angstrom, 27.0 * U.boltzmann)}
    return mol, ff


def pressures_isotherm(n=32, lo=1e-2, hi=1e2):
    return np.logspace(np.log10(lo), np.log10(hi), n) * U.bar


def pressures_uneven(n=8, lo=0.1, hi=50.0):
    """The 8 pressures of each framework of the uneven batch, log-spaced."""
    return np.logspace(np.log10(lo), np.log10(hi), n) * U.bar


def framework_program(host, npoints, version='WBII', images=None):
    """A Program for a synthetic host: united-atom methane at 300 K, FMT
    ``version`` + mean-field WDA + LJ external potential (periodic images
    summed when a cell edge is shorter than twice the cutoff, or as asked)."""
    from .program import Program
    from .system import NanoporousHost, SphericalLJGuest
    from .eos import MFAEOS
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost(host.name, host, host.ff),
                    SphericalLJGuest('CH4_UA', GUEST_MASS, GUEST_SIGMA, GUEST_EPSILON))
    prog.set_grid(npoints=list(npoints))
    prog.init_free_energy(TEMPERATURE)
    if version:
        prog.fener.add_hard_sphere(version)
    prog.fener.add_wdav(MFAEOS(GUEST_MASS, GUEST_SIGMA, GUEST_EPSILON))
    if images is None:
        images = bool(np.min(np.linalg.norm(host.rvecs, axis=1)) < 2 * CUTOFF)
    prog.fener.add_external_potential(cutoff=CUTOFF, images=images)
    return prog


def chempots_at(pressures):
    """mu(p) at 300 K from the MBWR equation of state (repo EOS)."""
    from .eos import ModifiedBenedictWebbRubinEOS
    eos = ModifiedBenedictWebbRubinEOS(GUEST_MASS, GUEST_SIGMA, GUEST_EPSILON)
    return np.array([eos.calculate_mu(TEMPERATURE, p) for p in pressures])
