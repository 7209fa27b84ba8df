"""CPU oracle: an FP64 NumPy/SciPy restatement of the reference cDFT hot path.

THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  It is the checker that the HIP
path is compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``cmmdft_amd`` never does.

It mirrors the reference's arithmetic and operation order (scipy complex
``fftn``/``ifftn`` with the reference's phase factor, two-branch ``np.where``
kernels, SLSQP-constrained Anderson mixing) so that it is also a faithful CPU
timing baseline.  Every function cites the reference file:line it restates
(paths relative to the reference repository root).

Pinning (what checks each function independently of this file; the tests named
live in ``tests/test_oracle_*.py``):

* Grid / FFT ........... numpy.fft + brute-force DFT on 8^3..15^3 grids, the
                          phase algebra, Parseval, round trips.
* HS weights ........... closed-form k=0 limits, Roth (2010) appendix B forms
                          evaluated independently, Lanczos closed form.
* HS value/derive ...... bulk limits: Rosenfeld FMT -> Percus-Yevick
                          compressibility, WBII/MFMT -> Carnahan-Starling
                          free energy and excess chemical potential; finite-
                          difference consistency of derive() with value().
* MFA .................. bulk limit c = 2 a rho with a = compute_vdw_a
                          (uniform density), a against the radial integral of
                          the shifted LJ potential, the supercell sampling
                          against a real-space fold, the phase-displaced
                          kernel against a direct sum on a tiny grid.
* WDA / EOS ............ bulk limit (weighted density == density), EOS
                          derivative by finite differences.  MBWR: the 32
                          coefficients equal the reference's literals
                          (fixture with checksum); the equation's critical
                          point is the published T*c = 1.313, rho*c = 0.310,
                          P*c = 0.13; its low-density limit is the LJ second
                          virial coefficient by quadrature; a planted sign
                          flip or swap of any coefficient fails these.
* LJ external potential  single-atom analytic values, brute-force image sum.
* Orientational average  Lebedev tables = the reference's (fixture) and exact
                          for spherical monomials up to their degree; rotation
                          matrices orthonormal with R e_z = the Lebedev point;
                          one-site guest -> the spherical LJ field; high- and
                          low-temperature limits (weighted mean, minimum); an
                          explicit per-orientation loop with the combined
                          matrices written as R1 diag(c - s, s + c, 1).
* EL solvers ........... ideal-gas closed form, bulk fixed point
                          mu = kT ln(Lambda^3 rho) + f_ex'(rho), solver-
                          independence of the fixed point (Picard vs Anderson),
                          exact simplex-QP vs SLSQP.
* mu(p) ................ EOS residual at the root.

Parity unpinned: the unit constants' last digits (third-party ``molmod``,
absent) and the exact SLSQP iteration path (only the converged fixed point is
pinned).
"""
from __future__ import annotations

import math
import os
import time

import numpy as np
import scipy.fft as sfft
import scipy.optimize as opt
from scipy.special import logsumexp

from cmmdft_amd import units as U

# ---------------------------------------------------------------------------
# Grid (cmmdft/system.py:231-332)
# ---------------------------------------------------------------------------


class Grid:
    """Periodic grid centred at the origin (cmmdft/system.py:231-295).

    ``rvecs`` rows are the cell vectors; the spacing uses the column norms
    (system.py:262-266), the points are frac . rvecs with frac in [-1/2, 1/2)
    (system.py:276-282).
    """

    def __init__(self, rvecs, npoints):
        rvecs = np.asarray(rvecs, dtype=np.float64)
        if isinstance(npoints, (int, np.integer)):
            npoints = [int(npoints)] * 3
        self.rvecs = rvecs
        self.npoints = [int(n) for n in npoints]
        self.npts = int(np.prod(self.npoints))
        self.volume = abs(float(np.linalg.det(rvecs)))
        self.spacings = [np.linalg.norm(rvecs[:, a]) / self.npoints[a] for a in range(3)]
        self.dr = self.volume / self.npts
        frac = [np.linspace(-0.5, 0.5, num=n, endpoint=False) for n in self.npoints]
        fg = np.asarray(np.meshgrid(*frac, indexing='ij'))
        pts = np.zeros(self.npoints + [4])
        pts[..., :3] = np.einsum('ab,aijk->ijkb', rvecs, fg)
        pts[..., 3] = np.sqrt((pts[..., :3] ** 2).sum(-1))
        self.points = pts
        k1 = [2 * np.pi * np.fft.fftfreq(self.npoints[a], d=self.spacings[a]) for a in range(3)]
        kg = np.meshgrid(*k1, indexing='ij')
        kp = np.zeros(self.npoints + [4])
        for a in range(3):
            kp[..., a] = kg[a]
        kp[..., 3] = np.sqrt(kp[..., 0] ** 2 + kp[..., 1] ** 2 + kp[..., 2] ** 2)
        self.kpoints = kp
        # system.py:291: sum_a k_a h_a N_a  (= 2*pi*m_a summed)
        self.scalprod = sum(kp[..., a] * self.spacings[a] * self.npoints[a] for a in range(3))
        kcut = 2 * np.pi / np.array(self.spacings)
        self.sigma_lanczos = (np.sinc(kp[..., 0] / kcut[0]) * np.sinc(kp[..., 1] / kcut[1])
                              * np.sinc(kp[..., 2] / kcut[2]))
        self._phase = np.exp(1j * np.pi * self.scalprod)

    def integrate(self, data):                       # system.py:307-308
        return np.sum(data) * self.dr

    def fft(self, rdata):                            # system.py:310-311
        return sfft.fftn(rdata) * self._phase / self.npts

    def fftn(self, rdata):                           # system.py:313-320
        F = sfft.fftn(rdata, axes=(0, 1, 2))
        return F * (self._phase / self.npts)[..., None]

    def ifft(self, fdata):                           # system.py:322-323
        return sfft.ifftn(fdata * np.conj(self._phase)).real * self.npts

    def ifftn(self, fdata):                          # system.py:325-332
        F = sfft.ifftn(fdata * np.conj(self._phase)[..., None], axes=(0, 1, 2))
        return F.real * self.npts


def cubic_cell(a):
    return np.eye(3) * a


# ---------------------------------------------------------------------------
# Guest helpers (cmmdft/system.py:115-117, 154-158)
# ---------------------------------------------------------------------------


def thermal_wavelength(mass, temperature):        # system.py:115-117
    kT = U.boltzmann * temperature
    return U.planck / np.sqrt(2 * np.pi * mass * kT)


def bh_radius(sigma, epsilon, temperature):       # system.py:154-158
    beta = 1 / (U.boltzmann * temperature)
    Tt = 1 / beta / epsilon
    return sigma * (1 + 0.2977 * Tt) / (1 + 0.33163 * Tt + 0.0010477 * Tt ** 2) / 2


# ---------------------------------------------------------------------------
# Equations of state, pointwise (cmmdft/eos.py)
# ---------------------------------------------------------------------------


class EOS:
    """Base protocol (eos.py:22-79): per-particle excess free energy a(rho)
    and its derivative; per-volume quantities derived from them."""

    def __init__(self, mass):
        self.mass = mass
        self.temperature = None

    def set_temperature(self, T):                  # eos.py:28-30
        self.temperature = T
        self.wvl = U.planck / np.sqrt(2 * np.pi * self.mass * U.boltzmann * T)

    def fex_volume(self, rho):                     # eos.py:47-49
        return rho * self.fex_particle(rho)

    def dfex_volume(self, rho):                    # eos.py:59-63
        return rho * self.dfex_particle(rho) + self.fex_particle(rho)

    def chempot(self, rho):                        # eos.py:32-34
        kT = U.boltzmann * self.temperature
        return kT * np.log(self.wvl ** 3 * rho) + self.dfex_volume(rho)

    def pressure(self, rho):                       # eos.py:39-41
        kT = U.boltzmann * self.temperature
        return kT * rho + rho ** 2 * self.dfex_particle(rho)

    def rough_grid(self, n):                       # eos.py:81-83
        return np.logspace(-10, 0, n) / U.angstrom ** 3

    def densities_from_pressures(self, pressures, n_rough=10000):   # eos.py:122-157
        grid = self.rough_grid(n_rough)
        kT = U.boltzmann * self.temperature
        pgrid = kT * grid + grid ** 2 * self.dfex_particle(grid)
        out = np.full((len(pressures), 3), np.nan)
        for i, p in enumerate(pressures):
            sols = []
            f = lambda r: kT * r + r ** 2 * self.dfex_particle(r) - p
            for j in range(1, n_rough):
                if pgrid[j - 1] <= p <= pgrid[j]:
                    sols.append(opt.brentq(f, grid[j - 1], grid[j]))
            if len(sols) > 3:
                raise ValueError('more than 3 density branches')
            out[i, :len(sols)] = sorted(sols)
        return out

    def densities_from_chempots(self, chempots, n_rough=1000):     # eos.py:85-120
        grid = self.rough_grid(n_rough)
        kT = U.boltzmann * self.temperature
        mugrid = kT * np.log(self.wvl ** 3 * grid) + self.fex_particle(grid) + grid * self.dfex_particle(grid)
        out = np.full((len(chempots), 2), np.nan)
        for i, mu in enumerate(chempots):
            sols = []
            f = lambda r: kT * np.log(self.wvl ** 3 * r) + self.fex_particle(r) + r * self.dfex_particle(r) - mu
            for j in range(1, n_rough):
                if mugrid[j - 1] <= mu <= mugrid[j]:
                    sols.append(opt.brentq(f, grid[j - 1], grid[j]))
            out[i, :len(sols)] = sorted(sols)[:2]
        return out

    def calculate_mu(self, T, p):                  # eos.py:209-215
        self.set_temperature(T)
        rho = self.densities_from_pressures([p])[0][0]
        return self.chempot(rho)


class MFAEOS(EOS):                                 # eos.py:523-546
    def __init__(self, mass, sigma=None, epsilon=None, a=None):
        super().__init__(mass)
        self.a = a if a is not None else -16 / 9 * np.pi * epsilon * sigma ** 3

    def fex_particle(self, rho):
        return self.a * rho

    def dfex_particle(self, rho):
        return self.a + 0.0 * rho


class CarnahanStarlingEOS(EOS):                    # eos.py:471-520
    def __init__(self, mass, Rhs):
        super().__init__(mass)
        self.R = float(Rhs)
        self.eta = 4 * np.pi * self.R ** 3 / 3

    def rough_grid(self, n):                       # eos.py:500-504
        end = np.log(U.angstrom ** 3 / self.eta) / np.log(10) - 0.01
        return np.logspace(-10, end, n) / U.angstrom ** 3

    def fex_particle(self, rho):
        kT = U.boltzmann * self.temperature
        x = self.eta * rho
        return kT * (4 * x - 3 * x ** 2) / (1 - x) ** 2

    def dfex_particle(self, rho):
        kT = U.boltzmann * self.temperature
        x = self.eta * rho
        return 2 * kT * self.eta * (2 - x) / (1 - x) ** 3


class MBWREOS(EOS):
    """Modified Benedict-Webb-Rubin LJ EOS (Johnson, Zollweg & Gubbins 1993),
    restating eos.py:295-447 (coefficient table eos.py:313-345)."""

    X = [0.8623085097507421, 2.976218765822098, -8.402230115796038, 0.1054136629203555,
         -0.8564583828174598, 1.582759470107601, 0.7639421948305453, 1.753173414312048,
         2.798291772190376e+3, -4.8394220260857657e-2, 0.9963265197721935, -3.698000291272493e+1,
         2.084012299434647e+1, 8.305402124717285e+1, -9.574799715203068e+2, -1.477746229234994e+2,
         6.398607852471505e+1, 1.603993673294834e+1, 6.805916615864377e+1, -2.791293578795945e+3,
         -6.245128304568454, -8.116836104958410e+3, 1.488735559561229e+1, -1.059346754655084e+4,
         -1.131607632802822e+2, -8.867771540418822e+3, -3.986982844450543e+1, -4.689270299917261e+3,
         2.593535277438717e+2, -2.694523589434903e+3, -7.218487631550215e+2, 1.721802063863269e+2]
    gamma = 3.0

    def __init__(self, mass, sigma, epsilon):
        super().__init__(mass)
        self.sigma = sigma
        self.epsilon = epsilon

    def rough_grid(self, n):                       # eos.py:449-451
        return np.logspace(-10, 0, n) * 1.5 / self.sigma ** 3

    def set_temperature(self, T):                  # eos.py:347-370
        super().set_temperature(T)
        x = [None] + self.X
        Tr = U.boltzmann * T / self.epsilon
        self.a_coef = [
            x[1] * Tr + x[2] * np.sqrt(Tr) + x[3] + x[4] / Tr + x[5] / Tr ** 2,
            x[6] * Tr + x[7] + x[8] / Tr + x[9] / Tr ** 2,
            x[10] * Tr + x[11] + x[12] / Tr,
            x[13],
            x[14] / Tr + x[15] / Tr ** 2,
            x[16] / Tr,
            x[17] / Tr + x[18] / Tr ** 2,
            x[19] / Tr ** 2,
        ]
        self.b_coef = [
            x[20] / Tr ** 2 + x[21] / Tr ** 3,
            x[22] / Tr ** 2 + x[23] / Tr ** 4,
            x[24] / Tr ** 2 + x[25] / Tr ** 3,
            x[26] / Tr ** 2 + x[27] / Tr ** 4,
            x[28] / Tr ** 2 + x[29] / Tr ** 3,
            x[30] / Tr ** 2 + x[31] / Tr ** 3 + x[32] / Tr ** 4,
        ]

    def _G(self, rho):                             # eos.py:372-382
        rr = rho * self.sigma ** 3
        F = np.exp(-self.gamma * rr ** 2)
        ig = 1.0 / (2 * self.gamma)
        G = [ig * (1 - F)]
        for n in range(2, 7):
            G.append(-ig * (F * rr ** (2 * (n - 1)) - 2 * (n - 1) * G[-1]))
        return G

    def fex_particle(self, rho):                   # eos.py:397-411
        rr = rho * self.sigma ** 3
        Ar = 0.0
        for i, ai in enumerate(self.a_coef):
            Ar = Ar + ai / (i + 1) * rr ** (i + 1)
        for bi, Gi in zip(self.b_coef, self._G(rho)):
            Ar = Ar + bi * Gi
        return Ar * self.epsilon

    def dfex_particle(self, rho):                  # eos.py:413-424
        rr = rho * self.sigma ** 3
        dA = 0.0
        for i, ai in enumerate(self.a_coef):
            dA = dA + ai * rr ** i * self.sigma ** 3
        F = np.exp(-self.gamma * rr ** 2)
        for t, bi in enumerate(self.b_coef):
            dA = dA + bi * self.sigma ** 3 * rr ** (2 * t + 1) * F
        return dA * self.epsilon


class SumOfEOS(EOS):                               # eos.py:226-265
    def __init__(self, mass, eos_list, factors=None):
        super().__init__(mass)
        self.eos_list = list(eos_list)
        self.factors = list(factors) if factors is not None else [1.0] * len(eos_list)

    def set_temperature(self, T):
        for e in self.eos_list:
            e.set_temperature(T)
        super().set_temperature(T)

    def fex_particle(self, rho):
        return sum(f * e.fex_particle(rho) for e, f in zip(self.eos_list, self.factors)) + 0.0 * rho

    def dfex_particle(self, rho):
        return sum(f * e.dfex_particle(rho) for e, f in zip(self.eos_list, self.factors)) + 0.0 * rho

    def fex_volume(self, rho):
        return sum(f * e.fex_volume(rho) for e, f in zip(self.eos_list, self.factors)) + 0.0 * rho

    def dfex_volume(self, rho):
        return sum(f * e.dfex_volume(rho) for e, f in zip(self.eos_list, self.factors)) + 0.0 * rho


# ---------------------------------------------------------------------------
# Hard-sphere FMT (cmmdft/functionals.py:379-767)
# ---------------------------------------------------------------------------

FMT_NAMES = ['FMT', 'aFMT', 'tFMT', 'atFMT', 'taFMT']
MFMT_NAMES = ['MFMT', 'aMFMT', 'tMFMT', 'atMFMT', 'taMFMT']
WBII_NAMES = ['WBII', 'aWBII', 'tWBII', 'atWBII', 'taWBII']


def hs_weights(grid, R, version):
    """k-space FMT weights (functionals.py:421-487)."""
    omega = grid.kpoints[..., 3] * R
    nz = ~np.isclose(omega, 0)
    sl = grid.sigma_lanczos
    w0 = np.ones_like(omega, dtype=np.complex128)
    w0[nz] = np.sinc(omega[nz] / np.pi)
    w0 *= sl
    w1 = R * w0
    w2 = 4.0 * np.pi * R ** 2 * w0
    j2 = np.ones_like(omega, dtype=np.complex128)
    o = omega[nz]
    j2[nz] = 3 * (np.sin(o) - o * np.cos(o)) / o ** 3
    j2 *= sl
    w3 = (4.0 * np.pi * R ** 3 / 3.0) * j2
    wv2 = -1j * w3[..., None] * grid.kpoints[..., :3]
    wv2[~nz] = 0.0
    wv1 = wv2 / (4 * np.pi * R)
    out = {'scalar': [w0, w1, w2, w3], 'vector': [wv1, wv2]}
    if 't' in version:                              # functionals.py:457-487
        kx, ky, kz, K = (grid.kpoints[..., i] for i in range(4))
        K2 = K ** 2
        with np.errstate(invalid='ignore', divide='ignore'):
            H = [np.where(K > 0, kx * kx / K2, 1 / 3), np.where(K > 0, kx * ky / K2, 0.0),
                 np.where(K > 0, kx * kz / K2, 0.0), np.where(K > 0, ky * ky / K2, 1 / 3),
                 np.where(K > 0, ky * kz / K2, 0.0), np.where(K > 0, kz * kz / K2, 1 / 3)]
        B = -4 * np.pi * R ** 2 * (j2 - w0)
        diag = [1 / 3, 0.0, 0.0, 1 / 3, 0.0, 1 / 3]
        out['tensor'] = [B * (h - d) for h, d in zip(H, diag)]
    return out


def phi1(n3, ln3):                                  # functionals.py:720-721
    return -ln3


def dphi1(n3):                                      # functionals.py:723-724
    return 1.0 / (1.0 - n3)


def phi2(n3, ln3, n3_2, version):                  # functionals.py:726-732
    if version in FMT_NAMES or version in MFMT_NAMES:
        return 1 / (1.0 - n3)
    return np.where(n3 <= 1e-8, (1 + n3_2 / 9) / (1 - n3),
                    (5 * n3 - n3_2 + 2 * (1 - n3) * ln3) / (3 * (n3 - n3_2)))


def dphi2(n3, ln3, n3_2, version):                 # functionals.py:734-741
    q = 1 - 2 * n3 + n3_2
    if version in FMT_NAMES or version in MFMT_NAMES:
        return 1 / q
    return np.where(n3 <= 1e-8, (1 + 2 * n3 / 9 + n3_2 / 18) / q,
                    -2 * (n3 - 3 * n3_2 + q * ln3) / (3 * n3_2 * q))


def phi3(n3, ln3, n3_2, n3_3, version):            # functionals.py:743-754
    q = 1 - 2 * n3 + n3_2
    if version in FMT_NAMES:
        return 1 / (24 * np.pi * q)
    if version in MFMT_NAMES:
        return np.where(n3 <= 1e-8, (1.0 - 2 * n3 / 9 - n3_2 / 18) / (24 * np.pi * q),
                        (n3 + q * ln3) / (36 * np.pi * n3_2 * q))
    return np.where(n3 <= 1e-8, (1 - 4 * n3 / 9 + n3_2 / 18) / (24 * np.pi * q),
                    -2 * (n3 - 3 * n3_2 + n3_3 + ln3 * q) / ((3 * n3_2) * 24 * np.pi * q))


def dphi3(n3, ln3, n3_2, n3_3, version):           # functionals.py:756-767
    c = 1 - 3 * n3 + 3 * n3_2 - n3_3
    if version in FMT_NAMES:
        return 1 / (12 * np.pi * c)
    if version in MFMT_NAMES:
        return np.where(n3 <= 1e-8, (8 / 3 - 0.5 * n3 - 0.1 * n3_2) / (36 * np.pi * c),
                        -(2 * n3 - 5 * n3_2 + n3_3 + 2 * c * ln3) / (36 * np.pi * n3_3 * c))
    return np.where(n3 <= 1e-8, (7 / 3 - n3 / 2 + n3_2 / 10) / (36 * np.pi * c),
                    (2 * n3 - 5 * n3_2 + 6 * n3_3 - n3_2 * n3_2 + 2 * c * ln3) / (36 * np.pi * n3_3 * c))


class HardSphere:
    """HardSphereFunctional restated (functionals.py:379-618)."""
    name = 'HardSphere'

    def __init__(self, grid, R, version, temperature):
        assert version in FMT_NAMES + MFMT_NAMES + WBII_NAMES
        self.grid, self.R, self.version = grid, R, version
        self.beta = 1 / (U.boltzmann * temperature)
        self.w = hs_weights(grid, R, version)
        self._key = None

    def densities(self, krho):                      # functionals.py:489-532, 534-552
        g, w, v = self.grid, self.w, self.version
        n = [g.ifft(krho * w['scalar'][a]) for a in range(4)]
        n[3] = np.clip(n[3], 1e-30, 0.99)
        nv1 = g.ifftn(krho[..., None] * w['vector'][0])
        nv2 = g.ifftn(krho[..., None] * w['vector'][1])
        xi = None
        if 'a' in v:
            xi = np.clip((nv2[..., 0] ** 2 + nv2[..., 1] ** 2 + nv2[..., 2] ** 2) / (n[2] ** 2 + 1e-16), 0.0, 1)
        n3 = n[3]
        d = dict(n0=n[0], n1=n[1], n2=n[2], n3=n3, ln3=np.log(1 - n3), n3_2=n3 * n3,
                 n3_3=n3 * n3 * n3, nv1=nv1, nv2=nv2, xi=xi, nt=None)
        if 't' in v:
            t = [g.ifft(krho * wt) for wt in w['tensor']]
            xx, xy, xz, yy, yz, zz = t
            tr2 = xx ** 2 + yy ** 2 + zz ** 2 + 2 * (xy ** 2 + xz ** 2 + yz ** 2)
            tr3 = (xx ** 3 + yy ** 3 + zz ** 3 + 3 * (xx * xy * xy + xx * xz * xz + yy * xy * xy + yy * yz * yz
                   + zz * xz * xz + zz * yz * yz) + 6 * xy * xz * yz)
            d['nt'] = [xx, xy, xz, yy, yz, zz, tr2, tr3]
        return d

    def _dens(self, krho):                          # set_density cache, functionals.py:565-572
        if self._key is None or not np.array_equal(krho, self._key):
            self._key = krho
            self._d = self.densities(krho)
        return self._d

    @staticmethod
    def _q3(d, version):
        n2, nv2 = d['n2'], d['nv2']
        v2 = nv2[..., 0] ** 2 + nv2[..., 1] ** 2 + nv2[..., 2] ** 2
        q = n2 ** 3 * (1 - d['xi']) ** 3 if 'a' in version else n2 ** 3 - 3.0 * n2 * v2
        if 't' in version:
            xx, xy, xz, yy, yz, zz, tr2, tr3 = d['nt']
            vx, vy, vz = nv2[..., 0], nv2[..., 1], nv2[..., 2]
            q = q + 4.5 * (xx * vx ** 2 + yy * vy ** 2 + zz * vz ** 2 + 2 * xy * vx * vy + 2 * xz * vx * vz
                           + 2 * yz * vy * vz)
            q = q - 4.5 * tr3
        return q

    def phi(self, d):                               # functionals.py:620-643
        v = self.version
        n3, ln3, n3_2, n3_3 = d['n3'], d['ln3'], d['n3_2'], d['n3_3']
        dot12 = (d['nv1'] * d['nv2']).sum(-1)
        out = d['n0'] * phi1(n3, ln3)
        out = out + (d['n1'] * d['n2'] - dot12) * phi2(n3, ln3, n3_2, v)
        return out + self._q3(d, v) * phi3(n3, ln3, n3_2, n3_3, v)

    def dphi(self, d):
        """Partial derivatives of Phi (functionals.py:645-717)."""
        v = self.version
        n3, ln3, n3_2, n3_3 = d['n3'], d['ln3'], d['n3_2'], d['n3_3']
        n2, nv1, nv2, xi = d['n2'], d['nv1'], d['nv2'], d['xi']
        p2 = phi2(n3, ln3, n3_2, v)
        p3 = phi3(n3, ln3, n3_2, n3_3, v)
        v2 = (nv2 ** 2).sum(-1)
        dn0 = phi1(n3, ln3)
        dn1 = n2 * p2
        if 'a' in v:
            dn2 = d['n1'] * p2 + (3 * n2 ** 2 * (1 + xi) * (1 - xi) ** 2) * p3
        else:
            dn2 = d['n1'] * p2 + 3 * (n2 ** 2 - v2) * p3
        dot12 = (nv2 * nv1).sum(-1)
        dn3 = d['n0'] * dphi1(n3) + (d['n1'] * n2 - dot12) * dphi2(n3, ln3, n3_2, v)
        dn3 = dn3 + self._q3(d, v) * dphi3(n3, ln3, n3_2, n3_3, v)
        dv1 = -nv2 * p2[..., None]
        fac = -6 * n2 * (1 - xi) ** 2 * p3 if 'a' in v else -6 * n2 * p3
        dv2 = -nv1 * p2[..., None] + nv2 * fac[..., None]
        dt = None
        if 't' in v:
            xx, xy, xz, yy, yz, zz, tr2, tr3 = d['nt']
            vx, vy, vz = nv2[..., 0], nv2[..., 1], nv2[..., 2]
            gv = np.stack([2 * (xx * vx + xy * vy + xz * vz), 2 * (xy * vx + yy * vy + yz * vz),
                           2 * (xz * vx + yz * vy + zz * vz)], axis=-1)
            dv2 = dv2 + 4.5 * gv * p3[..., None]
            gt = np.stack([vx * vx - 3 * (xx * xx + xy * xy + xz * xz),
                           (vx * vy - 3 * (xx * xy + yy * xy + xz * yz)) * 2,
                           (vx * vz - 3 * (xx * xz + zz * xz + xy * yz)) * 2,
                           vy * vy - 3 * (yy * yy + xy * xy + yz * yz),
                           (vy * vz - 3 * (yy * yz + zz * yz + xy * xz)) * 2,
                           vz * vz - 3 * (zz * zz + xz * xz + yz * yz)], axis=-1)
            dt = 4.5 * gt * p3[..., None]
        return [dn0, dn1, dn2, dn3], [dv1, dv2], dt

    def get_n3(self, krho):                         # functionals.py:554-556
        return self.grid.ifft(krho * self.w['scalar'][3])

    def derive(self, krho):                         # functionals.py:574-609
        g, w = self.grid, self.w
        d = self._dens(krho)
        ds, dv, dt = self.dphi(d)
        acc = 0.0
        for f, wk in zip(ds, w['scalar']):
            acc = acc + g.fft(f) * wk
        for f, wk in zip(dv, w['vector']):
            kf = g.fftn(f)
            acc = acc - (kf[..., 0] * wk[..., 0] + kf[..., 1] * wk[..., 1] + kf[..., 2] * wk[..., 2])
        if dt is not None:
            kf = g.fftn(dt)
            acc = acc + sum(kf[..., i] * w['tensor'][i] for i in range(6))
        return g.ifft(acc) / self.beta

    def value(self, krho, local=False):             # functionals.py:611-618
        phi = self.phi(self._dens(krho))
        return phi / self.beta if local else self.grid.integrate(phi) / self.beta


class WDAV:
    """WDAVFunctional restated (functionals.py:1116-1184)."""
    name = 'WDA-V'

    def __init__(self, grid, R, eos, temperature):
        self.grid, self.R, self.eos = grid, R, eos
        eos.set_temperature(temperature)
        self.D = 2 * R
        omega = grid.kpoints[..., 3] * self.D
        nz = ~np.isclose(omega, 0)
        kw = np.zeros_like(omega, dtype=np.complex128)
        o = omega[nz]
        kw[nz] = 3 * (np.sin(o) - o * np.cos(o)) / o ** 3
        kw[~nz] = 1.0
        self.kw = kw * grid.sigma_lanczos

    def weighted_density(self, krho):               # functionals.py:1157-1158
        return self.grid.ifft(krho * self.kw)

    def derive(self, krho, wd=None):                # functionals.py:1160-1174
        if wd is None:
            wd = self.weighted_density(krho)
        d = self.eos.dfex_volume(wd)
        return self.grid.ifft(self.grid.fft(d) * self.kw)

    def value(self, krho, wd=None, local=False):    # functionals.py:1176-1184
        if wd is None:
            wd = self.weighted_density(krho)
        phi = self.eos.fex_volume(wd)
        return phi if local else self.grid.integrate(phi)


class ExtPot:
    """ExternalPotential restated (functionals.py:941-1083) for a stored grid
    potential, with the clamp of FreeEnergy.add_external_potential
    (functionals.py:247-251)."""
    name = 'ExtPot'

    def __init__(self, grid, potential, limit_potential=1e4 * U.kjmol):
        V = np.array(potential, dtype=np.float64).reshape(grid.npoints)
        V[~np.isfinite(V)] = limit_potential
        V[V > limit_potential] = limit_potential
        self.grid, self.potential = grid, V

    def derive(self, krho):
        return self.potential

    def value(self, krho, local=False):
        rho = self.grid.ifft(krho)
        return rho * self.potential if local else self.grid.integrate(rho * self.potential)


class MFA:
    """MFAFunctional restated (functionals.py:770-939).  With tail
    corrections the potential lives on the supercell grid (Grid.supercell,
    system.py:297-301: cell vectors with column a times reps[a]) and the
    convolution samples every reps-th supercell wave vector."""
    name = 'MFA'

    def __init__(self, grid, tailcorrections=False, repetitions=(2, 2, 2)):   # :778-800
        self.tailcorrections = tailcorrections
        self.repetitions = list(repetitions)
        if tailcorrections:
            self.small_grid = grid
            r = np.asarray(self.repetitions)
            self.grid = Grid(grid.rvecs * r, list(np.asarray(grid.npoints) * r))
        else:
            self.grid = grid
        self.potential = None
        self.kpotential = None

    def generate_potential_lj(self, sigma, epsilon, rmin=None, limit_potential=0, cutoff=None):   # :878-910
        if rmin is None:
            rmin = sigma
        r = self.grid.points[..., 3]
        pot = np.full(r.shape, limit_potential, dtype=np.float64)
        mask = r > rmin
        x = np.zeros_like(r)
        x[mask] = sigma / r[mask]
        pot[mask] = 4 * epsilon * (x[mask] ** 12 - x[mask] ** 6)
        if cutoff is not None:
            rc_mask = r <= cutoff
            rc6 = (sigma / cutoff) ** 6
            shift = 4 * epsilon * (rc6 ** 2 - rc6)
            pot[rc_mask] -= shift
            pot[~rc_mask] = 0.0
        self.potential = pot
        self.kpotential = self.grid.fft(self.potential) * self.grid.sigma_lanczos

    def load_potential(self, potential):            # :809-812 (array instead of a file)
        self.potential = np.array(potential, dtype=np.float64)
        assert self.grid.points.shape[:3] == self.potential.shape
        self.kpotential = self.grid.fft(self.potential)

    def compute_vdw_a(self):                        # :814-822
        self.a = 0.5 * self.grid.integrate(self.potential)
        return self.a

    def derive(self, krho):                         # :912-926
        if self.tailcorrections:
            r0, r1, r2 = self.repetitions
            return self.small_grid.ifft(krho * self.kpotential[::r0, ::r1, ::r2]) * self.grid.volume
        return self.grid.ifft(krho * self.kpotential) * self.grid.volume

    def value(self, krho, local=False):             # :928-939
        grid = self.small_grid if self.tailcorrections else self.grid
        rho = grid.ifft(krho)
        if local:
            return 0.5 * rho * self.derive(krho)
        return 0.5 * grid.integrate(rho * self.derive(krho))


# ---------------------------------------------------------------------------
# LJ external potential (cmmdft/extpot_calculator.py:156-222)
# ---------------------------------------------------------------------------


def lennard_jones(r, sigma, epsilon, cutoff=12 * U.angstrom):   # extpot_calculator.py:156-172
    r = np.asarray(r, dtype=np.float64)
    inside = r < cutoff
    V = np.zeros_like(r)
    s6 = (sigma / r[inside]) ** 6
    rc6 = (sigma / cutoff) ** 6
    V[inside] = 4 * epsilon * (s6 * s6 - s6) - 4 * epsilon * (rc6 ** 2 - rc6)
    return V


def external_potential(points, atom_sigma, atom_epsilon, pos, rvecs, sigma_f, epsilon_f,
                       cutoff=12 * U.angstrom):
    """Per-atom minimum-image LJ sum over grid points
    (extpot_calculator.py:186-222).  ``atom_sigma``/``atom_epsilon`` are the
    per-atom host parameters (FF_dict looked up through ffatype_ids)."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    V = np.zeros(len(pts))
    X, Y, Z = pts.T
    L = np.linalg.norm(np.asarray(rvecs), axis=1)               # row norms (:202)
    for i in range(len(pos)):
        sm = 0.5 * (atom_sigma[i] + sigma_f)
        em = np.sqrt(atom_epsilon[i] * epsilon_f)
        rx = X - pos[i, 0]
        ry = Y - pos[i, 1]
        rz = Z - pos[i, 2]
        rx = rx - L[0] * np.round(rx / L[0])
        ry = ry - L[1] * np.round(ry / L[1])
        rz = rz - L[2] * np.round(rz / L[2])
        R = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2 + 1e-16)
        V += lennard_jones(R, sm, em, cutoff=cutoff)
    return V


def lattice_images(rvecs, cutoff, extent):
    """Integer lattice translations n (rows) that can bring an atom within
    ``cutoff`` of a point when atoms and points both lie within ``extent``
    of the origin: |n . rvecs| <= 2 extent + cutoff is necessary, so every
    n with |n_a| <= (2 extent + cutoff) / d_a (d_a the spacing of the lattice
    planes normal to reciprocal axis a) is enumerated."""
    R = np.asarray(rvecs, dtype=np.float64)
    vol = abs(np.linalg.det(R))
    d = [vol / np.linalg.norm(np.cross(R[(a + 1) % 3], R[(a + 2) % 3])) for a in range(3)]
    K = [int(np.ceil((2 * extent + cutoff) / d[a])) for a in range(3)]
    rng = [np.arange(-k, k + 1) for k in K]
    return np.array(np.meshgrid(*rng, indexing='ij')).reshape(3, -1).T


def external_potential_images(points, atom_sigma, atom_epsilon, pos, rvecs, sigma_f, epsilon_f,
                              cutoff=12 * U.angstrom):
    """LJ sum over every periodic image of every host atom within the cut-off,
    a brute-force restatement of what the reference's minimum-image sum
    (extpot_calculator.py:186-222) approximates and its periodic-shift
    version (extpot_calculator.py:224-225, unimplemented there) was meant to
    give: V(p) = sum_i sum_n LJ(|p - x_i - n . rvecs|) over R < cutoff with
    R = sqrt(r^2 + 1e-16), Lorentz-Berthelot mixing and the cut-off shift as
    in ``lennard_jones``.  ``points`` is an (nx, ny, nz, 3) grid of an
    orthorhombic cell (each image is then evaluated on the index block of
    the grid inside its cut-off cube only) or any (M, 3) array (all points
    for every image)."""
    pts = np.asarray(points, dtype=np.float64)
    R = np.asarray(rvecs, dtype=np.float64)
    pos = np.asarray(pos, dtype=np.float64)
    extent = max(np.max(np.linalg.norm(pts.reshape(-1, 3), axis=1)), np.max(np.linalg.norm(pos, axis=1)))
    imgs = lattice_images(R, cutoff, extent) @ R
    grid = pts.ndim == 4 and np.allclose(R, np.diag(np.diag(R)))
    if grid:
        ax = [pts[:, 0, 0, 0], pts[0, :, 0, 1], pts[0, 0, :, 2]]
        V = np.zeros(pts.shape[:3])
    else:
        P = pts.reshape(-1, 3)
        V = np.zeros(len(P))
    for i in range(len(pos)):
        sm = 0.5 * (atom_sigma[i] + sigma_f)
        em = np.sqrt(atom_epsilon[i] * epsilon_f)
        for t in imgs:
            q = pos[i] + t
            if grid:
                lo = [np.searchsorted(ax[a], q[a] - cutoff, side='left') for a in range(3)]
                hi = [np.searchsorted(ax[a], q[a] + cutoff, side='right') for a in range(3)]
                if any(h <= l for l, h in zip(lo, hi)):
                    continue
                dx = (ax[0][lo[0]:hi[0]] - q[0])[:, None, None]
                dy = (ax[1][lo[1]:hi[1]] - q[1])[None, :, None]
                dz = (ax[2][lo[2]:hi[2]] - q[2])[None, None, :]
                r = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2 + 1e-16)
                V[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += lennard_jones(r, sm, em, cutoff=cutoff)
            else:
                d = P - q
                r = np.sqrt(np.sum(d * d, axis=1) + 1e-16)
                V += lennard_jones(r, sm, em, cutoff=cutoff)
    return V.reshape(-1)


# ---------------------------------------------------------------------------
# Non-spherical guests: orientationally averaged potential
# (cmmdft/extpot_calculator.py:348-389, 572-646)
# ---------------------------------------------------------------------------

_LEBEDEV_FIXTURE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'tests', 'golden',
                                'lebedev_ref.npz')


def lebedev_table(degree):                          # AngGrid.py:150-169, 256-287
    """The reference's Lebedev points and weights x 4 pi (its tables, kept as a
    test fixture); degrees 3, 5, 7, 9, 11, 13 and 17."""
    with np.load(_LEBEDEV_FIXTURE, allow_pickle=False) as z:
        return z['points_%d' % degree], z['weights_%d' % degree] * 4 * np.pi


def rotation_matrix(degree, dimension):             # extpot_calculator.py:348-377
    if dimension == 2:
        theta = np.linspace(0, 2 * np.pi, degree, endpoint=False)
        c, s = np.cos(theta), np.sin(theta)
        rot_2 = np.array([[c, -s, np.zeros_like(c)], [s, c, np.zeros_like(c)],
                          [np.zeros_like(c), np.zeros_like(c), np.ones_like(c)]])
        return rot_2.transpose(2, 0, 1), 1 / (degree * 4 * np.pi)
    xyz, w = lebedev_table(degree)
    phi1 = np.arctan2(np.sqrt(xyz[:, 1] ** 2 + xyz[:, 0] ** 2), xyz[:, 2])
    phi2 = np.arctan2(xyz[:, 1], xyz[:, 0])
    c1, s1 = np.cos(phi1), np.sin(phi1)
    c2, s2 = np.cos(phi2), np.sin(phi2)
    zeros = np.zeros(len(phi1))
    rot = np.array([[c1 * c2, -s2, s1 * c2], [c1 * s2, c2, s1 * s2], [-s1, zeros, c1]])
    return rot.transpose(2, 0, 1), w


def orientations(degree):                           # extpot_calculator.py:627-631
    R1, weights1 = rotation_matrix(degree, 3)
    R2, weights2 = rotation_matrix(degree, 2)
    combined_rot = np.einsum('aij,bjk->abij', R1, R2).reshape(-1, 3, 3)
    expanded_weights = np.repeat(weights1 * weights2, len(R2))
    return combined_rot, expanded_weights


def effective_potential(points, guest_pos, guest_masses, guest_sigma, guest_epsilon, host_sigma, host_epsilon,
                        host_pos, rvecs, beta, degree=11, cutoff=12 * U.angstrom, limit_potential=1e4 * U.kjmol,
                        float32=True):
    """generate_effective_potential (:615-646) with get_external_potential
    generators and effective_potential_vectorized (:572-613) at the (..., 3)
    ``points``; per guest atom i its LJ parameters guest_sigma[i],
    guest_epsilon[i] (the generator of its type).  ``float32=True`` keeps the
    reference's float32 per-atom array (:595) and hence its float32 sum and
    -beta*pot (a Python-float beta does not promote a float32 array); False
    holds them in float64."""
    gens = [lambda c, a=a: external_potential(c, host_sigma, host_epsilon, host_pos, rvecs, guest_sigma[a],
                                              guest_epsilon[a], cutoff=cutoff) for a in range(len(guest_pos))]
    return effective_potential_gen(points, guest_pos, guest_masses, gens, beta, degree, limit_potential, float32)


def effective_potential_gen(points, guest_pos, guest_masses, gens, beta, degree=11, limit_potential=1e4 * U.kjmol,
                            float32=True):
    """The same average with one generator per guest atom (a function of
    (M, 3) positions, e.g. an Interpolator method)."""
    pts = np.asarray(points, dtype=np.float64)
    shape = pts.shape[:-1]
    position_shifts = pts.reshape(-1, 3)
    rotations, weights = orientations(degree)
    m = position_shifts.shape[0]
    natom = len(guest_pos)
    nrot = rotations.shape[0]
    masses = np.asarray(guest_masses, dtype=np.float64).reshape(natom, 1)
    total_mass = np.sum(masses)
    neutral_pos = np.asarray(guest_pos)[None, :, :] + position_shifts[:, None, :]
    COMs = np.sum(neutral_pos * masses[None, :, :], axis=1) / total_mass
    rel_pos = neutral_pos - COMs[:, None, :]
    rotated = np.einsum('rij,mnj->mnri', rotations, rel_pos) + COMs[:, None, None, :]
    pot = np.zeros((m, natom, nrot), dtype=np.float32 if float32 else np.float64)
    for a in range(natom):
        coords = rotated[:, a, :, :].reshape(m * nrot, 3)
        vals = gens[a](coords)
        pot[:, a, :] = vals.reshape(m, nrot)
    pot = np.sum(pot, axis=1)
    log_sum = logsumexp(-float(beta) * pot, b=weights, axis=1)
    result = -log_sum / beta
    result = np.where(np.isnan(result), limit_potential, result)
    return result.reshape(shape)


def lj_derivatives(points, atom_sigma, atom_epsilon, pos, rvecs, sigma_f, epsilon_f, spacings,
                   cutoff=12 * U.angstrom):                   # extpot_calculator.py:264-346
    """(8, M): V and its derivatives in unit-cube form, with the clamps."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    Vext, dVdx, dVdy, dVdz = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    dVdxy, dVdxz, dVdyz, dVdxyz = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    X, Y, Z = pts.T
    dx, dy, dz = spacings
    L = np.linalg.norm(np.asarray(rvecs), axis=1)
    for i in range(len(pos)):
        sm = 0.5 * (atom_sigma[i] + sigma_f)
        em = np.sqrt(atom_epsilon[i] * epsilon_f)
        rx = X - pos[i, 0]
        ry = Y - pos[i, 1]
        rz = Z - pos[i, 2]
        rx -= L[0] * (rx / L[0]).round()
        ry -= L[1] * (ry / L[1]).round()
        rz -= L[2] * (rz / L[2]).round()
        R = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2 + 1e-16)
        inside = R < cutoff
        V, dV, ddV, dddV = (np.zeros_like(R) for _ in range(4))
        r6 = (sm / R[inside]) ** 6
        r12 = r6 * r6
        rc6 = (sm / cutoff) ** 6
        V[inside] = 4 * em * (r12 - r6) - 4 * em * (rc6 ** 2 - rc6)
        dV[inside] = 24 * em * (r6 - 2 * r12) / R[inside] ** 2
        ddV[inside] = 96 * em * (7 * r12 - 2 * r6) / R[inside] ** 4
        dddV[inside] = 384 * em * (5 * r6 - 28 * r12) / R[inside] ** 6
        Vext += V
        dVdx += dV * rx
        dVdy += dV * ry
        dVdz += dV * rz
        dVdxy += ddV * rx * ry
        dVdxz += ddV * rx * rz
        dVdyz += ddV * ry * rz
        dVdxyz += dddV * rx * ry * rz
    max_value = 1e+6 * U.kjmol
    V_mask = Vext > max_value
    Vext = np.clip(Vext, -max_value, max_value)
    dVdx = np.clip(dVdx, -max_value, max_value)
    dVdy = np.clip(dVdy, -max_value, max_value)
    dVdz = np.clip(dVdz, -max_value, max_value)
    dVdxy[V_mask] = 0.0
    dVdxz[V_mask] = 0.0
    dVdyz[V_mask] = 0.0
    dVdxyz[V_mask] = 0.0
    dVdx *= dx
    dVdy *= dy
    dVdz *= dz
    dVdxy *= dx * dy
    dVdxz *= dx * dz
    dVdyz *= dy * dz
    dVdxyz *= dx * dy * dz
    return np.array([Vext, dVdx, dVdy, dVdz, dVdxy, dVdxz, dVdyz, dVdxyz])


_CORNERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
_DERIVS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def tricubic_coefficients():
    """The 64 x 64 matrix of extpot_calculator.py:20-147: a = C X maps the 8
    channels at the 8 corners (column corner + 8 channel) to the coefficients
    a[i + 4j + 16k] of x^i y^j z^k.  Built here as the inverse of the
    constraint matrix (value and derivatives of each monomial at each
    corner); it is integer-valued and its bytes are checked against the
    reference's table by checksum."""
    def dmono(p, d, x):
        if d > p:
            return 0.0
        c = 1
        for q in range(d):
            c *= p - q
        return c * x ** (p - d)
    B = np.zeros((64, 64))
    for ci, c in enumerate(_CORNERS):
        for di, d in enumerate(_DERIVS):
            for i in range(4):
                for j in range(4):
                    for k in range(4):
                        B[ci + 8 * di, i + 4 * j + 16 * k] = dmono(i, d[0], c[0]) * dmono(j, d[1], c[1]) * \
                            dmono(k, d[2], c[2])
    return np.round(np.linalg.inv(B))


class Interpolator:                                 # extpot_calculator.py:391-548
    def __init__(self, grid_values, grid_origin, grid_spacing):
        self.grid_values = np.asarray(grid_values, dtype=np.float64)
        self.origin = np.array(grid_origin)
        self.spacing = np.array(grid_spacing)
        self.coeff = tricubic_coefficients()

    def _frac(self, positions):
        s = (positions - self.origin) / self.spacing
        ix = np.floor(s).astype(int)
        return ix, s - ix

    def _corners(self, ix, Nx, Ny, Nz):
        ix0, iy0, iz0 = ix[:, 0] % Nx, ix[:, 1] % Ny, ix[:, 2] % Nz
        return [((ix0 + dx) % Nx, (iy0 + dy) % Ny, (iz0 + dz) % Nz) for dx, dy, dz in _CORNERS]

    @staticmethod
    def estimate_all_derivatives(V):
        Vx = (np.roll(V, -1, axis=0) - np.roll(V, 1, axis=0)) / (2)
        Vy = (np.roll(V, -1, axis=1) - np.roll(V, 1, axis=1)) / (2)
        Vz = (np.roll(V, -1, axis=2) - np.roll(V, 1, axis=2)) / (2)
        Vxy = (np.roll(Vx, -1, axis=1) - np.roll(Vx, 1, axis=1)) / (2)
        Vxz = (np.roll(Vx, -1, axis=2) - np.roll(Vx, 1, axis=2)) / (2)
        Vyz = (np.roll(Vy, -1, axis=2) - np.roll(Vy, 1, axis=2)) / (2)
        Vxyz = (np.roll(Vxy, -1, axis=2) - np.roll(Vxy, 1, axis=2)) / (2)
        return np.stack([V, Vx, Vy, Vz, Vxy, Vxz, Vyz, Vxyz], axis=0)

    def trilinear(self, positions):
        positions = np.atleast_2d(positions)
        values = self.grid_values[0] if self.grid_values.ndim == 4 else self.grid_values
        Nx, Ny, Nz = values.shape
        ix, rd = self._frac(positions)
        rx, ry, rz = rd[:, 0], rd[:, 1], rd[:, 2]
        V = np.zeros((positions.shape[0], 8))
        for c, (xi, yi, zi) in enumerate(self._corners(ix, Nx, Ny, Nz)):
            V[:, c] = values[xi, yi, zi]
        return np.sum(V * np.array([(1 - rx) * (1 - ry) * (1 - rz), rx * (1 - ry) * (1 - rz),
                                    (1 - rx) * ry * (1 - rz), rx * ry * (1 - rz), (1 - rx) * (1 - ry) * rz,
                                    rx * (1 - ry) * rz, (1 - rx) * ry * rz, rx * ry * rz]).T, axis=1)

    def tricubic(self, positions, estimate_derivatives=False):
        positions = np.atleast_2d(positions)
        if self.grid_values.ndim == 3:
            values = self.estimate_all_derivatives(self.grid_values)
        elif estimate_derivatives:
            values = self.estimate_all_derivatives(self.grid_values[0])
        else:
            values = self.grid_values
        Nx, Ny, Nz = values.shape[1:]
        N = positions.shape[0]
        ix, rd = self._frac(positions)
        rx, ry, rz = rd[:, 0], rd[:, 1], rd[:, 2]
        X = np.zeros((N, 64))
        for c, (xi, yi, zi) in enumerate(self._corners(ix, Nx, Ny, Nz)):
            for d in range(8):
                X[:, c + d * 8] = values[d, xi, yi, zi]
        result = np.zeros(N)
        a = X @ self.coeff.T
        for i in range(4):
            ui = rx ** i
            for j in range(4):
                vj = ry ** j
                for k in range(4):
                    result += a[:, i + 4 * j + 16 * k] * ui * vj * rz ** k
        return result


def min_distance(points, pos, rvecs):              # program.py:145-177
    """Distance of each point to the nearest atom, the difference vector
    wrapped by yaff's Cell.mic (fractional coordinates rounded to the
    nearest integer)."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    R = np.asarray(rvecs, dtype=np.float64)
    G = np.linalg.inv(R).T                     # reciprocal vectors as rows
    best = np.full(len(pts), np.inf)
    for a in np.asarray(pos, dtype=np.float64).reshape(-1, 3):
        d = pts - a
        f = d @ G.T
        f -= np.round(f)
        d = f @ R
        best = np.minimum(best, np.sqrt(np.sum(d * d, axis=1)))
    return best


def regions(dis, epot, temperature, energy_cutoff=0.55, range_cutoff=3.4 * U.angstrom, mof_cutoff=5):
    """Program.calc_regions (program.py:179-217)."""
    range_mask = dis < range_cutoff
    crit = np.amin(epot) - energy_cutoff * np.amin(epot)
    energy_mask = epot < crit
    mask_mof = epot > mof_cutoff * U.boltzmann * temperature
    mask_site = (energy_mask + range_mask) * (~mask_mof)
    mask_empty = (~energy_mask) * (~range_mask) * (~mask_mof)
    return mask_site, mask_mof, mask_empty


# ---------------------------------------------------------------------------
# Euler-Lagrange solvers (cmmdft/solver.py:18-604)
# ---------------------------------------------------------------------------


class FloatingPointFailure(FloatingPointError):
    pass


class NoSolutionError(Exception):
    pass


class Solver:
    """Generic fixed-point driver (solver.py:18-345).

    ``parts`` is the list of functionals (HardSphere, WDAV, ExtPot, ...);
    ``beta``/``wavelength`` come from the free-energy container
    (functionals.py:55-71).  ``criterion`` may be a list; besides the
    reference's 'riue'/'res'/'der' this restatement also accepts 'maxdrho'
    (max |rho_new - rho| in angstrom^-3), which both this oracle and the HIP
    path implement identically.
    """
    name = 'SOLVER'

    def __init__(self, grid, parts, beta, wavelength, nsteps=500, threshold=1e-6, criterion='RIUE',
                 a_tol=1e-6, r_tol=1e-4, track_energy=True, verbose=False):
        self.grid, self.parts, self.beta, self.wavelength = grid, parts, beta, wavelength
        self.nsteps = nsteps
        crits = criterion if isinstance(criterion, list) else [criterion]
        ths = threshold if isinstance(threshold, list) else [threshold] * len(crits)
        for c in crits:
            assert c.lower() in ['riue', 'res', 'der', 'maxdrho']
        if not isinstance(criterion, list) and criterion.lower() == 'der':
            ths = [1]
        self.criterion = [c.lower() for c in crits]
        self.threshold = ths
        self.a_tol, self.r_tol = a_tol, r_tol
        self.track_energy = track_energy
        self.verbose = verbose
        self.lower_density = 1e-12
        self.iphase = 0
        self.history = []

    # solver.py:90-101
    def _initiate(self, mu):
        self.fugacity = np.exp(self.beta * mu) / self.beta / self.wavelength ** 3
        self.chempot = mu
        self.curr_step = 0
        self.omega0 = None
        self.history = []

    def clip(self, rho):                            # solver.py:146-148
        return np.where(rho < self.lower_density, 1e-30, rho)

    def grho(self, C1):                             # solver.py:128-129
        return self.beta * np.exp(-self.beta * C1) * self.fugacity

    def c1(self, rho, krho=None):                   # solver.py:137-144
        if krho is None:
            krho = self.grid.fft(rho)
        C1 = np.zeros(self.grid.npoints)
        for p in self.parts:
            C1 += p.derive(krho)
        return C1

    def omega(self, rho, krho):                     # solver.py:103-119
        g = self.grid
        N = g.integrate(rho)
        rc = self.clip(rho)
        Fid = g.integrate(rc * (np.log(self.wavelength ** 3 * rc) - 1.0)).real / self.beta
        G = Fid - self.chempot * N
        for p in self.parts:
            G += p.value(krho)
        self.omega0 = G
        return G

    def converged(self, rho_new, krho_new, C1_new, rho, N_new):   # solver.py:167-234
        g = self.grid
        self.IUE = g.integrate(np.abs(rho_new - rho)).real
        self.RIUE = self.IUE / N_new if N_new > 0 else np.nan
        self.maxdrho = float(np.max(np.abs(rho_new - rho))) * U.angstrom ** 3
        if self.track_energy:
            self.omega(rho_new, krho_new)
        ok = True
        self.RES = self.DER = np.nan
        for c, th in zip(self.criterion, self.threshold):
            if c == 'riue':
                val = self.RIUE
            elif c == 'res':
                val = self.RES = np.linalg.norm(self.grho(C1_new) - rho_new) / np.sqrt(N_new) / np.sqrt(g.npts)
            elif c == 'der':
                mask = np.isclose(rho_new, 0)
                dO = np.log(self.wavelength ** 3 * self.clip(rho_new)) / self.beta + C1_new - self.chempot
                val = self.DER = np.linalg.norm((np.abs(rho_new) * self.beta * dO
                                                 / (self.a_tol + self.r_tol * np.abs(rho_new)))[~mask]) \
                    / np.sqrt(g.npts)
            else:
                val = self.maxdrho
            ok = ok and (val < th)
        if self.IUE == 0 and np.isnan(self.RIUE):
            ok = True
        self.history.append((self.curr_step, N_new, self.RIUE, self.maxdrho, self.omega0))
        if self.verbose:
            print('step %4d N=%.10e RIUE=%.3e maxdrho=%.3e' % (self.curr_step, N_new, self.RIUE, self.maxdrho))
        return ok

    def _solve(self, mu, rho):                      # solver.py:236-315
        self._initiate(mu)
        krho = self.grid.fft(rho)
        C1 = self.c1(rho, krho)
        self.omega0 = self.omega(rho, krho)
        converged = False
        N_new, rho_new = None, rho
        for istep in range(self.nsteps):
            self.curr_step = istep + 1
            rho_new, krho_new, C1_new = self.update_rho(rho, krho, C1)
            if not np.all(np.isfinite(rho_new)):
                raise FloatingPointFailure
            N_new = self.grid.integrate(rho_new).real
            if self.converged(rho_new, krho_new, C1_new, rho, N_new):
                converged = True
                break
            rho, C1, krho = rho_new.copy(), C1_new.copy(), krho_new.copy()
        return N_new, rho_new, converged

    def solve(self, mu, rho):                       # solver.py:317-345
        self.correction_factor = 1
        while self.correction_factor >= 1 / 4:
            try:
                return self._solve(mu, rho)
            except FloatingPointError:
                self.correction_factor /= 2
                self.iphase += 1
        self.correction_factor = 1
        raise NoSolutionError('no solution at mu=%g' % mu)


class Picard(Solver):
    """Picard solver (solver.py:347-457)."""
    name = 'PICARD'

    def __init__(self, *a, alpha_mix=0.1, method='hybrid', **kw):
        super().__init__(*a, **kw)
        self.alpha_mix = alpha_mix
        self.correction_factor = 1
        self.thresh = 1 * U.kjmol
        self.update_rho = self.update_rho_hybrid if method == 'hybrid' else self.update_rho_static

    def update_rho_static(self, rho, krho, C1):     # solver.py:392-401
        G = self.grho(C1)
        a = self.alpha_mix * self.correction_factor
        rn = (1.0 - a) * rho + a * G
        rn[rn < 1e-10 / U.angstrom ** 3] = 0.0
        kn = self.grid.fft(rn)
        return rn, kn, self.c1(rn, kn)

    def alpha_max(self, rho, krho, G):              # solver.py:150-165
        hs = [p for p in self.parts if p.name == 'HardSphere'][0]
        kG = self.grid.fft(G)
        m0 = np.max(hs.get_n3(krho)).real
        m1 = np.max(hs.get_n3(kG)).real
        return np.min([abs((1 - m0) / ((m1 - m0) + 1e-16)), 1])

    def update_rho_hybrid(self, rho, krho, C1):     # solver.py:403-457
        prev = self.omega0
        G = self.grho(C1)
        amax = self.alpha_max(rho, krho, G)
        if np.isclose(amax, 0):
            aopt, lo, hi = 0, 0, 0
        else:
            a1 = 0.45 * amax
            r1 = (1 - a1) * rho + a1 * G
            o1 = self.omega(r1, self.grid.fft(r1))
            a2 = 0.9 * amax if o1 <= prev else 0.225 * amax
            r2 = (1 - a2) * rho + a2 * G
            o2 = self.omega(r2, self.grid.fft(r2))
            c, b, a = np.polyfit([0, a1, a2], [prev, o1, o2], 2)
            al = np.linspace(-max(a1, a2) / 4, max(a1, a2), 10000)
            om = a + b * al + c * al ** 2
            aopt = al[np.where(om == np.min(om))[0][0]]
            lo, hi = np.min(om) / U.kjmol, np.max(om) / U.kjmol
        if aopt <= 0:
            if hi - lo > self.thresh:
                def f(alpha):
                    rt = (1 - alpha) * rho + alpha * G
                    return self.omega(rt, self.grid.fft(rt))
                bounds = opt.Bounds(0.01 * amax, 0.9 * amax)
                aopt = opt.minimize(f, [self.alpha_mix * amax], bounds=bounds, method='SLSQP',
                                    options={'ftol': 1e-8}).x
            else:
                aopt = self.alpha_mix * amax
        if aopt <= 0 or np.isclose(aopt, 0):
            aopt = self.alpha_mix * amax
        aopt = float(np.asarray(aopt).ravel()[0])
        rn = (1 - aopt * self.correction_factor) * rho + aopt * self.correction_factor * G
        rn = self.clip(rn)
        kn = self.grid.fft(rn)
        C1n = self.c1(rn, kn)
        self.omega(rn, kn)
        return rn, kn, C1n


def simplex_lsq_slsqp(R):
    """The reference's Anderson coefficient problem: minimise ||sum_i a_i r_i||
    over a in [0,1]^m with sum a = 1, solved by SLSQP (solver.py:581-591).
    ``R`` is (m, npts)."""
    mk = R.shape[0]
    fun = lambda a: np.linalg.norm(np.einsum('i,ij->j', a, R))
    return opt.minimize(fun, np.full(mk, 1 / mk), method='SLSQP', tol=1e-10, bounds=opt.Bounds(0, 1),
                        constraints=opt.LinearConstraint(np.ones(mk), 1, 1)).x


class Anderson(Picard):
    """Anderson / hybrid-Anderson (solver.py:459-604)."""
    name = 'ANDERSON'

    def __init__(self, *a, method='hybridanderson', m=5, damping=0.3, delta=0.2, damping_max=0.8,
                 damping_min=0.01, adaptive_damping=True, damping_factors=(1.5, 0.5), **kw):
        super().__init__(*a, method=method, **kw)
        self.method = method
        self.m, self.damping0 = m, damping
        self.damping_max, self.damping_min = damping_max, damping_min
        self.adaptive, self.factors, self.delta = adaptive_damping, damping_factors, delta
        self.update_rho = self.update_rho_anderson_driver
        # test switch: the Anderson density of these steps is treated as
        # having a non-finite Boltzmann factor (a planted overflow)
        self.force_nonfinite = ()

    def _initiate(self, mu):                        # solver.py:515-525
        super()._initiate(mu)
        n = self.grid.npts
        self.prev_rhos = np.zeros((self.m, n))
        self.prev_G = np.zeros((self.m, n))
        self.and_true = False
        self.it_eps0 = np.nan
        self.damping = self.damping0

    def _damping(self):                             # solver.py:537-544
        r = np.linalg.norm(self.prev_G[-1] - self.prev_rhos[-1])
        rp = np.linalg.norm(self.prev_G[-2] - self.prev_rhos[-2])
        if r < rp:
            self.damping = min(self.damping * self.factors[0], self.damping_max)
        else:
            self.damping = max(self.damping * self.factors[1], self.damping_min)

    def update_rho_anderson_driver(self, rho, krho, C1):   # solver.py:546-579
        prev = self.omega0
        G = self.grho(C1)
        res = np.linalg.norm(G - rho)
        self.prev_rhos = np.roll(self.prev_rhos, -1, axis=0)
        self.prev_rhos[-1] = rho.ravel()
        self.prev_G = np.roll(self.prev_G, -1, axis=0)
        self.prev_G[-1] = G.ravel()
        if self.curr_step == 0:
            self.it_eps = 0
        else:
            self.it_eps = res / self.grid.integrate(rho)
            if self.curr_step < 3:
                self.it_eps0 = self.it_eps
        cond = ('hybrid' not in self.method.lower()) or (self.it_eps <= self.it_eps0 * self.delta
                                                        and self.curr_step > 4) or self.and_true
        if cond:
            try:
                rn, kn, C1n = self.anderson_step()
                Gn = self.grho(C1n)
                if self.curr_step in self.force_nonfinite:
                    Gn = np.full_like(Gn, np.inf)
                self.and_true = True
                if np.isinf(Gn).any() or np.isnan(Gn).any():
                    rn, kn, C1n = self.update_rho_hybrid(rho, krho, C1)
                else:
                    On = self.omega(rn, kn)
                    if On > prev * 0.8:
                        rn, kn, C1n = self.update_rho_hybrid(rho, krho, C1)
            except FloatingPointError:
                rn, kn, C1n = self.update_rho_hybrid(rho, krho, C1)
        else:
            rn, kn, C1n = self.update_rho_hybrid(rho, krho, C1)
        return rn, kn, C1n

    def anderson_step(self):                        # solver.py:581-604
        mk = min(self.curr_step, self.m)
        R = self.prev_G[-mk:] - self.prev_rhos[-mk:]
        al = simplex_lsq_slsqp(R)
        rr = np.einsum('i,ij->j', al, self.prev_rhos[-mk:]).reshape(self.grid.npoints)
        Gr = np.einsum('i,ij->j', al, self.prev_G[-mk:]).reshape(self.grid.npoints)
        if self.adaptive:
            self._damping()
        d = self.damping * self.correction_factor
        rn = (1 - d) * rr + d * Gr
        rn = self.clip(rn)
        kn = self.grid.fft(rn)
        return rn, kn, self.c1(rn, kn)


# ---------------------------------------------------------------------------
# Initial density (cmmdft/program.py:219-306)
# ---------------------------------------------------------------------------


def initial_density(grid, mu, beta, wavelength, V=None, Ninit=None, temperature=None):
    if Ninit is None:                               # program.py:296-305
        Vd = np.zeros(grid.npoints) if V is None else V
        return np.exp(beta * (mu - Vd)) / wavelength ** 3
    if isinstance(Ninit, float):                    # program.py:278-290
        if V is not None:
            return Ninit * np.exp(-np.maximum(V, 0) / U.boltzmann / temperature)
        return np.full(grid.npoints, Ninit)
    return np.asarray(Ninit, dtype=np.float64)


def timed(fn, *a, **kw):
    t0 = time.perf_counter()
    out = fn(*a, **kw)
    return out, time.perf_counter() - t0
