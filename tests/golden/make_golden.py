"""Writes tests/golden/oracle_small.npz: inputs and expected outputs of the
cDFT hot path on a small anisotropic grid, computed by the oracle (the FP64
restatement of the reference in oracle/cdft_oracle.py).  The reference ships
no test vectors for this path and cannot be imported here (molmod and yaff
are absent), so these fixtures pin the oracle against later drift and let the
GPU tests compare the HIP path with fixed numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cmmdft_amd import units as U, workloads as W   # noqa: E402
from oracle import cdft_oracle as O                 # noqa: E402

SHAPE = (16, 12, 10)
T = 300.0
VERSIONS = ('FMT', 'WBII', 'aWBII', 'MFMT')


def compute():
    host = W.host_case1()
    og = O.Grid(host.rvecs, SHAPE)
    sig, eps = host.atom_params()
    V = O.external_potential(og.points[..., :3], sig, eps, host.pos, host.rvecs, W.GUEST_SIGMA,
                             W.GUEST_EPSILON).reshape(SHAPE)
    beta = 1 / (U.boltzmann * T)
    lam = O.thermal_wavelength(W.GUEST_MASS, T)
    R = O.bh_radius(W.GUEST_SIGMA, W.GUEST_EPSILON, T)
    rng = np.random.default_rng(2024)
    rho = 2.4e-3 / U.angstrom ** 3 * np.exp(-beta * np.clip(V, -50 * U.kjmol, 50 * U.kjmol))
    rho *= 1 + 0.1 * rng.random(SHAPE)
    krho = og.fft(rho)
    out = {'shape': np.array(SHAPE), 'rvecs': host.rvecs, 'rho': rho, 'V': V, 'R': np.array(R)}
    for v in VERSIONS:
        hs = O.HardSphere(og, R, v, T)
        out['c_' + v] = hs.derive(krho)
        out['F_' + v] = np.array(hs.value(krho))
    wda = O.WDAV(og, R, O.MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON), T)
    out['c_WDA_MFA'] = wda.derive(krho)
    out['F_WDA_MFA'] = np.array(wda.value(krho))
    # converged WBII + MFA-WDA + LJ solution at 10 bar (Picard static, tight)
    mu = O.MBWREOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON).calculate_mu(T, 10 * U.bar)
    parts = [O.HardSphere(og, R, 'WBII', T), wda, O.ExtPot(og, V)]
    rho0 = O.initial_density(og, mu, beta, lam, V=parts[2].potential)
    s = O.Picard(og, parts, beta, lam, nsteps=20000, alpha_mix=0.05, method='static', criterion='RIUE',
                 threshold=1e-13)
    N, rho_c, conv = s.solve(mu, rho0)
    assert conv
    out['mu'] = np.array(mu)
    out['rho_conv'] = rho_c
    out['N_conv'] = np.array(N)
    out['omega_conv'] = np.array(s.omega(rho_c, og.fft(rho_c)))
    return out


if __name__ == '__main__':
    fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'oracle_small.npz')
    np.savez_compressed(fn, **compute())
    print('wrote', fn, os.path.getsize(fn), 'bytes')
