"""Writes tests/golden/oracle_effpot.npz: inputs and expected outputs of the
orientational average for rigid guests, computed by the oracle (the FP64
restatement in oracle/cdft_oracle.py of generate_effective_potential,
get_external_potential_derivatives and the tricubic Interpolator,
cmmdft/extpot_calculator.py:264-646), on small grids of the case-1 host:
CO2 (degree 5) and ethane (degree 3) in float64 and in the reference's
float32 form, an 8-channel LJ derivative table, and tricubic / trilinear
interpolation of it at seeded points.  The reference ships no vectors for
this path; these pin the oracle against later drift and give the GPU tests
fixed numbers.

Run from the repository root:  python tests/golden/make_golden_effpot.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cmmdft_amd import units as U, workloads as W   # noqa: E402
from cmmdft_amd.system import ChkSystem              # noqa: E402
from cmmdft_amd.extpot import read_pars_file         # noqa: E402
from oracle import cdft_oracle as O                  # noqa: E402

INPUTS = os.path.join(ROOT, 'tests', 'golden', 'inputs')
T = 300.0


def guest(name):
    chk = ChkSystem.from_file(os.path.join(INPUTS, '%s_AA.chk' % name))
    ff = read_pars_file(os.path.join(INPUTS, '%s_pars_vdw_AA_trappe.txt' % name))
    gs = np.array([ff[chk.ffatypes[t]][0] for t in chk.ffatype_ids])
    ge = np.array([ff[chk.ffatypes[t]][1] for t in chk.ffatype_ids])
    return chk, gs, ge


def compute():
    host = W.host_case1()
    sig, eps = host.atom_params()
    beta = 1 / (U.boltzmann * T)
    out = {'rvecs': host.rvecs}
    for name, degree, shape in (('co2', 5, (8, 8, 8)), ('c2h6', 3, (6, 7, 5))):
        chk, gs, ge = guest(name)
        og = O.Grid(host.rvecs, shape)
        for f32 in (False, True):
            V = O.effective_potential(og.points[..., :3], chk.pos, chk.masses, gs, ge, sig, eps, host.pos,
                                      host.rvecs, beta, degree=degree, cutoff=W.CUTOFF, float32=f32)
            out['effpot_%s_%s' % (name, 'f32' if f32 else 'f64')] = V
        out['shape_%s' % name] = np.array(shape)
        out['degree_%s' % name] = np.array(degree)
    tshape = (10, 9, 8)
    og = O.Grid(host.rvecs, tshape)
    sf, ef = 3.05 * U.angstrom, 79.0 * U.boltzmann
    tab = O.lj_derivatives(og.points[..., :3], sig, eps, host.pos, host.rvecs, sf, ef, og.spacings,
                           cutoff=W.CUTOFF).reshape((8,) + tshape)
    out['table_shape'] = np.array(tshape)
    out['table'] = tab
    rng = np.random.default_rng(77)
    pts = (rng.random((64, 3)) - 0.5) @ host.rvecs
    origin = og.points[0, 0, 0, :3]
    h = np.array(og.spacings)
    # a smooth, clipped table (the raw LJ table spans 1e6 kJ/mol near atoms)
    smooth = np.clip(tab, -0.05, 0.05)
    it = O.Interpolator(smooth, origin, h)
    out['interp_points'] = pts
    out['interp_origin'] = origin
    out['interp_spacing'] = h
    out['interp_table'] = smooth
    out['tricubic'] = it.tricubic(pts)
    out['trilinear'] = it.trilinear(pts)
    out['tricubic_estimated'] = it.tricubic(pts, estimate_derivatives=True)
    return out


if __name__ == '__main__':
    np.savez(os.path.join(ROOT, 'tests', 'golden', 'oracle_effpot.npz'), **compute())
