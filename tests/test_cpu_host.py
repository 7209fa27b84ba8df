"""CPU tests: the C ABI library loads and exports every declared symbol, and
the host-side logic (constructors, validation, file naming, parameter files,
synthetic workloads, EOS state points) behaves like the reference's."""
import os
import re

import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'cmmdft_hip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(cmm_\w+)\s*\(', txt, re.M)))


def test_library_exports_every_declared_symbol():
    from cmmdft_amd import _lib
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.EXPORTS, 'no ctypes prototype for %s' % s
    assert lib.cmm_version() >= 10000


def test_functional_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of cmm_functional against the C compiler's layout
    of the header's struct: size and every member's offset."""
    import ctypes as C
    import shutil
    import subprocess
    from cmmdft_amd._lib import CmmFunctional
    names = [n for n, _ in CmmFunctional._fields_]
    assert names[:5] == ['hs_base', 'hs_antisym', 'hs_tensor', 'use_wda', 'n_eos']
    cc = shutil.which('gcc') or shutil.which('cc')
    if cc is None:
        pytest.skip('no C compiler')
    src = tmp_path / 'layout.c'
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cmmdft_hip.h"\nint main(void) {\n'
                   '  printf("%zu\\n", sizeof(cmm_functional));\n' +
                   ''.join('  printf("%%zu\\n", offsetof(cmm_functional, %s));\n' % n for n in names) +
                   '  return 0;\n}\n')
    exe = tmp_path / 'layout'
    subprocess.run([cc, '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == C.sizeof(CmmFunctional)
    assert out[1:] == [getattr(CmmFunctional, n).offset for n in names]


def test_api_rejects_bad_arguments_without_gpu():
    import ctypes as C
    from cmmdft_amd import _lib
    lib = _lib.load()
    assert lib.cmm_plan_destroy(None) == _lib.CMM_OK
    h = C.c_void_p()
    rv = np.eye(3)
    assert lib.cmm_plan_create(7, 8, 8, rv.ctypes.data_as(C.c_void_p), 1, C.byref(h)) == _lib.CMM_E_ARG  # 7 = prime
    assert 'axis length' in lib.cmm_last_error().decode()
    assert lib.cmm_plan_create(8, 8, 8, None, 1, C.byref(h)) == _lib.CMM_E_ARG
    with pytest.raises(ValueError):
        _lib.check(_lib.CMM_E_ARG)


def test_solver_constructor_validation():
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import NanoporousHost, SphericalLJGuest
    from cmmdft_amd.solver import Anderson, Picard
    host = W.host_case1()
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost('h', host, host.ff), SphericalLJGuest('g', W.GUEST_MASS, W.GUEST_SIGMA,
                                                                           W.GUEST_EPSILON))
    prog.set_grid(npoints=16)
    prog.init_free_energy(300.0)
    prog.fener.add_hard_sphere('WBII')
    with pytest.raises(AssertionError):
        Picard(prog, criterion='bogus')
    with pytest.raises(AssertionError):
        Picard(prog, criterion=['riue', 'res'], threshold=[1e-6])
    s = Anderson(prog, criterion=['riue', 'res'], threshold=1e-7, m=5)
    assert s.threshold == [1e-7, 1e-7] and s.criterion == ['riue', 'res']
    with pytest.raises(AssertionError):
        Anderson(prog, m=9)
    with pytest.raises(AssertionError):
        prog.fener.add_hard_sphere('XYZ')
    # reference default damping and method
    a = Anderson(prog)
    assert a.damping == 0.3 and a.Anderson_method == 'hybridanderson' and a.alpha_mix == 0.1


def test_file_suffix_format():
    from cmmdft_amd.program import get_file_suffix
    assert get_file_suffix(-31.5 * U.kjmol, 300.0) == '-31.50000kJmol_300.00000K'
    assert get_file_suffix([1 * U.kjmol, 2 * U.kjmol], 77.0) == '1.00000kJmol_2.00000kJmol_77.00000K'


def test_read_pars_file(tmp_path):
    from cmmdft_amd.extpot import read_pars_file
    fn = tmp_path / 'pars.txt'
    fn.write_text('# comment\nLJ:UNIT SIGMA angstrom\nLJ:UNIT EPSILON kcalmol\nLJ:SCALE 1 0.0\n'
                  'LJ:PARS      ME     3.730  0.29411\nLJ:PARS  C 3.4 0.1  # trailing\n')
    ff = read_pars_file(str(fn))
    assert np.isclose(ff['ME'][0], 3.730 * U.angstrom) and np.isclose(ff['ME'][1], 0.29411 * U.kcalmol)
    assert set(ff) == {'ME', 'C'}


def test_units_consistency():
    # 148 K expressed as kcal/mol ~ 0.294 (the TraPPE methane epsilon)
    assert abs(148 * U.boltzmann / U.kcalmol - 0.2941) < 2e-3
    assert abs(U.angstrom - 1.8897261) < 1e-6


def test_workload_generators():
    h1 = W.host_case1()
    assert h1.natom == 64
    L = np.linalg.norm(h1.rvecs, axis=1)
    d = h1.pos[:, None, :] - h1.pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(h1.natom) * 1e9
    assert r.min() >= 3.0 * U.angstrom
    assert np.allclose(h1.pos.mean(0), 0, atol=1e-10)
    h2 = W.host_case2()
    assert h2.natom == 276 and set(h2.ffatype_ids) == {0, 1, 2, 3}
    h5 = W.host_case5(3)
    assert all(W.smooth5(n) == n for n in h5.notes['npoints'])
    assert W.smooth5(127) in (125, 128)


def test_product_eos_matches_oracle_state_points():
    from oracle import cdft_oracle as O
    from cmmdft_amd import eos as E
    m, s, e = W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON
    pe, oe = E.ModifiedBenedictWebbRubinEOS(m, s, e), O.MBWREOS(m, s, e)
    for p in [0.01, 1.0, 10.0, 100.0]:
        assert np.isclose(pe.calculate_mu(300.0, p * U.bar), oe.calculate_mu(300.0, p * U.bar), rtol=1e-13)
    R = O.bh_radius(s, e, 300.0)
    pc, oc = E.CarnahanStarlingEOS(m, R), O.CarnahanStarlingEOS(m, R)
    pc.set_temperature(300.0)
    oc.set_temperature(300.0)
    r = np.logspace(-6, -2, 7)
    assert np.allclose(pc.derivative_excess_free_energy_volume(r), oc.dfex_volume(r), rtol=1e-14)
    terms = E.SumOfEOS(m, [pe, pc, E.MFAEOS(m, s, e)], [1, -1, -1])
    terms.set_temperature(300.0)
    kinds = [k for k, f, par in terms.device_terms()]
    facs = [f for k, f, par in terms.device_terms()]
    assert kinds == [E.EOS_MBWR, E.EOS_CS, E.EOS_MFA] and facs == [1, -1, -1]


def test_radix_candidates_multiply_to_length():
    from cmmdft_amd.autotune import radix_candidates
    for n in (8, 24, 64, 72, 96, 100, 128, 160, 250, 256):
        c = radix_candidates(n)
        assert c, n
        assert len(set(c)) == len(c)
        lo = min(len(s) for s in c)
        for s in c:
            assert int(np.prod(s)) == n and set(s) <= {2, 3, 4, 5, 8} and len(s) <= lo + 1
    # 128 = 8*8*2 (3 orders), 8*4*4 (3), 8*4*2*2 (12), 4*4*4*2 (4)
    assert len(radix_candidates(128)) == 22


def test_adsorption_unit_conversion_and_aif_layout(tmp_path):
    """Unit factors from masses and volume (tools.py:368-399): 1 molecule per
    cell of mass M (amu) is 1000/M mol/kg and 100 m/M wt%; AIF text layout
    (tools.py:401-456)."""
    from cmmdft_amd import units as U
    from cmmdft_amd.calculator import convert_units, write_aif
    m_guest, m_host, vol = 16.04 * U.amu, 2000.0 * U.amu, 4000.0 * U.angstrom ** 3
    cv = convert_units(m_guest, m_host, vol)
    assert np.isclose(cv.conversion_factor('au/uc', 'mol/kg'), 1000.0 / 2000.0)
    assert np.isclose(cv.conversion_factor('au/uc', 'wt%'), 100 * 16.04 / 2000.0)
    assert np.isclose(cv.conversion_factor('au/uc', 'mg/g'), 1000 * 16.04 / 2000.0)
    # 'cm3/cm3' as the reference defines it: m_host rho_stp / (m_guest rho_host),
    # the inverse of the STP volume per cell volume of one molecule per cell
    stp = 22414.0 / 6.0221415e23 / 4000.0e-24
    assert np.isclose(cv.conversion_factor('au/uc', 'cm3/cm3'), 1 / stp, rtol=1e-4)
    fn = tmp_path / 'x.aif'
    write_aif(fn, 'blk', [('_a', 'x'), ('_b', ['f1', 'f2'])], '_p_', ['u', 'v'], [['1', '2'], ['3', '4']])
    assert fn.read_text() == "data_blk\n_a x\n_b ['f1', 'f2']\n\nloop_\n_p_u\n_p_v\n1 3 \n2 4 \n"
