"""ctypes binding of libcmmdft_hip.so (C ABI declared in include/cmmdft_hip.h).

The shared library is built in-tree (``__graft_entry__.build()`` or
``cmmdft_amd/csrc/build.sh``).  There is no fallback: if the library or a GPU
is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libcmmdft_hip.so')

CMM_OK, CMM_E_ARG, CMM_E_HIP, CMM_E_NONFINITE, CMM_E_UNSUPPORTED = 0, 1, 2, 3, 4
NSCAL = 32
MMAX = 8

# scalar slots per instance (see cmm_kernels.h)
S_PHI, S_FEX, S_N3MAX = 0, 1, 2
S_NCUR, S_FID, S_EXT = 3, 4, 5
S_PIC_NNEW, S_PIC_IUE, S_PIC_DMAX, S_PIC_NONFIN = 6, 7, 8, 9
S_AND_NONFIN, S_GRAM = 6, 7
S_CHK_NONFIN, S_CHK_RES2, S_CHK_DER2, S_CHK_DERN2 = 6, 7, 8, 9
S_PN, S_PIUE, S_PDMAX, S_PNONFIN, S_PFID, S_PEXT = 16, 17, 18, 19, 20, 21
S_ADJ_PQ, S_ADJ_RR, S_ADJ_LAM = 24, 25, 27
S_FMFA = 28
S_ADAMP = 30


class CmmFunctional(C.Structure):
    _fields_ = [
        ('hs_base', C.c_int), ('hs_antisym', C.c_int), ('hs_tensor', C.c_int), ('use_wda', C.c_int),
        ('n_eos', C.c_int),
        ('eos_kind', C.c_int * 4), ('eos_factor', C.c_double * 4), ('eos_par', (C.c_double * 16) * 4),
        ('R', C.c_double), ('D', C.c_double), ('kT', C.c_double), ('n3_series', C.c_double),
        ('wda_pointwise', C.c_int),
    ]


EXPORTS = {
    'cmm_version': ([], C.c_int),
    'cmm_last_error': ([], C.c_char_p),
    'cmm_plan_create': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    'cmm_plan_destroy': ([C.c_void_p], C.c_int),
    'cmm_plan_set_stream': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_plan_set_tiles': ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    'cmm_plan_get_info': ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    'cmm_plan_set_radix': ([C.c_void_p, C.c_int, C.c_int, C.c_void_p], C.c_int),
    'cmm_plan_set_mfa': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_mfa_lj_potential': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_fold': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_field_reduce': ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_double, C.c_double, C.c_void_p,
                          C.c_void_p, C.c_void_p], C.c_int),
    'cmm_mfa_table': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p], C.c_int),
    'cmm_plan_get_radix': ([C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p], C.c_int),
    'cmm_plan_set_active': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_set_functional': ([C.c_void_p, C.POINTER(CmmFunctional)], C.c_int),
    'cmm_set_instances': ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_fft_r2c': ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_fft_c2r': ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_seed': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double], C.c_int),
    'cmm_derive': ([C.c_void_p], C.c_int),
    'cmm_value': ([C.c_void_p], C.c_int),
    'cmm_weighted_densities': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_launch_pass': ([C.c_void_p, C.c_int], C.c_int),
    'cmm_finish_c1': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_double], C.c_int),
    'cmm_finish_picard': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                           C.c_double, C.c_double], C.c_int),
    'cmm_finish_anderson': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                             C.c_void_p, C.c_double, C.c_double], C.c_int),
    'cmm_finish_check': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                          C.c_double], C.c_int),
    'cmm_boltzmann': ([C.c_void_p, C.c_longlong, C.c_double, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_anderson_init': ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_anderson_store_row': ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    'cmm_local_density': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_hs_weight_table': ([C.c_void_p, C.c_double, C.c_void_p], C.c_int),
    'cmm_anderson_coef': ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                           C.c_double, C.c_double, C.c_double], C.c_int),
    'cmm_anderson_mix': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                          C.c_void_p, C.c_void_p, C.c_int, C.c_double], C.c_int),
    'cmm_anderson_state': ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_anderson_set_damping': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_trial': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                   C.c_int], C.c_int),
    'cmm_get_scalars': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_batch_reset': ([C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int], C.c_int),
    'cmm_batch_check': ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    'cmm_batch_state': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_batch_release': ([C.c_void_p], C.c_int),
    'cmm_graph_begin': ([C.c_void_p], C.c_int),
    'cmm_graph_end': ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    'cmm_graph_launch': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_graph_destroy': ([C.c_void_p], C.c_int),
    'cmm_synchronize': ([C.c_void_p], C.c_int),
    'cmm_lj_grid': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                     C.c_double, C.c_int, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_hessian_apply': ([C.c_void_p, C.c_void_p], C.c_int),
    'cmm_adjoint_seed': ([C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p,
                          C.c_void_p, C.c_void_p], C.c_int),
    'cmm_adjoint_final': ([C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_adjoint_update': ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_adjoint_lambda': ([C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_lj_grad_workspace': ([C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong)], C.c_int),
    'cmm_lj_grid_grad': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                          C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                          C.c_void_p, C.c_void_p], C.c_int),
    'cmm_lj_grid_grad_noalloc': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_void_p], C.c_int),
    'cmm_effpot': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                    C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                    C.c_double, C.c_double, C.c_int, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_lj_derivs': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_int,
                       C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p],
                      C.c_int),
    'cmm_interp_table': ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_interp_points': ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_longlong,
                           C.c_void_p, C.c_void_p], C.c_int),
    'cmm_effpot_interp': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_int, C.c_void_p,
                           C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    'cmm_min_distance': ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
                         C.c_int),
    'cmm_lj_points': ([C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                       C.c_double, C.c_int, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
}

_lib = None


def load():
    """Load the library (once) and declare every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('libcmmdft_hip.so is not built (%s); run __graft_entry__.build()' % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in EXPORTS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(status):
    if status == CMM_OK:
        return
    msg = load().cmm_last_error().decode()
    if status == CMM_E_ARG:
        raise ValueError(msg)
    if status == CMM_E_NONFINITE:
        raise FloatingPointError(msg)
    if status == CMM_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _ints(vals):
    a = (C.c_int * max(1, len(vals)))(*vals)
    return a


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('the HIP path needs a GPU (torch.cuda.is_available() is False)')
    return torch


class Plan:
    """A grid plan for ``nbatch`` instances (owner of the device workspace)."""

    def __init__(self, npoints, rvecs, nbatch=1):
        torch = _torch()
        self.lib = load()
        self.npoints = tuple(int(n) for n in npoints)
        self.nbatch = int(nbatch)
        self.rvecs = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(3, 3))
        h = C.c_void_p()
        check(self.lib.cmm_plan_create(self.npoints[0], self.npoints[1], self.npoints[2],
                                       self.rvecs.ctypes.data_as(C.c_void_p), self.nbatch, C.byref(h)))
        self.h = h
        self.device = torch.device('cuda', torch.cuda.current_device())
        self._scal = np.zeros(self.nbatch * NSCAL)
        self.bind_stream()
        # tuned FFT radix sequences and tiles of this grid shape, if any
        from . import autotune
        self.tuned = autotune.apply(self)

    def bind_stream(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        self.stream = s
        check(self.lib.cmm_plan_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def __del__(self):
        try:
            if getattr(self, 'h', None):
                self.lib.cmm_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def info(self):
        out = np.zeros(12, dtype=np.int64)
        check(self.lib.cmm_plan_get_info(self.h, out.ctypes.data_as(C.c_void_p), 12))
        return dict(zip(['nx', 'ny', 'nz', 'nbatch', 'tb_x', 'tb_y', 'p_z', 'p_phi', 'workspace_bytes', 'nzh', 'nf',
                         'wda_kspace'], [int(v) for v in out]))

    @property
    def nfields(self):
        """Weighted-density fields of the functional set: 6, or 12 for the
        tensor variant."""
        return self.info['nf']

    def set_tiles(self, tb_x=0, tb_y=0, p_z=0, p_phi=0):
        check(self.lib.cmm_plan_set_tiles(self.h, tb_x, tb_y, p_z, p_phi))

    def set_radix(self, axis, radices):
        r = _ints([int(v) for v in radices])
        check(self.lib.cmm_plan_set_radix(self.h, int(axis), len(radices), r))

    def radix(self, axis):
        n = C.c_int()
        r = (C.c_int * 8)()
        check(self.lib.cmm_plan_get_radix(self.h, int(axis), C.byref(n), r))
        return [int(r[i]) for i in range(n.value)]

    def set_active(self, mask):
        self._active = mask
        check(self.lib.cmm_plan_set_active(self.h, _ptr(mask)))

    def set_functional(self, f: CmmFunctional):
        self._func = f
        self._mfa = None
        check(self.lib.cmm_set_functional(self.h, C.byref(f)))

    def set_mfa(self, kd):
        """Mean-field kernel table (complex half spectrum from ``mfa_table``)
        or None; kept referenced while set."""
        nx, ny, nz = self.npoints
        if kd is not None and kd.numel() * kd.element_size() != 16 * nx * ny * (nz // 2 + 1):
            raise ValueError('set_mfa: table must hold nx x ny x (nz/2+1) complex values')
        check(self.lib.cmm_plan_set_mfa(self.h, _ptr(kd)))
        self._mfa = kd

    @property
    def has_mfa(self):
        return getattr(self, '_mfa', None) is not None

    def mfa_table(self, uhat, reps, rvecs_sc, lanczos, scale, kd):
        rv = np.ascontiguousarray(np.asarray(rvecs_sc, dtype=np.float64).reshape(9))
        check(self.lib.cmm_mfa_table(self.h, _ptr(uhat), _ints([int(r) for r in reps]),
                                     rv.ctypes.data_as(C.c_void_p), int(bool(lanczos)), float(scale), _ptr(kd)))

    def set_instances(self, fug, alpha):
        fug = np.ascontiguousarray(np.broadcast_to(np.asarray(fug, dtype=np.float64), (self.nbatch,)))
        alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float64), (self.nbatch,)))
        check(self.lib.cmm_set_instances(self.h, fug.ctypes.data_as(C.c_void_p), alpha.ctypes.data_as(C.c_void_p)))

    # ---- transforms -------------------------------------------------------
    def fft_r2c(self, x, X):
        check(self.lib.cmm_fft_r2c(self.h, _ptr(x), _ptr(X)))

    def fft_c2r(self, X, x):
        check(self.lib.cmm_fft_c2r(self.h, _ptr(X), _ptr(x)))

    # ---- functional chain -------------------------------------------------
    def seed(self, rho, vext=None, vext_batched=False, lam3=1.0):
        check(self.lib.cmm_seed(self.h, _ptr(rho), _ptr(vext), int(vext_batched), float(lam3)))

    def derive(self):
        check(self.lib.cmm_derive(self.h))

    def value(self):
        check(self.lib.cmm_value(self.h))

    def launch_pass(self, which):
        check(self.lib.cmm_launch_pass(self.h, int(which)))

    def weighted_densities(self, out):
        """Weighted densities of the seeded density into out
        (nbatch x nfields x npts values)."""
        if out.numel() != self.nbatch * self.nfields * int(np.prod(self.npoints)):
            raise ValueError('weighted_densities: out must hold nbatch x %d x npts values' % self.nfields)
        check(self.lib.cmm_weighted_densities(self.h, _ptr(out)))

    def local_density(self, nfld=None, rbar=None, phi=None, fex=None, dfex=None):
        check(self.lib.cmm_local_density(self.h, _ptr(nfld), _ptr(rbar), _ptr(phi), _ptr(fex), _ptr(dfex)))

    def hs_weight_table(self, R, out):
        if out.numel() != 2 * int(np.prod(self.npoints)):
            raise ValueError('hs_weight_table: out must hold 2 x npts values')
        check(self.lib.cmm_hs_weight_table(self.h, float(R), _ptr(out)))

    def finish_c1(self, rho, vext, vext_batched, c1, lam3=1.0):
        check(self.lib.cmm_finish_c1(self.h, _ptr(rho), _ptr(vext), int(vext_batched), _ptr(c1), float(lam3)))

    def finish_picard(self, rho, vext, vext_batched, rho_out, c1_out, beta, lam3, zero_thr):
        check(self.lib.cmm_finish_picard(self.h, _ptr(rho), _ptr(vext), int(vext_batched), _ptr(rho_out),
                                         _ptr(c1_out), float(beta), float(lam3), float(zero_thr)))

    def finish_anderson(self, rho, vext, vext_batched, rhist, m, slot_new, c1_out, beta, lam3):
        check(self.lib.cmm_finish_anderson(self.h, _ptr(rho), _ptr(vext), int(vext_batched), _ptr(rhist), int(m),
                                           int(slot_new), _ptr(c1_out), float(beta), float(lam3)))

    def finish_check(self, rho, vext, vext_batched, beta, lam3, a_tol=1e-6, r_tol=1e-4):
        check(self.lib.cmm_finish_check(self.h, _ptr(rho), _ptr(vext), int(vext_batched), float(beta), float(lam3),
                                        float(a_tol), float(r_tol)))

    def anderson_store_row(self, m, slot_new):
        check(self.lib.cmm_anderson_store_row(self.h, int(m), int(slot_new)))

    def anderson_init(self, damping, cf):
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(damping, dtype=np.float64), (self.nbatch,)))
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(cf, dtype=np.float64), (self.nbatch,)))
        check(self.lib.cmm_anderson_init(self.h, d.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p)))

    def anderson_coef(self, m, order, slot_new, slot_prev, adaptive, fac_up, fac_down, dmin, dmax):
        check(self.lib.cmm_anderson_coef(self.h, int(m), len(order), _ints(order), int(slot_new), int(slot_prev),
                                         int(adaptive), float(fac_up), float(fac_down), float(dmin), float(dmax)))

    def anderson_mix(self, rhohist, order_rho, rhist, order, rho_cur, rho_out, vext, vext_batched, lam3):
        check(self.lib.cmm_anderson_mix(self.h, _ptr(rhohist), _ints(order_rho), _ptr(rhist), _ints(order),
                                        len(order), _ptr(rho_cur), _ptr(rho_out), _ptr(vext), int(vext_batched),
                                        float(lam3)))

    def anderson_set_damping(self, damping):
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(damping, dtype=np.float64), (self.nbatch,)))
        check(self.lib.cmm_anderson_set_damping(self.h, d.ctypes.data_as(C.c_void_p)))

    def anderson_state(self):
        d = np.zeros(self.nbatch)
        a = np.zeros(self.nbatch * MMAX)
        check(self.lib.cmm_anderson_state(self.h, d.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p)))
        return d, a.reshape(self.nbatch, MMAX)

    def trial(self, rho, rhist, slot, rho_out, vext, vext_batched, lam3, clip):
        check(self.lib.cmm_trial(self.h, _ptr(rho), _ptr(rhist), int(slot), _ptr(rho_out), _ptr(vext),
                                 int(vext_batched), float(lam3), int(clip)))

    # ---- linear response ---------------------------------------------------
    def hessian_apply(self, nfld):
        check(self.lib.cmm_hessian_apply(self.h, _ptr(nfld)))

    def adjoint_seed(self, rho, beta, dr, init, parity, y, r, p):
        check(self.lib.cmm_adjoint_seed(self.h, _ptr(rho), float(beta), float(dr), int(init), int(parity), _ptr(y),
                                        _ptr(r), _ptr(p)))

    def adjoint_final(self, rho, beta, p, q):
        check(self.lib.cmm_adjoint_final(self.h, _ptr(rho), float(beta), _ptr(p), _ptr(q)))

    def adjoint_update(self, parity, y, r, p, q):
        check(self.lib.cmm_adjoint_update(self.h, int(parity), _ptr(y), _ptr(r), _ptr(p), _ptr(q)))

    def adjoint_lambda(self, rho, beta, y, lam):
        check(self.lib.cmm_adjoint_lambda(self.h, _ptr(rho), float(beta), _ptr(y), _ptr(lam)))

    # ---- batched solves and graphs -----------------------------------------
    def batch_reset(self, mu, beta, lam3, dr, tol, nsteps):
        mu = np.ascontiguousarray(np.asarray(mu, dtype=np.float64))
        assert mu.shape == (self.nbatch,)
        check(self.lib.cmm_batch_reset(self.h, mu.ctypes.data_as(C.c_void_p), float(beta), float(lam3), float(dr),
                                       float(tol), int(nsteps)))

    def batch_check(self, new_slot, init=False):
        check(self.lib.cmm_batch_check(self.h, int(new_slot), int(init)))

    def batch_state(self):
        nb = self.nbatch
        a, it, st, fs = (np.zeros(nb, dtype=np.int32) for _ in range(4))
        om = np.zeros(nb)
        check(self.lib.cmm_batch_state(self.h, a.ctypes.data_as(C.c_void_p), it.ctypes.data_as(C.c_void_p),
                                       st.ctypes.data_as(C.c_void_p), fs.ctypes.data_as(C.c_void_p),
                                       om.ctypes.data_as(C.c_void_p)))
        return dict(active=a, iters=it, status=st, final_slot=fs, omega=om)

    def batch_release(self):
        check(self.lib.cmm_batch_release(self.h))

    def graph_begin(self):
        check(self.lib.cmm_graph_begin(self.h))

    def graph_end(self):
        ex = C.c_void_p()
        check(self.lib.cmm_graph_end(self.h, C.byref(ex)))
        return ex

    def graph_launch(self, ex):
        check(self.lib.cmm_graph_launch(self.h, ex))

    def graph_destroy(self, ex):
        check(self.lib.cmm_graph_destroy(ex))

    def scalars(self):
        check(self.lib.cmm_get_scalars(self.h, self._scal.ctypes.data_as(C.c_void_p)))
        return self._scal.reshape(self.nbatch, NSCAL).copy()

    def synchronize(self):
        check(self.lib.cmm_synchronize(self.h))


def boltzmann(c1, beta, fug, out, stream=None):
    """out = beta exp(-beta c1) fug on the device (cmm_boltzmann)."""
    torch = _torch()
    lib = load()
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_boltzmann(_ptr(c1), int(c1.numel()), float(beta), float(fug), _ptr(out),
                            C.c_void_p(s.cuda_stream)))


def lj_grid(npoints, rvecs, atoms_dev, natom, sigma_f, epsilon_f, cutoff, images, limit, out, stream=None):
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_lj_grid(int(npoints[0]), int(npoints[1]), int(npoints[2]), rv.ctypes.data_as(C.c_void_p),
                          _ptr(atoms_dev), int(natom), float(sigma_f), float(epsilon_f), float(cutoff), int(images),
                          float(limit), _ptr(out), C.c_void_p(s.cuda_stream)))


def lj_points(points_dev, M, rvecs, atoms_dev, natom, sigma_f, epsilon_f, cutoff, images, limit, out, stream=None):
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_lj_points(_ptr(points_dev), int(M), rv.ctypes.data_as(C.c_void_p), _ptr(atoms_dev), int(natom),
                            float(sigma_f), float(epsilon_f), float(cutoff), int(images), float(limit), _ptr(out),
                            C.c_void_p(s.cuda_stream)))


def effpot(npoints, points_dev, M, rvecs, atoms_dev, natom, gtype, gsig, geps, offsets_dev, weights_dev, nrot, dmax,
           beta, cutoff, images, nanval, out, stream=None):
    """Orientationally averaged potential (cmm_effpot): on the grid of
    ``npoints`` when ``points_dev`` is None, else at the M points."""
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    gt = np.ascontiguousarray(np.asarray(gtype, dtype=np.int32))
    gs = np.ascontiguousarray(np.asarray(gsig, dtype=np.float64))
    ge = np.ascontiguousarray(np.asarray(geps, dtype=np.float64))
    s = stream if stream is not None else torch.cuda.current_stream()
    n = [int(v) for v in npoints] if npoints is not None else [1, 1, 1]
    check(lib.cmm_effpot(n[0], n[1], n[2], _ptr(points_dev) if points_dev is not None else None, int(M),
                         rv.ctypes.data_as(C.c_void_p), _ptr(atoms_dev), int(natom), len(gt),
                         gt.ctypes.data_as(C.c_void_p), len(gs), gs.ctypes.data_as(C.c_void_p),
                         ge.ctypes.data_as(C.c_void_p), _ptr(offsets_dev), _ptr(weights_dev), int(nrot), float(dmax),
                         float(beta), float(cutoff), int(images), float(nanval), _ptr(out),
                         C.c_void_p(s.cuda_stream)))


def lj_derivs(npoints, points_dev, M, rvecs, atoms_dev, natom, sigma_f, epsilon_f, cutoff, images, spacings, vmax,
              out, stream=None):
    """cmm_lj_derivs: out (8, npts) on the grid of ``npoints`` (points_dev None) or at M points."""
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    h = np.ascontiguousarray(np.asarray(spacings, dtype=np.float64).reshape(3))
    s = stream if stream is not None else torch.cuda.current_stream()
    n = [int(v) for v in npoints] if npoints is not None else [1, 1, 1]
    check(lib.cmm_lj_derivs(n[0], n[1], n[2], _ptr(points_dev) if points_dev is not None else None, int(M),
                            rv.ctypes.data_as(C.c_void_p), _ptr(atoms_dev), int(natom), float(sigma_f),
                            float(epsilon_f), float(cutoff), int(images), h.ctypes.data_as(C.c_void_p), float(vmax),
                            _ptr(out), C.c_void_p(s.cuda_stream)))


def interp_table(values_dev, n, estimate, table, stream=None):
    torch = _torch()
    lib = load()
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_interp_table(_ptr(values_dev), int(n[0]), int(n[1]), int(n[2]), int(estimate), _ptr(table),
                               C.c_void_p(s.cuda_stream)))


def interp_points(table, n, origin, spacing, mode, points_dev, M, out, stream=None):
    torch = _torch()
    lib = load()
    s = stream if stream is not None else torch.cuda.current_stream()
    o = np.ascontiguousarray(np.asarray(origin, dtype=np.float64).reshape(3))
    h = np.ascontiguousarray(np.asarray(spacing, dtype=np.float64).reshape(3))
    check(lib.cmm_interp_points(_ptr(table), _ints([int(v) for v in n]), o.ctypes.data_as(C.c_void_p),
                                h.ctypes.data_as(C.c_void_p), int(mode), _ptr(points_dev), int(M), _ptr(out),
                                C.c_void_p(s.cuda_stream)))


def effpot_interp(npoints, points_dev, M, rvecs, gtype, tables, tn, torigin, tspacing, tmode, offsets_dev,
                  weights_dev, nrot, beta, nanval, out, stream=None):
    """cmm_effpot_interp: ``tables`` a list of device tables (one per guest
    type) with their sizes ``tn``, origins, spacings and modes."""
    torch = _torch()
    lib = load()
    s = stream if stream is not None else torch.cuda.current_stream()
    nt = len(tables)
    ptrs = (C.c_void_p * nt)(*[t.data_ptr() for t in tables])
    gt = np.ascontiguousarray(np.asarray(gtype, dtype=np.int32))
    tn_ = np.ascontiguousarray(np.asarray(tn, dtype=np.int32).reshape(nt, 3))
    to = np.ascontiguousarray(np.asarray(torigin, dtype=np.float64).reshape(nt, 3))
    th = np.ascontiguousarray(np.asarray(tspacing, dtype=np.float64).reshape(nt, 3))
    tm = np.ascontiguousarray(np.asarray(tmode, dtype=np.int32).reshape(nt))
    rv = np.ascontiguousarray(np.asarray(rvecs if rvecs is not None else np.eye(3), dtype=np.float64).reshape(9))
    n = [int(v) for v in npoints] if npoints is not None else [1, 1, 1]
    check(lib.cmm_effpot_interp(n[0], n[1], n[2], _ptr(points_dev) if points_dev is not None else None, int(M),
                                rv.ctypes.data_as(C.c_void_p), len(gt), gt.ctypes.data_as(C.c_void_p), nt,
                                C.cast(ptrs, C.c_void_p), tn_.ctypes.data_as(C.c_void_p),
                                to.ctypes.data_as(C.c_void_p), th.ctypes.data_as(C.c_void_p),
                                tm.ctypes.data_as(C.c_void_p), _ptr(offsets_dev), _ptr(weights_dev), int(nrot),
                                float(beta), float(nanval), _ptr(out), C.c_void_p(s.cuda_stream)))


def min_distance(npoints, rvecs, pos_dev, natom, out, stream=None):
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_min_distance(int(npoints[0]), int(npoints[1]), int(npoints[2]), rv.ctypes.data_as(C.c_void_p),
                               _ptr(pos_dev), int(natom), _ptr(out), C.c_void_p(s.cuda_stream)))


def lj_grid_grad(npoints, rvecs, atoms_dev, types_dev, natom, ntypes, sigma_f, epsilon_f, cutoff, images, limit, w,
                 stream=None):
    """Device tensor (ntypes, 2): sum_i w_i dV_i/dsigma_t, sum_i w_i dV_i/depsilon_t;
    ``w`` is one weight field (npts) or several (nw, npts), summed per point."""
    nw = 1 if w.dim() == 1 else int(w.shape[0])
    torch = _torch()
    lib = load()
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    s = stream if stream is not None else torch.cuda.current_stream()
    nwork = C.c_longlong()
    check(lib.cmm_lj_grad_workspace(int(npoints[0]), int(npoints[1]), int(npoints[2]), int(ntypes), C.byref(nwork)))
    work = torch.empty(int(nwork.value), dtype=torch.float64, device=w.device)
    grad = torch.empty((int(ntypes), 2), dtype=torch.float64, device=w.device)
    check(lib.cmm_lj_grid_grad(int(npoints[0]), int(npoints[1]), int(npoints[2]), rv.ctypes.data_as(C.c_void_p),
                               _ptr(atoms_dev), _ptr(types_dev), int(natom), int(ntypes), float(sigma_f),
                               float(epsilon_f), float(cutoff), int(images), float(limit), _ptr(w), nw, _ptr(work),
                               _ptr(grad), C.c_void_p(s.cuda_stream)))
    return grad


def mfa_lj_potential(ns, rvecs, reps, sigma, epsilon, rmin, limit, cutoff, stream=None):
    """Device tensor (ns / reps): the mean-field LJ potential on the grid of
    ``ns`` points of the cell ``rvecs``, folded over ``reps`` images
    (cutoff None: no cutoff)."""
    torch = _torch()
    lib = load()
    n = [int(a) // int(r) for a, r in zip(ns, reps)]
    out = torch.empty(n, dtype=torch.float64, device='cuda')
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(9))
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_mfa_lj_potential(_ints([int(a) for a in ns]), rv.ctypes.data_as(C.c_void_p),
                                   _ints([int(r) for r in reps]), float(sigma), float(epsilon), float(rmin),
                                   float(limit), float(cutoff if cutoff is not None else 1.0),
                                   int(cutoff is not None), _ptr(out), C.c_void_p(s.cuda_stream)))
    return out


def fold(field, reps, stream=None):
    """Device tensor: ``field`` (3-D, device) summed over its reps images."""
    torch = _torch()
    lib = load()
    ns = [int(a) for a in field.shape]
    n = [a // int(r) for a, r in zip(ns, reps)]
    out = torch.empty(n, dtype=torch.float64, device=field.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_fold(_ints(ns), _ints([int(r) for r in reps]), _ptr(field.contiguous()), _ptr(out),
                       C.c_void_p(s.cuda_stream)))
    return out


FIELD_OPS = {'sum': 0, 'boltzmann': 1, 'boltzmann_clip': 2, 'min': 3, 'max': 4, 'sumsq': 5, 'fid': 6, 'fid_track': 7}


def field_reduce(x, op, beta=1.0, lam3=1.0, mask=None, stream=None):
    """Deterministic device reduction of a field (see cmm_field_reduce);
    returns a Python float."""
    torch = _torch()
    lib = load()
    x = x.reshape(-1).contiguous()
    if mask is not None:
        mask = mask.reshape(-1).contiguous().to(torch.uint8)
        if mask.numel() != x.numel():
            raise ValueError('mask must have the shape of the field')
    work = torch.empty(1024, dtype=torch.float64, device=x.device)
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.cmm_field_reduce(_ptr(x), _ptr(mask), int(x.numel()), FIELD_OPS[op], float(beta), float(lam3),
                               _ptr(work), _ptr(out), C.c_void_p(s.cuda_stream)))
    return float(out.item())
