// C ABI of libcmmdft_hip.so: plans, launchers and error handling.
// Declared in include/cmmdft_hip.h.  One translation unit: the kernels are
// included here so the launch templates are instantiated next to their use.
#include "cmm_kernels.hip"
#include "cmm_lj.hip"
#include "cmm_effpot.hip"
#include "cmm_adjoint.hip"
#include "cmm_mfa.hip"
#include "cmm_calc.hip"
#include "../../include/cmmdft_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace cmm;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return fail(CMM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(what)                                                                           \
  do {                                                                                            \
    hipError_t e_ = hipGetLastError();                                                            \
    if (e_ != hipSuccess) return fail(CMM_E_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct cmm_plan {
  int nx, ny, nz, nzh, nb;
  double rvecs[9];
  hipStream_t stream;
  double2* tw[3];
  double* ktab[3];
  double* ltab[3];
  GridDev g;
  FuncDev fp;
  int has_func;
  double2* rz;
  double2* F;
  double* kwt;       // the functional's weights w2, w3, w_D on this grid, rows x <= nx/2 [3][..] (x passes)
  long long nhalf;   // complex elements per half spectrum
  long long npts;
  double* partials;
  long long pstride;
  double* scal;
  InstDev* inst;
  double* gram;
  double* damping;
  double* cf;
  double* coef;
  double* alpha_out;
  const int* active;
  int tbx, tby, pz, pphi;
  int nf;           // weighted-density fields of the functional: 6, or 12 with the tensor variant
  int fcap;         // fields the workspace F holds per instance
  size_t ws_bytes;
  // batched-solve bookkeeping (allocated on first use)
  int* b_active;
  int* b_iters;
  int* b_status;
  int* b_final;
  double* b_omega_prev;
  double* b_omega;
  double* b_mu;
  BatchDev bd;
};

namespace {

// stage divisors of a radix sequence (already in f.rad[0..nst))
void radix_finish(Fft1D& f) {
  int Ns = 1;
  for (int s = 0; s < f.nst; ++s) {
    f.fns[s] = make_fastdiv((unsigned)Ns);
    f.fnq[s] = make_fastdiv((unsigned)(f.n / f.rad[s]));
    Ns *= f.rad[s];
  }
}

// default radix sequence: 5s, 3s, then 8s, 4s, 2s
bool radix_plan(int n, Fft1D& f) {
  f.n = n;
  f.nst = 0;
  int m = n;
  while (m % 5 == 0 && f.nst < CMM_MAX_STAGES) { f.rad[f.nst++] = 5; m /= 5; }
  while (m % 3 == 0 && f.nst < CMM_MAX_STAGES) { f.rad[f.nst++] = 3; m /= 3; }
  while (m % 8 == 0 && f.nst < CMM_MAX_STAGES) { f.rad[f.nst++] = 8; m /= 8; }
  while (m % 4 == 0 && f.nst < CMM_MAX_STAGES) { f.rad[f.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && f.nst < CMM_MAX_STAGES) { f.rad[f.nst++] = 2; m /= 2; }
  if (m != 1) return false;
  radix_finish(f);
  return true;
}

double sinc_np(double x) {   // numpy.sinc: sin(pi x)/(pi x), 1 at 0
  if (x == 0.0) return 1.0;
  const double y = M_PI * x;
  return std::sin(y) / y;
}

int largest_even_le(int v) { return std::max(2, v - (v & 1)); }

void default_tiles(cmm_plan* p) {
  const int budget = VPT * NT;               // complex values a block stages
  // x passes: one tile of nx * tb_x columns per NTX threads (all fields in LDS;
  // launches with twelve fields cap tb_x by LDS)
  p->tbx = std::max(1, std::min(32, NTX / p->nx));
  // y passes: tb_y columns, and the paired block's two rounds of
  // ceil(tb_y / 2) columns within the same staging limit
  int tby = std::min(32, std::min(budget / p->ny, 4096 / p->ny));
  if ((tby & 1) && (tby + 1) * p->ny > budget) --tby;
  p->tby = std::max(1, tby);
  // simple z pass: (P/2)*nz <= budget and nz*((P/2)|1)*16 <= 64 KB
  int pz = std::min(32, 2 * budget / p->nz);
  while (pz > 2 && (size_t)p->nz * ((pz / 2) | 1) * 16 > 65536) pz -= 2;
  p->pz = largest_even_le(pz);
  // 6-field z pass: 3*P*nz <= budget, nz*((3P)|1)*16 <= 64 KB; nz = 128:
  // the compile-time tile of 6 pencils (measured at 128^3, 32 instances,
  // five fields: 1.29 ms against 1.59 ms for 4 pencils and 1.49 ms for 8)
  int pphi = std::min(16, budget / (3 * p->nz));
  while (pphi > 2 && (size_t)p->nz * ((3 * pphi) | 1) * 16 > 65536) pphi -= 2;
  p->pphi = (p->nz == 128) ? 6 : largest_even_le(pphi);
}

RedDev red_of(cmm_plan* p) { return RedDev{p->partials, p->pstride, p->scal}; }

// totals of the partials the previous kernel stored (see k_reduce_partials)
int launch_reduce(cmm_plan* p, long long nblocks, int nq, unsigned maxmask, int slot, unsigned skipmask = 0u) {
  if ((long long)nq * nblocks > p->pstride) return fail(CMM_E_ARG, "partials beyond the per-instance stride");
  hipLaunchKernelGGL(k_reduce_partials, dim3(p->nb, nq), dim3(NT), 0, p->stream, red_of(p), (int)nblocks, maxmask,
                     slot, p->active, skipmask);
  hipError_t e_ = hipGetLastError();
  if (e_ != hipSuccess) return fail(CMM_E_HIP, std::string("k_reduce_partials: ") + hipGetErrorString(e_));
  return CMM_OK;
}

long long zblocks(cmm_plan* p, int P) { return ((long long)p->nx * p->ny + P - 1) / P; }

ZTile ztile(cmm_plan* p, int P, int nf) {
  ZTile t;
  t.P = P;
  t.B = P / 2;
  t.S = nf * t.B;
  t.ES = t.S | 1;
  t.fn = make_fastdiv((unsigned)p->nz);
  t.fnzh = make_fastdiv((unsigned)p->nzh);
  t.fS = make_fastdiv((unsigned)t.S);
  return t;
}

size_t zlds(cmm_plan* p, int P, int nf) { return (size_t)p->nz * ((nf * (P / 2)) | 1) * sizeof(double2); }

}  // namespace

// The exported functions take C linkage from their declarations in cmmdft_hip.h.

int cmm_version(void) { return 10000; }

const char* cmm_last_error(void) { return g_err.c_str(); }

int cmm_plan_create(int nx, int ny, int nz, const double* rvecs, int nbatch, cmm_plan** out) {
  if (!out || !rvecs) return fail(CMM_E_ARG, "null argument");
  if (nx < 2 || ny < 2 || nz < 2 || nbatch < 1 || nbatch > 4096) return fail(CMM_E_ARG, "bad grid or batch size");
  if (nx > CMM_MAX_N || ny > CMM_MAX_N || nz > CMM_MAX_N)
    return fail(CMM_E_ARG, "grid axis longer than 256 is not supported");
  cmm_plan* p = new cmm_plan();
  std::memset(p, 0, sizeof(*p));
  p->nx = nx; p->ny = ny; p->nz = nz; p->nzh = nz / 2 + 1; p->nb = nbatch;
  std::memcpy(p->rvecs, rvecs, 9 * sizeof(double));
  p->stream = 0;
  int ns[3] = {nx, ny, nz};
  Fft1D* fs[3] = {&p->g.fx, &p->g.fy, &p->g.fz};
  KAxis* ks[3] = {&p->g.kx, &p->g.ky, &p->g.kz};
  for (int a = 0; a < 3; ++a) {
    const int n = ns[a];
    if (!radix_plan(n, *fs[a])) { delete p; return fail(CMM_E_ARG, "axis length must be 2^a 3^b 5^c"); }
    std::vector<double2> tw(n);
    for (int j = 0; j < n; ++j) {
      long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
      tw[j] = make_double2((double)cosl(ang), (double)(-sinl(ang)));
    }
    // spacing from the column norm (cell vectors are the rows)
    const double colnorm = std::sqrt(rvecs[0 * 3 + a] * rvecs[0 * 3 + a] + rvecs[1 * 3 + a] * rvecs[1 * 3 + a] +
                                     rvecs[2 * 3 + a] * rvecs[2 * 3 + a]);
    const double h = colnorm / n;
    const double val = 1.0 / (n * h);
    std::vector<double> kt(n), lt(n);
    const double kcut = 2.0 * M_PI / h;
    for (int m = 0; m < n; ++m) {
      const int fm = (m <= (n - 1) / 2) ? m : m - n;
      kt[m] = 2.0 * M_PI * (fm * val);
      lt[m] = sinc_np(kt[m] / kcut);
    }
    if (hipMalloc(&p->tw[a], n * sizeof(double2)) != hipSuccess || hipMalloc(&p->ktab[a], n * sizeof(double)) != hipSuccess ||
        hipMalloc(&p->ltab[a], n * sizeof(double)) != hipSuccess) {
      cmm_plan_destroy(p);
      return fail(CMM_E_HIP, "hipMalloc tables");
    }
    (void)hipMemcpy(p->tw[a], tw.data(), n * sizeof(double2), hipMemcpyHostToDevice);
    (void)hipMemcpy(p->ktab[a], kt.data(), n * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(p->ltab[a], lt.data(), n * sizeof(double), hipMemcpyHostToDevice);
    fs[a]->tw = p->tw[a];
    ks[a]->k = p->ktab[a];
    ks[a]->lanczos = p->ltab[a];
    ks[a]->nyq = (n % 2 == 0) ? n / 2 : -1;
  }
  p->g.nx = nx; p->g.ny = ny; p->g.nz = nz; p->g.nzh = p->nzh;
  p->npts = (long long)nx * ny * nz;
  p->nhalf = (long long)nx * ny * p->nzh;
  p->g.inv_npts = 1.0 / (double)p->npts;
  default_tiles(p);
  // workspace: rz (1) + F (6) half spectra per instance
  size_t half_bytes = (size_t)p->nhalf * sizeof(double2) * nbatch;
  // (x-pass reductions have one partial per tile: at most ny * nzh)
  long long maxblocks = std::max(zblocks(p, 2), (long long)ny * p->nzh);
  p->pstride = 16 * maxblocks;
  hipError_t e = hipSuccess;
  e = (e == hipSuccess) ? hipMalloc(&p->rz, half_bytes) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->F, 6 * half_bytes) : e;
  p->nf = 6;
  p->fcap = 6;
  e = (e == hipSuccess) ? hipMalloc(&p->partials, (size_t)p->pstride * nbatch * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->scal, (size_t)NSCAL * nbatch * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->inst, (size_t)nbatch * sizeof(InstDev)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->gram, (size_t)nbatch * CMM_MMAX * CMM_MMAX * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->damping, (size_t)nbatch * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->cf, (size_t)nbatch * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->coef, (size_t)nbatch * 2 * CMM_MMAX * sizeof(double)) : e;
  e = (e == hipSuccess) ? hipMalloc(&p->alpha_out, (size_t)nbatch * CMM_MMAX * sizeof(double)) : e;
  if (e != hipSuccess) {
    cmm_plan_destroy(p);
    return fail(CMM_E_HIP, std::string("workspace allocation: ") + hipGetErrorString(e));
  }
  (void)hipMemset(p->scal, 0, (size_t)NSCAL * nbatch * sizeof(double));
  (void)hipMemset(p->gram, 0, (size_t)nbatch * CMM_MMAX * CMM_MMAX * sizeof(double));
  (void)hipMemset(p->coef, 0, (size_t)nbatch * 2 * CMM_MMAX * sizeof(double));
  std::vector<InstDev> iv(nbatch, InstDev{0.0, 0.0});
  (void)hipMemcpy(p->inst, iv.data(), nbatch * sizeof(InstDev), hipMemcpyHostToDevice);
  p->ws_bytes = 7 * half_bytes;
  if (hipDeviceSynchronize() != hipSuccess) {
    cmm_plan_destroy(p);
    return fail(CMM_E_HIP, "plan initialisation");
  }
  *out = p;
  return CMM_OK;
}

int cmm_plan_destroy(cmm_plan* p) {
  if (!p) return CMM_OK;
  for (int a = 0; a < 3; ++a) {
    if (p->tw[a]) (void)hipFree(p->tw[a]);
    if (p->ktab[a]) (void)hipFree(p->ktab[a]);
    if (p->ltab[a]) (void)hipFree(p->ltab[a]);
  }
  void* bufs[] = {p->rz, p->F, p->kwt, p->partials, p->scal, p->inst, p->gram, p->damping, p->cf, p->coef,
                  p->alpha_out, p->b_active, p->b_iters, p->b_status, p->b_final, p->b_omega_prev, p->b_omega,
                  p->b_mu};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete p;
  return CMM_OK;
}

int cmm_plan_set_stream(cmm_plan* p, void* stream) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  p->stream = (hipStream_t)stream;
  return CMM_OK;
}

int cmm_plan_set_tiles(cmm_plan* p, int tbx, int tby, int pz, int pphi) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  const int budget = VPT * NT;
  if (tbx > 0) {
    if (tbx * p->nx > NTX) return fail(CMM_E_ARG, "tb_x too large (nx * tb_x must not exceed 512)");
    p->tbx = tbx;
  }
  if (tby > 0) {
    const int tbe = tby + (tby & 1);   // the paired y block's two rounds of ceil(tb_y / 2) columns
    if (tbe * p->ny > budget || tbe * p->ny * 16 > 65536) return fail(CMM_E_ARG, "tb_y too large");
    p->tby = tby;
  }
  if (pz > 0) {
    if ((pz & 1) || (pz / 2) * p->nz > budget || zlds(p, pz, 1) > 65536) return fail(CMM_E_ARG, "p_z invalid");
    p->pz = pz;
  }
  if (pphi > 0) {
    // nz = 128: the compile-time z tiles (2, 4, 6 or 8 pencils); other
    // lengths: six fields of the tile within the staging and LDS limits
    const bool ok = (p->nz == 128) ? (pphi <= 8)
                                   : (3 * pphi * p->nz <= budget && zlds(p, pphi, 6) <= 65536);
    if ((pphi & 1) || !ok) return fail(CMM_E_ARG, "p_phi invalid");
    p->pphi = pphi;
  }
  return CMM_OK;
}

int cmm_plan_set_radix(cmm_plan* p, int axis, int nstages, const int* radices) {
  if (!p || !radices) return fail(CMM_E_ARG, "null argument");
  if (axis < 0 || axis > 2) return fail(CMM_E_ARG, "axis must be 0, 1 or 2");
  if (nstages < 1 || nstages > CMM_MAX_STAGES) return fail(CMM_E_ARG, "bad number of stages");
  Fft1D* f = axis == 0 ? &p->g.fx : (axis == 1 ? &p->g.fy : &p->g.fz);
  long long prod = 1;
  for (int s = 0; s < nstages; ++s) {
    const int r = radices[s];
    if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return fail(CMM_E_ARG, "radix must be 2, 3, 4, 5 or 8");
    prod *= r;
  }
  if (prod != f->n) return fail(CMM_E_ARG, "radices do not multiply to the axis length");
  f->nst = nstages;
  for (int s = 0; s < nstages; ++s) f->rad[s] = radices[s];
  radix_finish(*f);
  return CMM_OK;
}

int cmm_plan_get_radix(cmm_plan* p, int axis, int* nstages, int* radices) {
  if (!p || !nstages || !radices) return fail(CMM_E_ARG, "null argument");
  if (axis < 0 || axis > 2) return fail(CMM_E_ARG, "axis must be 0, 1 or 2");
  const Fft1D* f = axis == 0 ? &p->g.fx : (axis == 1 ? &p->g.fy : &p->g.fz);
  *nstages = f->nst;
  for (int s = 0; s < f->nst; ++s) radices[s] = f->rad[s];
  return CMM_OK;
}

int cmm_plan_get_info(cmm_plan* p, long long* info, int n) {
  if (!p || !info) return fail(CMM_E_ARG, "null argument");
  long long v[12] = {p->nx, p->ny, p->nz, p->nb, p->tbx, p->tby, p->pz, p->pphi, (long long)p->ws_bytes, p->nzh, p->nf,
                     (p->nf == 6 && p->fp.wl2 != 0.0) ? 1 : 0};
  for (int i = 0; i < n && i < 12; ++i) info[i] = v[i];
  return CMM_OK;
}

int cmm_plan_set_active(cmm_plan* p, const int* mask) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  p->active = mask;
  return CMM_OK;
}

int cmm_set_functional(cmm_plan* p, const cmm_functional* f) {
  if (!p || !f) return fail(CMM_E_ARG, "null argument");
  if (f->hs_base < -1 || f->hs_base > 2) return fail(CMM_E_ARG, "hs_base must be -1, 0, 1 or 2");
  if (f->n_eos < 0 || f->n_eos > 4) return fail(CMM_E_ARG, "n_eos must be 0..4");
  if (f->hs_base >= 0 && !(f->R > 0.0)) return fail(CMM_E_ARG, "hard-sphere radius must be positive");
  if (!(f->kT > 0.0)) return fail(CMM_E_ARG, "kT must be positive");
  FuncDev& d = p->fp;
  std::memset(&d, 0, sizeof(d));
  d.use_hs = f->hs_base >= 0;
  d.base = std::max(0, f->hs_base);
  d.anti = f->hs_antisym ? 1 : 0;
  d.tensor = (d.use_hs && f->hs_tensor) ? 1 : 0;
  const int nf = d.tensor ? 12 : 6;
  if (nf == 12 && (12 * p->nz > VPT * NT || 13 * 16 * p->nx > 65536 - CMM_MAX_N * 16))
    return fail(CMM_E_UNSUPPORTED, "the tensor variant needs nz <= 170 (twelve fields per z tile)");
  if (nf > p->fcap) {
    // grow the workspace to twelve half spectra per instance
    (void)hipDeviceSynchronize();
    double2* Fn = nullptr;
    HIPCHK(hipMalloc(&Fn, (size_t)nf * p->nhalf * p->nb * sizeof(double2)));
    if (p->F) (void)hipFree(p->F);
    p->F = Fn;
    p->fcap = nf;
    p->ws_bytes = (size_t)(nf + 1) * p->nhalf * p->nb * sizeof(double2);
  }
  p->nf = nf;
  d.use_wda = f->use_wda && f->n_eos > 0;
  d.neos = f->n_eos;
  for (int t = 0; t < 4; ++t) {
    d.eos_kind[t] = f->eos_kind[t];
    d.eos_fac[t] = f->eos_factor[t];
    for (int k = 0; k < 16; ++k) d.eos_par[t][k] = f->eos_par[t][k];
    if (t < f->n_eos && (f->eos_kind[t] < 0 || f->eos_kind[t] > 2)) return fail(CMM_E_ARG, "unknown EOS kind");
  }
  // Every EOS term mean field: f_ex = A rho_bar^2 is quadratic, its
  // derivative w_D * (2A rho_bar) linear in rho, and the x passes apply it in
  // k space (2A w_D^2 rho^ / N, energy by Parseval) unless the caller asks
  // for the pointwise route.  Six-field functionals only.
  if (d.use_wda && nf == 6 && !f->wda_pointwise) {
    bool lin = true;
    double A = 0.0;
    for (int t = 0; t < d.neos; ++t) {
      lin = lin && d.eos_kind[t] == 0;
      A += d.eos_fac[t] * d.eos_par[t][0];
    }
    if (lin) d.wl2 = 2.0 * A;
  }
  d.R = d.use_hs ? f->R : 1.0;
  if (!(f->n3_series >= 0.0 && f->n3_series <= 1e-3))
    return fail(CMM_E_ARG, "n3_series must lie in [0, 1e-3] (the series forms are second order)");
  d.n3_series = f->n3_series;
  d.D = f->D;
  d.kT = f->kT;
  d.c_w2 = 4.0 * M_PI * d.R * d.R;
  d.c_w3 = 4.0 * M_PI * d.R * d.R * d.R / 3.0;
  d.inv4piR2 = 1.0 / (4.0 * M_PI * d.R * d.R);
  d.inv4piR = 1.0 / (4.0 * M_PI * d.R);
  d.c_tor = 3.0 / d.R;
  // the x passes' weights of the half-spectrum elements with x <= nx/2 (those
  // of x and nx - x are equal), computed once per functional (they depend on
  // R, D and the grid) on the plan's stream
  const long long kwt_n = (long long)(p->nx / 2 + 1) * p->ny * p->nzh;   // rows x <= nx/2
  if (!p->kwt) HIPCHK(hipMalloc(&p->kwt, (size_t)3 * kwt_n * sizeof(double)));
  hipLaunchKernelGGL(k_kw_table, dim3((unsigned)std::min<long long>((p->nhalf + NT - 1) / NT, 65536)), dim3(NT), 0,
                     p->stream, p->g, d, p->kwt);
  LAUNCHCHK("k_kw_table");
  d.kwt = p->kwt;
  d.kwt_n = kwt_n;
  p->has_func = 1;
  return CMM_OK;
}

int cmm_plan_set_mfa(cmm_plan* p, const double* kd) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (kd && !p->has_func) return fail(CMM_E_ARG, "set the functional before the mean-field table");
  p->fp.kmfa = reinterpret_cast<const double2*>(kd);
  return CMM_OK;
}

int cmm_mfa_lj_potential(const int* ns, const double* rvecs, const int* reps, double sigma, double epsilon,
                         double rmin, double limit, double cutoff, int use_cutoff, double* out, void* stream) {
  if (!ns || !rvecs || !reps || !out) return fail(CMM_E_ARG, "null argument");
  int n[3];
  for (int a = 0; a < 3; ++a) {
    if (ns[a] < 1 || reps[a] < 1 || ns[a] % reps[a]) return fail(CMM_E_ARG, "grid points must be a multiple of reps");
    n[a] = ns[a] / reps[a];
  }
  if (!(sigma > 0.0)) return fail(CMM_E_ARG, "sigma must be positive");
  Cell9 c;
  for (int k = 0; k < 9; ++k) c.rv[k] = rvecs[k];
  MfaLJ lj;
  lj.sigma = sigma;
  lj.eps4 = 4.0 * epsilon;
  lj.rmin = rmin;
  lj.limit = limit;
  lj.cutoff = cutoff;
  lj.use_cutoff = use_cutoff ? 1 : 0;
  const double rc6 = std::pow(sigma / cutoff, 6.0);
  lj.shift = use_cutoff ? 4.0 * epsilon * (rc6 * rc6 - rc6) : 0.0;
  const long long npts = (long long)n[0] * n[1] * n[2];
  hipLaunchKernelGGL(k_mfa_lj, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream, out, n[0],
                     n[1], n[2], reps[0], reps[1], reps[2], c, lj);
  LAUNCHCHK("k_mfa_lj");
  return CMM_OK;
}

int cmm_fold(const int* ns, const int* reps, const double* in, double* out, void* stream) {
  if (!ns || !reps || !in || !out) return fail(CMM_E_ARG, "null argument");
  int n[3];
  for (int a = 0; a < 3; ++a) {
    if (ns[a] < 1 || reps[a] < 1 || ns[a] % reps[a]) return fail(CMM_E_ARG, "grid points must be a multiple of reps");
    n[a] = ns[a] / reps[a];
  }
  const long long npts = (long long)n[0] * n[1] * n[2];
  hipLaunchKernelGGL(k_fold, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream, in, out, n[0],
                     n[1], n[2], reps[0], reps[1], reps[2]);
  LAUNCHCHK("k_fold");
  return CMM_OK;
}

int cmm_mfa_table(cmm_plan* p, const double* uhat, const int* reps, const double* rvecs_sc, int lanczos,
                  double scale, double* kd) {
  if (!p || !uhat || !reps || !rvecs_sc || !kd) return fail(CMM_E_ARG, "null argument");
  const int n[3] = {p->nx, p->ny, p->nz};
  const int nh[3] = {p->nx, p->ny, p->nzh};
  std::vector<double2> ph(nh[0] + nh[1] + nh[2]);
  int off = 0;
  for (int a = 0; a < 3; ++a) {
    if (reps[a] < 1) return fail(CMM_E_ARG, "reps must be positive");
    // supercell axis a: N_sc = reps n, spacing from the column norm of the
    // supercell vectors (system.py:262-266); the grid's wave vector m is the
    // supercell's reps m, fftfreq index reps fm (system.py:284-291)
    const long long Nsc = (long long)reps[a] * n[a];
    const double col = std::sqrt(rvecs_sc[0 * 3 + a] * rvecs_sc[0 * 3 + a] + rvecs_sc[1 * 3 + a] * rvecs_sc[1 * 3 + a] +
                                 rvecs_sc[2 * 3 + a] * rvecs_sc[2 * 3 + a]);
    const double h = col / (double)Nsc;
    const double val = 1.0 / ((double)Nsc * h);
    for (int m = 0; m < nh[a]; ++m) {
      const long long fm = (m <= (n[a] - 1) / 2) ? m : m - n[a];
      const double k = 2.0 * M_PI * ((double)(reps[a] * fm) * val);
      const double theta = M_PI * (k * h * (double)Nsc);
      ph[off + m] = make_double2(std::cos(theta), std::sin(theta));
    }
    off += nh[a];
  }
  double2* dph = nullptr;
  HIPCHK(hipMalloc(&dph, ph.size() * sizeof(double2)));
  hipError_t e = hipMemcpyAsync(dph, ph.data(), ph.size() * sizeof(double2), hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfa_table, dim3((unsigned)((p->nhalf + 255) / 256)), dim3(256), 0, p->stream,
                       reinterpret_cast<const double2*>(uhat), reinterpret_cast<double2*>(kd), p->g, dph,
                       lanczos ? 1 : 0, scale);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  (void)hipFree(dph);
  if (e != hipSuccess) return fail(CMM_E_HIP, std::string("k_mfa_table: ") + hipGetErrorString(e));
  return CMM_OK;
}

int cmm_field_reduce(const double* x, const unsigned char* mask, long long n, int op, double beta, double lam3,
                     double* work, double* out, void* stream) {
  if (!x || !work || !out) return fail(CMM_E_ARG, "null argument");
  if (n < 1) return fail(CMM_E_ARG, "empty field");
  if (op < FOP_SUM || op > FOP_FID_TRACK) return fail(CMM_E_ARG, "unknown reduction");
  const long long want = (n + RED_NT - 1) / RED_NT;
  const int nb = (int)std::min<long long>(RED_BLOCKS, want);
  hipLaunchKernelGGL(k_field_reduce, dim3(nb), dim3(RED_NT), 0, (hipStream_t)stream, x, mask, n, op, beta, lam3, work);
  LAUNCHCHK("k_field_reduce");
  hipLaunchKernelGGL(k_field_reduce_final, dim3(1), dim3(64), 0, (hipStream_t)stream, work, nb, op, out);
  LAUNCHCHK("k_field_reduce_final");
  return CMM_OK;
}

int cmm_set_instances(cmm_plan* p, const double* fug, const double* alpha) {
  if (!p || !fug || !alpha) return fail(CMM_E_ARG, "null argument");
  std::vector<InstDev> iv(p->nb);
  for (int i = 0; i < p->nb; ++i) iv[i] = InstDev{fug[i], alpha[i]};
  HIPCHK(hipMemcpyAsync(p->inst, iv.data(), p->nb * sizeof(InstDev), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

// --------------------------------------------------------------------------
// launch helpers
// --------------------------------------------------------------------------
// The compile-time y pass (k_ypass128): ny = 128 with the radix sequence 8,
// 4, 4 and a tile of 8, 13 or 16 columns
static bool ypass128_ok(const cmm_plan* p) {
  const Fft1D& f = p->g.fy;
  return p->ny == 128 && f.nst == 3 && f.rad[0] == 8 && f.rad[1] == 4 && f.rad[2] == 4 &&
         (p->tby == 8 || p->tby == 13 || p->tby == 16);
}

template <int TB>
static void launch_ypass128(cmm_plan* p, dim3 grid, size_t lds, double2* F, long long is, int ntz, int per, int sgn,
                            unsigned fmap, bool pair) {
  auto kern = sgn < 0 ? (pair ? k_ypass128<-1, true, TB> : k_ypass128<-1, false, TB>)
                      : (pair ? k_ypass128<1, true, TB> : k_ypass128<1, false, TB>);
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, p->stream, F, is, p->nhalf, p->g, ntz, per, p->active, fmap);
}

// y passes over nf fields (field stride nhalf, instance stride is); fmap != 0
// lists the fields, one byte per block row; pair: the PAIR kernel (field 1
// carries field 3, see k_ypass)
static int launch_ypass_is(cmm_plan* p, double2* F, long long is, int nf, int sgn, unsigned fmap = 0,
                           bool pair = false) {
  const int ntz = (p->nzh + p->tby - 1) / p->tby;
  const long long ntiles = (long long)p->nx * ntz;
  const int per = (int)((ntiles + 7) / 8);
  dim3 grid((unsigned)(8 * per), nf, p->nb);
  // the paired block of the runtime kernel takes its columns in two rounds
  // of ceil(tby / 2), two sequences each
  size_t lds = (size_t)p->ny * (pair ? 2 * ((p->tby + 1) / 2) : p->tby) * sizeof(double2);
  const FastDiv fT = make_fastdiv((unsigned)p->tby);
  if (ypass128_ok(p)) {
    lds = (size_t)p->ny * p->tby * sizeof(double2);
    if (p->tby == 8) launch_ypass128<8>(p, grid, lds, F, is, ntz, per, sgn, fmap, pair);
    else if (p->tby == 13) launch_ypass128<13>(p, grid, lds, F, is, ntz, per, sgn, fmap, pair);
    else launch_ypass128<16>(p, grid, lds, F, is, ntz, per, sgn, fmap, pair);
    LAUNCHCHK("k_ypass128");
    return CMM_OK;
  }
  auto kern = sgn < 0 ? (pair ? k_ypass<-1, true> : k_ypass<-1, false>)
                      : (pair ? k_ypass<1, true> : k_ypass<1, false>);
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, p->stream, F, is, p->nhalf, p->g, p->tby, fT, ntz, per, p->active,
                     fmap);
  LAUNCHCHK("k_ypass");
  return CMM_OK;
}

static int launch_ypass(cmm_plan* p, double2* F, int nf, int sgn) {
  return launch_ypass_is(p, F, (nf == 1 && F == p->rz) ? p->nhalf : (long long)p->fcap * p->nhalf, nf, sgn);
}

// y passes of the functional chain's fields.  Six-field functional: the z
// component of n_v2 (field 4) is i kz times n3 on both sides of the y and x
// passes and is not carried (k_zphi forms it, see z_load_half_multi), and
// its y component (field 3) is formed from n3 (forward chain) or folded into
// G_n3 (back chain) by the paired y pass, so the x passes neither write nor
// read it: three block rows (n2, n3 with n_v2,y, n_v2,x) plus rho_bar's
// unless the WDA is linear (wda_kspace: rho_bar is never formed).  `full`
// (callers that read the weighted densities) carries every field.
static bool wda_kspace(const cmm_plan* p) { return p->nf == 6 && p->fp.wl2 != 0.0; }

static int launch_ypass_chain(cmm_plan* p, int sgn, bool full = false) {
  const long long is = (long long)p->fcap * p->nhalf;
  if (p->nf == 6 && !full)
    return wda_kspace(p) ? launch_ypass_is(p, p->F, is, 3, sgn, 0x020100u, true)
                         : launch_ypass_is(p, p->F, is, 4, sgn, 0x05020100u, true);
  return launch_ypass_is(p, p->F, is, p->nf, sgn);
}

static int launch_xpass(cmm_plan* p, double2* F, long long is, int sgn) {
  const long long nb = (long long)p->ny * p->nzh;
  dim3 grid((unsigned)((nb + p->tbx - 1) / p->tbx), 1, p->nb);
  size_t lds = (size_t)p->nx * p->tbx * sizeof(double2);
  const FastDiv fT = make_fastdiv((unsigned)p->tbx);
  if (sgn < 0)
    hipLaunchKernelGGL(k_xpass<-1>, grid, dim3(NT), lds, p->stream, F, is, p->nhalf, p->g, p->tbx, fT, p->active);
  else
    hipLaunchKernelGGL(k_xpass<1>, grid, dim3(NT), lds, p->stream, F, is, p->nhalf, p->g, p->tbx, fT, p->active);
  LAUNCHCHK("k_xpass");
  return CMM_OK;
}

// fields an x pass weights and transforms (XFields: four of the six, three
// with a linear WDA) in the chain's layout nfx (5, 6 or 12)
static int xfields(int nfx) { return nfx == 6 ? XFields<6>::nc : (nfx == 5 ? XFields<5>::nc : nfx); }
static int chain_nf(const cmm_plan* p, bool full = false) { return (wda_kspace(p) && !full) ? 5 : p->nf; }

// x passes of the functional chain (see k_xweights / k_xback): tiles of
// nx * tb columns with all fields in LDS (nl nx tb complex values for nl
// fields, the nx twiddles and the reduction scratch, at most 64 KB together)
static int xpass_tb(cmm_plan* p, int nl) {
  const int lim = (65536 - CMM_RED_SH * 8 - 16 * p->nx) / (16 * nl * p->nx);
  return std::max(1, std::min(p->tbx, lim));
}

// threads of an x-pass block: one per column element of the tile (T = nx *
// tb), and at least enough for the in-LDS FFT of all nl fields (nl * T
// values at VPT per thread), rounded up to whole wavefronts, at most NTX (the
// LDS limit of xpass_tb keeps nl * T <= VPT * NTX).  A smaller tile is then a
// smaller block, so more tiles are resident per CU and their load and FFT
// phases overlap (measured at 128^3: tb 2 with 256 threads beats tb 4 with
// 512 threads by 8 %, tb 2 with 512 threads loses 24 %).
static unsigned xpass_threads(int T, int nl) {
  const int need = std::max(T, (nl * T + VPT - 1) / VPT);
  return (unsigned)std::min(NTX, ((need + 63) / 64) * 64);
}

// Full-size x-pass blocks (NTX threads, the 128^3 tile of 4 columns of six
// fields) hold ~53 KB of LDS, so three fit on a CU when the kernel is held to
// 6 waves per SIMD (80 VGPRs).  That spills ~25 VGPRs to scratch and is still
// faster (measured at 128^3: k_xweights -16 %, k_xback -11 % for one
// instance, -10 % / -13 % for 32).  Blocks whose LDS admits only two per CU
// gain nothing from it, and smaller blocks already fit five per CU: both keep
// the unconstrained build.
constexpr int XOCC_FULL = 6;
constexpr size_t CU_LDS = 160 * 1024;   // LDS of one gfx950 CU

static bool xpass_full_occ(unsigned nt, size_t lds_bytes) {
  return nt == (unsigned)NTX && 3 * lds_bytes <= CU_LDS;
}

// The compile-time x tile (k_xweights / k_xback with TBC = tb): nx = 128
// with the radix sequence 8, 4, 4 and two or four columns; 0 = runtime tile.
// Its blocks are 128 tb threads, which is what xpass_threads gives for the
// five- and six-field chains at those widths.
static int xpass_tbc(const cmm_plan* p, int tb) {
  const Fft1D& f = p->g.fx;
  if (p->nx != 128 || f.nst != 3 || f.rad[0] != 8 || f.rad[1] != 4 || f.rad[2] != 4) return 0;
  return (tb == 2 || tb == 4) ? tb : 0;
}

// returns the block size (its waves each store one partial per energy)
template <int NF, bool MFA>
static unsigned launch_xweights_nf(cmm_plan* p, int tb, long long ntiles, int per, bool full) {
  constexpr int NC = XFields<NF>::nc;
  const int T = p->nx * tb;
  const unsigned nt = xpass_threads(T, NC);
  // nx rows of NC tb values (padded rows for the compile-time tile), then the twiddles
  const int tbc = XFields<NF>::red ? xpass_tbc(p, tb) : 0;
  const size_t lds = (size_t)(p->nx * (tbc ? x_row(NC * tbc) : NC * tb) + p->nx) * sizeof(double2);
  const bool red = MFA || XFields<NF>::wlin;
  const bool occ = xpass_full_occ(nt, lds + (red ? CMM_RED_SH * sizeof(double) : 0));
  auto kern = occ ? k_xweights<NF, MFA, XOCC_FULL> : k_xweights<NF, MFA, 1>;
  if constexpr (XFields<NF>::red) {
    if (tbc == 4) kern = occ ? k_xweights<NF, MFA, XOCC_FULL, 4> : k_xweights<NF, MFA, 1, 4>;
    if (tbc == 2) kern = occ ? k_xweights<NF, MFA, XOCC_FULL, 2> : k_xweights<NF, MFA, 1, 2>;
  }
  dim3 grid((unsigned)(8 * per), 1, p->nb);
  hipLaunchKernelGGL(kern, grid, dim3(nt), lds, p->stream,
                     p->rz, p->nhalf, p->F, (long long)p->fcap * p->nhalf, p->nhalf, p->g, p->fp, tb,
                     make_fastdiv((unsigned)tb), make_fastdiv((unsigned)(NC * tb)), make_fastdiv((unsigned)p->nzh), per,
                     ntiles, p->active, red_of(p), full ? 1 : 0);
  return nt;
}

// Energies reduced here: the MFA energy into S_FMFA; with a linear WDA its
// sum f_ex(rho_bar) into S_PHI + 1 (the z pass then leaves that slot alone).
static int launch_xweights(cmm_plan* p, bool full = false) {
  const long long nb = (long long)p->ny * p->nzh;
  const int nfx = chain_nf(p, full);
  const int tb = xpass_tb(p, xfields(nfx));
  const long long ntiles = (nb + tb - 1) / tb;
  const int per = (int)((ntiles + 7) / 8);
  const bool mfa = p->fp.kmfa != nullptr;
  unsigned nt;
  if (nfx == 12) {
    nt = mfa ? launch_xweights_nf<12, true>(p, tb, ntiles, per, full) : launch_xweights_nf<12, false>(p, tb, ntiles, per, full);
  } else if (nfx == 5) {
    nt = mfa ? launch_xweights_nf<5, true>(p, tb, ntiles, per, full) : launch_xweights_nf<5, false>(p, tb, ntiles, per, full);
  } else {
    nt = mfa ? launch_xweights_nf<6, true>(p, tb, ntiles, per, full) : launch_xweights_nf<6, false>(p, tb, ntiles, per, full);
  }
  LAUNCHCHK("k_xweights");
  const long long nparts = ntiles * (long long)((nt + 63) / 64);   // one partial per wave
  if (mfa) {
    const int rc = launch_reduce(p, nparts, 1, 0u, S_FMFA);
    if (rc) return rc;
  }
  return nfx == 5 ? launch_reduce(p, nparts, 2, 0u, S_PHI, 1u) : CMM_OK;
}

template <int NF, bool MFA>
static void launch_xback_nf(cmm_plan* p, int tb, long long ntiles, int per) {
  constexpr int NL = XFields<NF>::nc + ((MFA || XFields<NF>::wlin) ? 1 : 0);
  const int T = p->nx * tb;
  const unsigned nt = xpass_threads(T, NL);
  const int tbc = XFields<NF>::red ? xpass_tbc(p, tb) : 0;
  const size_t lds = (size_t)(p->nx * (tbc ? x_row(NL * tbc) : NL * tb) + p->nx) * sizeof(double2);
  const bool occ = xpass_full_occ(nt, lds);
  auto kern = occ ? k_xback<NF, MFA, XOCC_FULL> : k_xback<NF, MFA, 1>;
  if constexpr (XFields<NF>::red) {
    if (tbc == 4) kern = occ ? k_xback<NF, MFA, XOCC_FULL, 4> : k_xback<NF, MFA, 1, 4>;
    if (tbc == 2) kern = occ ? k_xback<NF, MFA, XOCC_FULL, 2> : k_xback<NF, MFA, 1, 2>;
  }
  dim3 grid((unsigned)(8 * per), 1, p->nb);
  hipLaunchKernelGGL(kern, grid, dim3(nt), lds, p->stream, p->F,
                     (long long)p->fcap * p->nhalf, p->nhalf, p->g, p->fp, tb, make_fastdiv((unsigned)tb),
                     make_fastdiv((unsigned)(NL * tb)), make_fastdiv((unsigned)T), make_fastdiv((unsigned)p->nzh), per,
                     ntiles, p->active, p->rz, p->nhalf);
}

static int launch_xback(cmm_plan* p) {
  const long long nb = (long long)p->ny * p->nzh;
  const bool mfa = p->fp.kmfa != nullptr;
  const int nfx = chain_nf(p);
  const int tb = xpass_tb(p, xfields(nfx) + ((mfa || nfx == 5) ? 1 : 0));
  const long long ntiles = (nb + tb - 1) / tb;
  const int per = (int)((ntiles + 7) / 8);
  if (nfx == 12) {
    if (mfa) launch_xback_nf<12, true>(p, tb, ntiles, per);
    else launch_xback_nf<12, false>(p, tb, ntiles, per);
  } else if (nfx == 5) {
    if (mfa) launch_xback_nf<5, true>(p, tb, ntiles, per);
    else launch_xback_nf<5, false>(p, tb, ntiles, per);
  } else {
    if (mfa) launch_xback_nf<6, true>(p, tb, ntiles, per);
    else launch_xback_nf<6, false>(p, tb, ntiles, per);
  }
  LAUNCHCHK("k_xback");
  return CMM_OK;
}

// z-pass pencils of the runtime-tile functional passes (k_zphi, k_zhess):
// the plan's p_phi within the staging and LDS limits of six fields; for
// twelve, half of it (at least 2) within those of twelve
static int zphi_p(cmm_plan* p) {
  int P6 = p->pphi;
  while (P6 > 2 && (3 * P6 * p->nz > VPT * NT || zlds(p, P6, 6) > 65536)) P6 -= 2;
  if (p->nf != 12) return P6;
  int P = std::max(2, p->pphi / 2);
  P -= P & 1;
  while (P > 2 && (6 * P * p->nz > VPT * NT || zlds(p, P, 12) > 65536)) P -= 2;
  return P;
}

int cmm_fft_r2c(cmm_plan* p, const double* x, double* X) {
  if (!p || !x || !X) return fail(CMM_E_ARG, "null argument");
  double2* Xc = reinterpret_cast<double2*>(X);
  dim3 grid((unsigned)zblocks(p, p->pz), p->nb);
  hipLaunchKernelGGL(k_zr2c, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, x, p->npts, Xc, p->nhalf, p->g,
                     ztile(p, p->pz, 1), p->active);
  LAUNCHCHK("k_zr2c");
  // y and x passes on the caller's buffer (instance stride nhalf)
  int rc = launch_ypass_is(p, Xc, p->nhalf, 1, -1);
  if (rc) return rc;
  return launch_xpass(p, Xc, p->nhalf, -1);
}

int cmm_fft_c2r(cmm_plan* p, const double* X, double* x) {
  if (!p || !x || !X) return fail(CMM_E_ARG, "null argument");
  // work on a copy in the workspace (rz) so the caller's spectrum survives
  HIPCHK(hipMemcpyAsync(p->rz, X, (size_t)p->nhalf * p->nb * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
  int rc = launch_xpass(p, p->rz, p->nhalf, 1);
  if (rc) return rc;
  rc = launch_ypass_is(p, p->rz, p->nhalf, 1, 1);
  if (rc) return rc;
  dim3 grid((unsigned)zblocks(p, p->pz), 1, p->nb);
  hipLaunchKernelGGL(k_zc2r, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->rz, p->nhalf, 0LL, x, p->npts, 0LL,
                     p->g, ztile(p, p->pz, 1), p->active);
  LAUNCHCHK("k_zc2r");
  return CMM_OK;
}

static ElDev el_base(cmm_plan* p, const double* rho, const double* vext, int vb, double lam3) {
  ElDev el;
  std::memset(&el, 0, sizeof(el));
  el.rho = rho;
  el.vext = vext;
  el.vext_batched = vb;
  el.fstride = p->npts;
  el.inst = p->inst;
  el.lam3 = lam3;
  el.rz = p->rz;
  el.rzstride = p->nhalf;
  return el;
}

template <int MODE>
static int launch_produce(cmm_plan* p, const ElDev& el, const char* name) {
  dim3 grid((unsigned)zblocks(p, p->pz), p->nb);
  hipLaunchKernelGGL(k_zproduce<MODE>, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->g, el, ztile(p, p->pz, 1),
                     p->active, red_of(p));
  LAUNCHCHK(name);
  return launch_reduce(p, grid.x, 7, 1u << 2, S_PN);
}

int cmm_seed(cmm_plan* p, const double* rho, const double* vext, int vb, double lam3) {
  if (!p || !rho) return fail(CMM_E_ARG, "null argument");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  return launch_produce<MODE_SEED>(p, el, "k_zproduce(seed)");
}

static int forward_chain(cmm_plan* p, bool full = false) {
  if (!p->has_func) return fail(CMM_E_ARG, "functional not set");
  int rc = launch_ypass(p, p->rz, 1, -1);
  if (rc) return rc;
  rc = launch_xweights(p, full);
  if (rc) return rc;
  return launch_ypass_chain(p, 1, full);
}

// nz = 128 and five or six fields: the compile-time tile kernel
template <int NF>
static bool launch_zphi128(cmm_plan* p, int store) {
  const int P = p->pphi;
  dim3 grid((unsigned)zblocks(p, P), p->nb);
  const size_t lds = zlds(p, P, NF);
  const long long Fs = (long long)p->fcap * p->nhalf;
  switch (P) {
    case 2: hipLaunchKernelGGL((k_zphi128<NF, 2>), grid, dim3(NT), lds, p->stream, p->F, Fs, p->nhalf, p->g, p->fp,
                               store, p->active, red_of(p)); break;
    case 4: hipLaunchKernelGGL((k_zphi128<NF, 4>), grid, dim3(NT), lds, p->stream, p->F, Fs, p->nhalf, p->g, p->fp,
                               store, p->active, red_of(p)); break;
    case 6: hipLaunchKernelGGL((k_zphi128<NF, 6>), grid, dim3(NT), lds, p->stream, p->F, Fs, p->nhalf, p->g, p->fp,
                               store, p->active, red_of(p)); break;
    case 8: hipLaunchKernelGGL((k_zphi128<NF, 8>), grid, dim3(NT), lds, p->stream, p->F, Fs, p->nhalf, p->g, p->fp,
                               store, p->active, red_of(p)); break;
    default: return false;
  }
  return true;
}

static int launch_zphi(cmm_plan* p, int store, bool reduce = true) {
  const bool wl = wda_kspace(p);
  if (p->nz == 128 && p->nf == 6) {
    const int P = p->pphi;
    if (wl ? launch_zphi128<5>(p, store) : launch_zphi128<6>(p, store)) {
      LAUNCHCHK("k_zphi128");
      // one partial per wave (wave_reduce_store)
      return reduce ? launch_reduce(p, zblocks(p, P) * (NT / 64), 3, 1u << 2, S_PHI, wl ? 2u : 0u) : CMM_OK;
    }
  }
  const int P = zphi_p(p);
  dim3 grid((unsigned)zblocks(p, P), p->nb);
  if (p->nf == 12)
    hipLaunchKernelGGL(k_zphi<12>, grid, dim3(NT), zlds(p, P, 12), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, p->g, p->fp, ztile(p, P, 12), store, p->active, red_of(p));
  else if (wl)
    hipLaunchKernelGGL(k_zphi<5>, grid, dim3(NT), zlds(p, P, 5), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, p->g, p->fp, ztile(p, P, 5), store, p->active, red_of(p));
  else
    hipLaunchKernelGGL(k_zphi<6>, grid, dim3(NT), zlds(p, P, 6), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, p->g, p->fp, ztile(p, P, 6), store, p->active, red_of(p));
  LAUNCHCHK("k_zphi");
  return reduce ? launch_reduce(p, (long long)grid.x * (NT / 64), 3, 1u << 2, S_PHI, wl ? 2u : 0u) : CMM_OK;
}

// back-convolution of the nf derivative fields in F: inverse y, the x pass
// with the weighted sum, inverse y of the sum (left in field 0 of F)
static int back_chain(cmm_plan* p) {
  int rc = launch_ypass_chain(p, -1);
  if (rc) return rc;
  rc = launch_xback(p);
  if (rc) return rc;
  return launch_ypass_is(p, p->F, (long long)p->fcap * p->nhalf, 1, 1);
}

int cmm_derive(cmm_plan* p) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  int rc = forward_chain(p);
  if (rc) return rc;
  rc = launch_zphi(p, 1);
  if (rc) return rc;
  return back_chain(p);
}

int cmm_hessian_apply(cmm_plan* p, const double* nfld) {
  if (!p || !nfld) return fail(CMM_E_ARG, "null argument");
  int rc = forward_chain(p);
  if (rc) return rc;
  const int P = zphi_p(p);
  dim3 grid((unsigned)zblocks(p, P), p->nb);
  if (p->nf == 12)
    hipLaunchKernelGGL(k_zhess<12>, grid, dim3(NT), zlds(p, P, 12), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, nfld, p->npts, p->g, p->fp, ztile(p, P, 12), p->active);
  else if (wda_kspace(p))
    hipLaunchKernelGGL(k_zhess<5>, grid, dim3(NT), zlds(p, P, 5), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, nfld, p->npts, p->g, p->fp, ztile(p, P, 5), p->active);
  else
    hipLaunchKernelGGL(k_zhess<6>, grid, dim3(NT), zlds(p, P, 6), p->stream, p->F, (long long)p->fcap * p->nhalf,
                       p->nhalf, nfld, p->npts, p->g, p->fp, ztile(p, P, 6), p->active);
  LAUNCHCHK("k_zhess");
  return back_chain(p);
}

int cmm_value(cmm_plan* p) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  int rc = forward_chain(p);
  if (rc) return rc;
  return launch_zphi(p, 0);
}

int cmm_launch_pass(cmm_plan* p, int which) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (!p->has_func) return fail(CMM_E_ARG, "functional not set");
  switch (which) {
    case 0: return launch_zphi(p, 1, false);
    case 1: return launch_xweights(p);
    case 2: return launch_xback(p);
    case 3: return launch_ypass_chain(p, -1);
    case 4: return launch_ypass(p, p->rz, 1, -1);
    default: return fail(CMM_E_ARG, "unknown pass");
  }
}

int cmm_weighted_densities(cmm_plan* p, double* out) {
  if (!p || !out) return fail(CMM_E_ARG, "null argument");
  int rc = forward_chain(p, true);
  if (rc) return rc;
  dim3 grid((unsigned)zblocks(p, p->pz), p->nf, p->nb);
  hipLaunchKernelGGL(k_zc2r, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->F, (long long)p->fcap * p->nhalf,
                     p->nhalf, out, (long long)p->nf * p->npts, p->npts, p->g, ztile(p, p->pz, 1), p->active);
  LAUNCHCHK("k_zc2r");
  return CMM_OK;
}

template <int MODE>
static int launch_final(cmm_plan* p, const ElDev& el, const char* name) {
  dim3 grid((unsigned)zblocks(p, p->pz), p->nb);
  hipLaunchKernelGGL(k_zfinal<MODE>, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->F, (long long)p->fcap * p->nhalf,
                     p->g, el,
                     ztile(p, p->pz, 1), p->active, red_of(p));
  LAUNCHCHK(name);
  constexpr int NQ = (MODE == MODE_ANDRES) ? (4 + CMM_MMAX) : 8;
  return launch_reduce(p, grid.x, NQ, MODE == MODE_PICARD ? (1u << 5) : 0u, S_NCUR);
}

int cmm_local_density(cmm_plan* p, const double* nfld, const double* rbar, double* phi, double* fex, double* dfex) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (!p->has_func) return fail(CMM_E_ARG, "functional not set");
  if (nfld && phi && !p->fp.use_hs) return fail(CMM_E_ARG, "the functional has no hard-sphere part");
  if (rbar && (fex || dfex) && !p->fp.use_wda) return fail(CMM_E_ARG, "the functional has no WDA part");
  const long long n = p->npts;
  const unsigned nb = (unsigned)std::min<long long>((n + NT - 1) / NT, 65536);
  for (int b = 0; b < p->nb; ++b) {
    const double* nf = nfld ? nfld + (long long)b * p->nf * n : nullptr;
    const double* rb = rbar ? rbar + (long long)b * n : nullptr;
    double* ph = phi ? phi + (long long)b * n : nullptr;
    double* f = fex ? fex + (long long)b * n : nullptr;
    double* df = dfex ? dfex + (long long)b * n : nullptr;
    if (p->nf == 12)
      hipLaunchKernelGGL(k_local<12>, dim3(nb), dim3(NT), 0, p->stream, nf, n, rb, n, p->fp, ph, f, df);
    else
      hipLaunchKernelGGL(k_local<6>, dim3(nb), dim3(NT), 0, p->stream, nf, n, rb, n, p->fp, ph, f, df);
    LAUNCHCHK("k_local");
  }
  return CMM_OK;
}

int cmm_hs_weight_table(cmm_plan* p, double R, double* out) {
  if (!p || !out) return fail(CMM_E_ARG, "null argument");
  if (!(R > 0.0)) return fail(CMM_E_ARG, "hard-sphere radius must be positive");
  const unsigned nb = (unsigned)std::min<long long>((p->npts + NT - 1) / NT, 65536);
  hipLaunchKernelGGL(k_weight_table, dim3(nb), dim3(NT), 0, p->stream, p->g, R, out);
  LAUNCHCHK("k_weight_table");
  return CMM_OK;
}

int cmm_finish_c1(cmm_plan* p, const double* rho, const double* vext, int vb, double* c1, double lam3) {
  if (!p || !rho || !c1) return fail(CMM_E_ARG, "null argument");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  el.c1out = c1;
  return launch_final<MODE_C1>(p, el, "k_zfinal(c1)");
}

int cmm_finish_picard(cmm_plan* p, const double* rho, const double* vext, int vb, double* rho_out, double* c1_out,
                      double beta, double lam3, double zero_thr) {
  if (!p || !rho || !rho_out) return fail(CMM_E_ARG, "null argument");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  el.rho_out = rho_out;
  el.c1out = c1_out;
  el.beta = beta;
  el.zero_thr = zero_thr;
  return launch_final<MODE_PICARD>(p, el, "k_zfinal(picard)");
}

int cmm_finish_anderson(cmm_plan* p, const double* rho, const double* vext, int vb, double* rhist, int m, int slot_new,
                        double* c1_out, double beta, double lam3) {
  if (!p || !rho || !rhist) return fail(CMM_E_ARG, "null argument");
  if (m < 1 || m > CMM_MMAX || slot_new < 0 || slot_new >= m) return fail(CMM_E_ARG, "bad history slot");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  el.rhist = rhist;
  el.hstride = (long long)p->nb * p->npts;
  el.m = m;
  el.slot_new = slot_new;
  el.c1out = c1_out;
  el.beta = beta;
  return launch_final<MODE_ANDRES>(p, el, "k_zfinal(anderson)");
}

int cmm_finish_check(cmm_plan* p, const double* rho, const double* vext, int vb, double beta, double lam3,
                     double a_tol, double r_tol) {
  if (!p || !rho) return fail(CMM_E_ARG, "null argument");
  if (!(beta > 0.0) || !(lam3 > 0.0)) return fail(CMM_E_ARG, "beta and lam3 must be positive");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  el.beta = beta;
  el.a_tol = a_tol;
  el.r_tol = r_tol;
  return launch_final<MODE_CHECK>(p, el, "k_zfinal(check)");
}

int cmm_boltzmann(const double* c1, long long n, double beta, double fug, double* out, void* stream) {
  if (!c1 || !out || n < 0) return fail(CMM_E_ARG, "null argument");
  if (n == 0) return CMM_OK;
  const long long want = (n + NT - 1) / NT;
  const unsigned nb = (unsigned)std::min<long long>(want, 65536);
  hipLaunchKernelGGL(k_boltzmann, dim3(nb), dim3(NT), 0, (hipStream_t)stream, c1, n, beta, fug, out);
  LAUNCHCHK("k_boltzmann");
  return CMM_OK;
}

int cmm_anderson_store_row(cmm_plan* p, int m, int slot_new) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (m < 1 || m > CMM_MMAX || slot_new < 0 || slot_new >= m) return fail(CMM_E_ARG, "bad history slot");
  AndDev a;
  std::memset(&a, 0, sizeof(a));
  a.gram = p->gram;
  a.m = m;
  a.mk = 1;
  a.slot_new = slot_new;
  a.slot_prev = -1;
  a.store_only = 1;
  hipLaunchKernelGGL(k_anderson_coef, dim3(p->nb), dim3(64), 0, p->stream, a, p->active, red_of(p));
  LAUNCHCHK("k_anderson_coef(store)");
  return CMM_OK;
}

int cmm_anderson_init(cmm_plan* p, const double* damping, const double* cf) {
  if (!p || !damping || !cf) return fail(CMM_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(p->damping, damping, p->nb * sizeof(double), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->cf, cf, p->nb * sizeof(double), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemsetAsync(p->gram, 0, (size_t)p->nb * CMM_MMAX * CMM_MMAX * sizeof(double), p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

int cmm_anderson_coef(cmm_plan* p, int m, int mk, const int* order, int slot_new, int slot_prev, int adaptive,
                      double fac_up, double fac_down, double dmin, double dmax) {
  if (!p || !order) return fail(CMM_E_ARG, "null argument");
  if (m < 1 || m > CMM_MMAX || mk < 1 || mk > m) return fail(CMM_E_ARG, "bad history length");
  AndDev a;
  std::memset(&a, 0, sizeof(a));
  a.gram = p->gram;
  a.m = m;
  a.mk = mk;
  for (int i = 0; i < mk; ++i) {
    if (order[i] < 0 || order[i] >= m) return fail(CMM_E_ARG, "bad history order");
    a.order[i] = order[i];
  }
  a.slot_new = slot_new;
  a.slot_prev = slot_prev;
  a.damping = p->damping;
  a.cf = p->cf;
  a.adaptive = adaptive;
  a.fac_up = fac_up;
  a.fac_down = fac_down;
  a.dmin = dmin;
  a.dmax = dmax;
  a.coef = p->coef;
  a.alpha_out = p->alpha_out;
  hipLaunchKernelGGL(k_anderson_coef, dim3(p->nb), dim3(64), 0, p->stream, a, p->active, red_of(p));
  LAUNCHCHK("k_anderson_coef");
  return CMM_OK;
}

int cmm_anderson_mix(cmm_plan* p, const double* rhohist, const int* order_rho, const double* rhist, const int* order,
                     int mk, const double* rho_cur, double* rho_out, const double* vext, int vb, double lam3) {
  if (!p || !rhohist || !rhist || !order || !order_rho || !rho_cur || !rho_out) return fail(CMM_E_ARG, "null argument");
  if (mk < 1 || mk > CMM_MMAX) return fail(CMM_E_ARG, "bad history length");
  ElDev el = el_base(p, rho_cur, vext, vb, lam3);
  el.rho_out = rho_out;
  el.rhohist = rhohist;
  el.hstride2 = (long long)p->nb * p->npts;
  el.rhist = const_cast<double*>(rhist);
  el.hstride = (long long)p->nb * p->npts;
  el.coef = p->coef;
  el.mk = mk;
  for (int i = 0; i < mk; ++i) {
    el.order[i] = order[i];
    el.order_rho[i] = order_rho[i];
  }
  return launch_produce<MODE_MIX>(p, el, "k_zproduce(mix)");
}

int cmm_anderson_set_damping(cmm_plan* p, const double* damping) {
  if (!p || !damping) return fail(CMM_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(p->damping, damping, p->nb * sizeof(double), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

int cmm_anderson_state(cmm_plan* p, double* damping, double* alpha) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (damping) HIPCHK(hipMemcpyAsync(damping, p->damping, p->nb * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  if (alpha)
    HIPCHK(hipMemcpyAsync(alpha, p->alpha_out, p->nb * CMM_MMAX * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

int cmm_trial(cmm_plan* p, const double* rho, const double* rhist, int slot, double* rho_out, const double* vext, int vb,
              double lam3, int clip) {
  if (!p || !rho || !rhist || !rho_out) return fail(CMM_E_ARG, "null argument");
  ElDev el = el_base(p, rho, vext, vb, lam3);
  el.rho_out = rho_out;
  el.rhist = const_cast<double*>(rhist);
  el.hstride = (long long)p->nb * p->npts;
  el.slot_new = slot;
  el.clip_trial = clip;
  return launch_produce<MODE_TRIAL>(p, el, "k_zproduce(trial)");
}

int cmm_get_scalars(cmm_plan* p, double* out) {
  if (!p || !out) return fail(CMM_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(out, p->scal, (size_t)NSCAL * p->nb * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

int cmm_synchronize(cmm_plan* p) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

static AdjDev adj_base(cmm_plan* p, const double* rho, double beta) {
  AdjDev a;
  std::memset(&a, 0, sizeof(a));
  a.rho = rho;
  a.fstride = p->npts;
  a.beta = beta;
  return a;
}

int cmm_adjoint_seed(cmm_plan* p, const double* rho, double beta, double dr, int init, int parity, double* y,
                     double* r, double* pvec) {
  if (!p || !rho || !r || !pvec || (init && !y)) return fail(CMM_E_ARG, "null argument");
  if (!(beta > 0.0)) return fail(CMM_E_ARG, "beta must be positive");
  AdjDev a = adj_base(p, rho, beta);
  a.dr = dr;
  a.y = y;
  a.r = r;
  a.p = pvec;
  a.parity = parity & 1;
  dim3 grid((unsigned)zblocks(p, p->pz), p->nb);
  if (init)
    hipLaunchKernelGGL(k_adj_seed<ADJ_INIT>, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->g, a,
                       ztile(p, p->pz, 1), p->rz, p->nhalf, p->active, red_of(p));
  else
    hipLaunchKernelGGL(k_adj_seed<ADJ_STEP>, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->g, a,
                       ztile(p, p->pz, 1), p->rz, p->nhalf, p->active, red_of(p));
  LAUNCHCHK("k_adj_seed");
  return init ? launch_reduce(p, grid.x, 1, 0u, S_ADJ_RR) : CMM_OK;
}

int cmm_adjoint_final(cmm_plan* p, const double* rho, double beta, const double* pvec, double* q) {
  if (!p || !rho || !pvec || !q) return fail(CMM_E_ARG, "null argument");
  AdjDev a = adj_base(p, rho, beta);
  a.p = const_cast<double*>(pvec);
  a.q = q;
  dim3 grid((unsigned)zblocks(p, p->pz), p->nb);
  hipLaunchKernelGGL(k_adj_final, grid, dim3(NT), zlds(p, p->pz, 1), p->stream, p->F, (long long)p->fcap * p->nhalf,
                     p->g, a,
                     ztile(p, p->pz, 1), p->active, red_of(p));
  LAUNCHCHK("k_adj_final");
  return launch_reduce(p, grid.x, 1, 0u, S_ADJ_PQ);
}

static dim3 flat_grid(cmm_plan* p) {
  long long nbk = (p->npts + NT * 8 - 1) / (NT * 8);
  nbk = std::min(nbk, p->pstride);
  return dim3((unsigned)std::max(1LL, nbk), p->nb);
}

int cmm_adjoint_update(cmm_plan* p, int parity, double* y, double* r, const double* pvec, const double* q) {
  if (!p || !y || !r || !pvec || !q) return fail(CMM_E_ARG, "null argument");
  AdjDev a = adj_base(p, nullptr, 1.0);
  a.y = y;
  a.r = r;
  a.p = const_cast<double*>(pvec);
  a.q = const_cast<double*>(q);
  a.parity = parity & 1;
  const dim3 grid = flat_grid(p);
  hipLaunchKernelGGL(k_adj_update, grid, dim3(NT), 0, p->stream, a, p->active, red_of(p));
  LAUNCHCHK("k_adj_update");
  return launch_reduce(p, grid.x, 1, 0u, S_ADJ_RR + ((parity & 1) ^ 1));
}

int cmm_adjoint_lambda(cmm_plan* p, const double* rho, double beta, const double* y, double* lam) {
  if (!p || !rho || !y || !lam) return fail(CMM_E_ARG, "null argument");
  AdjDev a = adj_base(p, rho, beta);
  a.y = const_cast<double*>(y);
  a.lam = lam;
  const dim3 grid = flat_grid(p);
  hipLaunchKernelGGL(k_adj_lambda, grid, dim3(NT), 0, p->stream, a, p->active, red_of(p));
  LAUNCHCHK("k_adj_lambda");
  return launch_reduce(p, grid.x, 1, 0u, S_ADJ_LAM);
}

static int lj_setup(int nx, int ny, int nz, const double* rvecs, double sigma_f, double epsilon_f, double cutoff,
                    int images, double limit, LJDev& d) {
  if (!rvecs) return fail(CMM_E_ARG, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(CMM_E_ARG, "bad sizes");
  if (!(cutoff > 0.0)) return fail(CMM_E_ARG, "cutoff must be positive");
  std::memset(&d, 0, sizeof(d));
  d.nx = nx; d.ny = ny; d.nz = nz;
  std::memcpy(d.rv, rvecs, 9 * sizeof(double));
  for (int a = 0; a < 3; ++a) {
    d.L[a] = std::sqrt(rvecs[3 * a] * rvecs[3 * a] + rvecs[3 * a + 1] * rvecs[3 * a + 1] + rvecs[3 * a + 2] * rvecs[3 * a + 2]);
    const int n = a == 0 ? nx : (a == 1 ? ny : nz);
    d.inv_n[a] = 1.0 / n;
  }
  if (images) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b && rvecs[3 * a + b] != 0.0) return fail(CMM_E_UNSUPPORTED, "image sum needs an orthorhombic cell");
    for (int a = 0; a < 3; ++a) d.nimg[a] = (int)std::ceil(cutoff / d.L[a]) + 1;
  }
  d.rc = cutoff;
  d.sigf = sigma_f;
  d.epsf = epsilon_f;
  d.images = images ? 1 : 0;
  d.limit = limit;
  return CMM_OK;
}

int cmm_lj_grid(int nx, int ny, int nz, const double* rvecs, const double* atoms, int natom, double sigma_f,
                double epsilon_f, double cutoff, int images, double limit, double* V, void* stream) {
  if (!V || (natom > 0 && !atoms) || natom < 0) return fail(CMM_E_ARG, "null argument");
  LJDev d;
  int rc = lj_setup(nx, ny, nz, rvecs, sigma_f, epsilon_f, cutoff, images, limit, d);
  if (rc) return rc;
  dim3 grid((nz + LJ_BZ - 1) / LJ_BZ, (ny + LJ_BY - 1) / LJ_BY, (nx + LJ_BX - 1) / LJ_BX);
  hipLaunchKernelGGL(k_lj, grid, dim3(LJ_NT), 0, (hipStream_t)stream, atoms, natom, d, V);
  LAUNCHCHK("k_lj");
  return CMM_OK;
}

int cmm_lj_points(const double* points, long long M, const double* rvecs, const double* atoms, int natom,
                  double sigma_f, double epsilon_f, double cutoff, int images, double limit, double* V, void* stream) {
  if (!V || !points || (natom > 0 && !atoms) || natom < 0 || M < 0) return fail(CMM_E_ARG, "null argument");
  if (M == 0) return CMM_OK;
  LJDev d;
  int rc = lj_setup(1, 1, 1, rvecs, sigma_f, epsilon_f, cutoff, images, limit, d);
  if (rc) return rc;
  dim3 grid((unsigned)((M + LJ_NT - 1) / LJ_NT));
  hipLaunchKernelGGL(k_lj_points, grid, dim3(LJ_NT), 0, (hipStream_t)stream, atoms, natom, d, points, M, V);
  LAUNCHCHK("k_lj_points");
  return CMM_OK;
}

static dim3 lj_grid_dims(int nx, int ny, int nz) {
  return dim3((nz + LJ_BZ - 1) / LJ_BZ, (ny + LJ_BY - 1) / LJ_BY, (nx + LJ_BX - 1) / LJ_BX);
}

int cmm_effpot(int nx, int ny, int nz, const double* points, long long M, const double* rvecs, const double* atoms,
               int natom, int ng, const int* gtype, int nt, const double* gsig, const double* geps,
               const double* offsets, const double* weights, int nrot, double dmax, double beta, double cutoff,
               int images, double nanval, double* V, void* stream) {
  if (!V || !gtype || !gsig || !geps || !offsets || !weights || (natom > 0 && !atoms) || natom < 0)
    return fail(CMM_E_ARG, "null argument");
  if (ng < 1 || ng > EP_MAXG) return fail(CMM_E_ARG, "guest atom count must be in [1, 16]");
  if (nt < 1 || nt > EP_MAXT) return fail(CMM_E_ARG, "guest atom type count must be in [1, 4]");
  if (nrot < 1) return fail(CMM_E_ARG, "no orientations");
  if (!(beta > 0.0)) return fail(CMM_E_ARG, "beta must be positive");
  if (!(dmax >= 0.0)) return fail(CMM_E_ARG, "dmax must be non-negative");
  EffDev e;
  std::memset(&e, 0, sizeof(e));
  const bool grid = points == nullptr;
  if (!grid && M < 0) return fail(CMM_E_ARG, "negative point count");
  int rc = lj_setup(grid ? nx : 1, grid ? ny : 1, grid ? nz : 1, rvecs, 0.0, 0.0, cutoff, images, 0.0, e.d);
  if (rc) return rc;
  e.nrot = nrot;
  e.ng = ng;
  e.nt = nt;
  for (int i = 0; i < ng; ++i) {
    if (gtype[i] < 0 || gtype[i] >= nt) return fail(CMM_E_ARG, "guest atom type out of range");
    e.gtype[i] = gtype[i];
  }
  for (int t = 0; t < nt; ++t) {
    e.gsig[t] = gsig[t];
    e.geps[t] = geps[t];
  }
  e.dmax = dmax;
  e.beta = beta;
  e.nanval = nanval;
  hipStream_t s = (hipStream_t)stream;
  if (grid) {
    dim3 g((nz + EP_BZ - 1) / EP_BZ, (ny + EP_BY - 1) / EP_BY, (nx + EP_BX - 1) / EP_BX);
    if (images) hipLaunchKernelGGL((k_effpot<true, true>), g, dim3(EP_NT), 0, s, atoms, natom, e, offsets, weights,
                                   nullptr, 0LL, V);
    else hipLaunchKernelGGL((k_effpot<true, false>), g, dim3(EP_NT), 0, s, atoms, natom, e, offsets, weights,
                            nullptr, 0LL, V);
  } else {
    if (M == 0) return CMM_OK;
    dim3 g((unsigned)((M + EP_NT - 1) / EP_NT));
    if (images) hipLaunchKernelGGL((k_effpot<false, true>), g, dim3(EP_NT), 0, s, atoms, natom, e, offsets, weights,
                                   points, M, V);
    else hipLaunchKernelGGL((k_effpot<false, false>), g, dim3(EP_NT), 0, s, atoms, natom, e, offsets, weights,
                            points, M, V);
  }
  LAUNCHCHK("k_effpot");
  return CMM_OK;
}

int cmm_lj_derivs(int nx, int ny, int nz, const double* points, long long M, const double* rvecs,
                  const double* atoms, int natom, double sigma_f, double epsilon_f, double cutoff, int images,
                  const double* spacings, double vmax, double* out, void* stream) {
  if (!out || !spacings || (natom > 0 && !atoms) || natom < 0) return fail(CMM_E_ARG, "null argument");
  const bool grid = points == nullptr;
  if (!grid && M < 0) return fail(CMM_E_ARG, "negative point count");
  LJDev d;
  int rc = lj_setup(grid ? nx : 1, grid ? ny : 1, grid ? nz : 1, rvecs, sigma_f, epsilon_f, cutoff, images, 0.0, d);
  if (rc) return rc;
  LJDerivArgs da;
  for (int a = 0; a < 3; ++a) da.h[a] = spacings[a];
  da.vmax = vmax;
  hipStream_t s = (hipStream_t)stream;
  if (grid) {
    dim3 g((nz + LJ_BZ - 1) / LJ_BZ, (ny + LJ_BY - 1) / LJ_BY, (nx + LJ_BX - 1) / LJ_BX);
    hipLaunchKernelGGL((k_lj_derivs<true>), g, dim3(LJ_NT), 0, s, atoms, natom, d, da, nullptr, 0LL, out);
  } else {
    if (M == 0) return CMM_OK;
    dim3 g((unsigned)((M + LJD_NT - 1) / LJD_NT));
    hipLaunchKernelGGL((k_lj_derivs<false>), g, dim3(LJD_NT), 0, s, atoms, natom, d, da, points, M, out);
  }
  LAUNCHCHK("k_lj_derivs");
  return CMM_OK;
}

int cmm_interp_table(const double* values, int nx, int ny, int nz, int estimate, double* table, void* stream) {
  if (!values || !table) return fail(CMM_E_ARG, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(CMM_E_ARG, "bad sizes");
  const long long n = (long long)nx * ny * nz;
  dim3 g((unsigned)((n + 255) / 256));
  if (estimate) hipLaunchKernelGGL(k_interp_table<1>, g, dim3(256), 0, (hipStream_t)stream, values, nx, ny, nz, table);
  else hipLaunchKernelGGL(k_interp_table<0>, g, dim3(256), 0, (hipStream_t)stream, values, nx, ny, nz, table);
  LAUNCHCHK("k_interp_table");
  return CMM_OK;
}

static int interp_tab(const double* table, const int* n, const double* origin, const double* spacing, int mode,
                      InterpTab& T) {
  if (!table || !n || !origin || !spacing) return fail(CMM_E_ARG, "null argument");
  if (mode != 0 && mode != 1) return fail(CMM_E_ARG, "interpolation mode must be 0 (tricubic) or 1 (trilinear)");
  T.tab = table;
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 1) return fail(CMM_E_ARG, "bad table sizes");
    if (!(spacing[a] > 0.0)) return fail(CMM_E_ARG, "table spacing must be positive");
    T.n[a] = n[a];
    T.origin[a] = origin[a];
    T.h[a] = spacing[a];
  }
  T.mode = mode;
  return CMM_OK;
}

int cmm_interp_points(const double* table, const int* n, const double* origin, const double* spacing, int mode,
                      const double* points, long long M, double* out, void* stream) {
  if (!points || !out || M < 0) return fail(CMM_E_ARG, "null argument");
  InterpTab T;
  int rc = interp_tab(table, n, origin, spacing, mode, T);
  if (rc) return rc;
  if (M == 0) return CMM_OK;
  hipLaunchKernelGGL(k_interp_points, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, (hipStream_t)stream, T, points,
                     M, out);
  LAUNCHCHK("k_interp_points");
  return CMM_OK;
}

int cmm_effpot_interp(int nx, int ny, int nz, const double* points, long long M, const double* rvecs, int ng,
                      const int* gtype, int nt, const double* const* tables, const int* tn, const double* torigin,
                      const double* tspacing, const int* tmode, const double* offsets, const double* weights,
                      int nrot, double beta, double nanval, double* V, void* stream) {
  if (!V || !gtype || !tables || !tn || !torigin || !tspacing || !tmode || !offsets || !weights)
    return fail(CMM_E_ARG, "null argument");
  if (ng < 1 || ng > EP_MAXG) return fail(CMM_E_ARG, "guest atom count must be in [1, 16]");
  if (nt < 1 || nt > EP_MAXT) return fail(CMM_E_ARG, "guest atom type count must be in [1, 4]");
  if (nrot < 1) return fail(CMM_E_ARG, "no orientations");
  if (!(beta > 0.0)) return fail(CMM_E_ARG, "beta must be positive");
  const bool grid = points == nullptr;
  if (grid && (!rvecs || nx < 1 || ny < 1 || nz < 1)) return fail(CMM_E_ARG, "bad grid");
  if (!grid && M < 0) return fail(CMM_E_ARG, "negative point count");
  EffInterpDev e;
  std::memset(&e, 0, sizeof(e));
  for (int t = 0; t < nt; ++t) {
    int rc = interp_tab(tables[t], tn + 3 * t, torigin + 3 * t, tspacing + 3 * t, tmode[t], e.tab[t]);
    if (rc) return rc;
  }
  e.nrot = nrot;
  e.ng = ng;
  for (int i = 0; i < ng; ++i) {
    if (gtype[i] < 0 || gtype[i] >= nt) return fail(CMM_E_ARG, "guest atom type out of range");
    e.gtype[i] = gtype[i];
  }
  e.beta = beta;
  e.nanval = nanval;
  hipStream_t s = (hipStream_t)stream;
  if (grid) {
    e.nx = nx;
    e.ny = ny;
    e.nz = nz;
    std::memcpy(e.rv, rvecs, 9 * sizeof(double));
    e.inv_n[0] = 1.0 / nx;
    e.inv_n[1] = 1.0 / ny;
    e.inv_n[2] = 1.0 / nz;
    dim3 g((nz + EP_BZ - 1) / EP_BZ, (ny + EP_BY - 1) / EP_BY, (nx + EP_BX - 1) / EP_BX);
    hipLaunchKernelGGL((k_effpot_interp<true>), g, dim3(EP_NT), 0, s, e, offsets, weights, nullptr, 0LL, V);
  } else {
    if (M == 0) return CMM_OK;
    dim3 g((unsigned)((M + EP_NT - 1) / EP_NT));
    hipLaunchKernelGGL((k_effpot_interp<false>), g, dim3(EP_NT), 0, s, e, offsets, weights, points, M, V);
  }
  LAUNCHCHK("k_effpot_interp");
  return CMM_OK;
}

int cmm_min_distance(int nx, int ny, int nz, const double* rvecs, const double* pos, int natom, double* out,
                     void* stream) {
  if (!rvecs || !out || (natom > 0 && !pos) || natom < 1) return fail(CMM_E_ARG, "null argument or no atoms");
  if (nx < 1 || ny < 1 || nz < 1) return fail(CMM_E_ARG, "bad sizes");
  MinDistDev d;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  std::memcpy(d.rv, rvecs, 9 * sizeof(double));
  // G = R^-T: rows of G are the reciprocal vectors (G R^T = I)
  const double* r = rvecs;
  const double det = r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) +
                     r[2] * (r[3] * r[7] - r[4] * r[6]);
  if (!(std::fabs(det) > 0.0)) return fail(CMM_E_ARG, "singular cell");
  // inverse of R (row-major), then transpose
  double inv[9];
  inv[0] = (r[4] * r[8] - r[5] * r[7]) / det;
  inv[1] = (r[2] * r[7] - r[1] * r[8]) / det;
  inv[2] = (r[1] * r[5] - r[2] * r[4]) / det;
  inv[3] = (r[5] * r[6] - r[3] * r[8]) / det;
  inv[4] = (r[0] * r[8] - r[2] * r[6]) / det;
  inv[5] = (r[2] * r[3] - r[0] * r[5]) / det;
  inv[6] = (r[3] * r[7] - r[4] * r[6]) / det;
  inv[7] = (r[1] * r[6] - r[0] * r[7]) / det;
  inv[8] = (r[0] * r[4] - r[1] * r[3]) / det;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) d.gv[3 * a + b] = inv[3 * b + a];
  d.inv_n[0] = 1.0 / nx;
  d.inv_n[1] = 1.0 / ny;
  d.inv_n[2] = 1.0 / nz;
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_min_dist, dim3((unsigned)((n + MD_NT - 1) / MD_NT)), dim3(MD_NT), 0, (hipStream_t)stream, pos,
                     natom, d, out);
  LAUNCHCHK("k_min_dist");
  return CMM_OK;
}

int cmm_lj_grad_workspace(int nx, int ny, int nz, int ntypes, long long* ndoubles) {
  if (!ndoubles || nx < 1 || ny < 1 || nz < 1 || ntypes < 1) return fail(CMM_E_ARG, "bad arguments");
  const dim3 g = lj_grid_dims(nx, ny, nz);
  *ndoubles = 2LL * ntypes * g.x * g.y * g.z;
  return CMM_OK;
}

int cmm_lj_grid_grad(int nx, int ny, int nz, const double* rvecs, const double* atoms, const int* types, int natom,
                     int ntypes, double sigma_f, double epsilon_f, double cutoff, int images, double limit,
                     const double* w, int nw, double* work, double* grad, void* stream) {
  if (!w || !work || !grad || (natom > 0 && (!atoms || !types)) || natom < 0 || ntypes < 1 || nw < 1)
    return fail(CMM_E_ARG, "null argument");
  LJDev d;
  int rc = lj_setup(nx, ny, nz, rvecs, sigma_f, epsilon_f, cutoff, images, limit, d);
  if (rc) return rc;
  const dim3 grid = lj_grid_dims(nx, ny, nz);
  const long long nblocks = (long long)grid.x * grid.y * grid.z;
  hipLaunchKernelGGL(k_lj_bwd, grid, dim3(LJ_NT), 0, (hipStream_t)stream, atoms, types, natom, ntypes, d, w, nw,
                     work);
  LAUNCHCHK("k_lj_bwd");
  hipLaunchKernelGGL(k_lj_bwd_reduce, dim3(2 * ntypes), dim3(LJ_NT), 0, (hipStream_t)stream, work, nblocks, grad);
  LAUNCHCHK("k_lj_bwd_reduce");
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free entry point.  Each GPU the
// library runs on holds one copy of this array until the process ends; its
// device address is looked up once per device with that device current
// (hipGetSymbolAddress) and used from then on -- the array's name is a
// host-side stand-in in host code, never passed to a kernel.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[CMM_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

int scratch_ptr(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CMM_E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CMM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return CMM_OK;
}
}  // namespace

int cmm_lj_grid_grad_noalloc(int nx, int ny, int nz, const double* rvecs, const double* atoms, const int* types,
                             int natom, int ntypes, double sigma_f, double epsilon_f, double cutoff, int images,
                             double limit, const double* w, int nw, double* grad, void* stream) {
  long long need = 0;
  int rc = cmm_lj_grad_workspace(nx, ny, nz, ntypes, &need);
  if (rc) return rc;
  if (need > (long long)CMM_NOALLOC_SCRATCH_DOUBLES)
    return fail(CMM_E_ARG, "cmm_lj_grid_grad_noalloc: the grid and type count need " + std::to_string(need) +
                               " doubles of scratch, above the fixed " +
                               std::to_string((long long)CMM_NOALLOC_SCRATCH_DOUBLES) +
                               "; use cmm_lj_grid_grad with a workspace");
  double* work = nullptr;
  rc = scratch_ptr(&work);
  if (rc) return rc;
  return cmm_lj_grid_grad(nx, ny, nz, rvecs, atoms, types, natom, ntypes, sigma_f, epsilon_f, cutoff, images, limit,
                          w, nw, work, grad, stream);
}

int cmm_batch_reset(cmm_plan* p, const double* mu, double beta, double lam3, double dr, double tol, int nsteps) {
  if (!p || !mu) return fail(CMM_E_ARG, "null argument");
  const int nb = p->nb;
  if (!p->b_active) {
    hipError_t e = hipSuccess;
    e = (e == hipSuccess) ? hipMalloc(&p->b_active, nb * sizeof(int)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_iters, nb * sizeof(int)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_status, nb * sizeof(int)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_final, nb * sizeof(int)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_omega_prev, nb * sizeof(double)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_omega, nb * sizeof(double)) : e;
    e = (e == hipSuccess) ? hipMalloc(&p->b_mu, nb * sizeof(double)) : e;
    if (e != hipSuccess) return fail(CMM_E_HIP, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  std::vector<int> ones(nb, 1), zeros(nb, 0);
  HIPCHK(hipMemcpyAsync(p->b_active, ones.data(), nb * sizeof(int), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->b_iters, zeros.data(), nb * sizeof(int), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->b_status, zeros.data(), nb * sizeof(int), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->b_final, zeros.data(), nb * sizeof(int), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->b_mu, mu, nb * sizeof(double), hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  BatchDev& b = p->bd;
  b.active = p->b_active;
  b.iters = p->b_iters;
  b.status = p->b_status;
  b.final_slot = p->b_final;
  b.omega_prev = p->b_omega_prev;
  b.omega = p->b_omega;
  b.mu = p->b_mu;
  b.beta = beta;
  b.kT = 1.0 / beta;
  b.lam3 = lam3;
  b.dr = dr;
  b.tol = tol;
  b.nsteps = nsteps;
  b.nb = nb;
  b.mfa = p->fp.kmfa != nullptr;
  p->active = p->b_active;
  return CMM_OK;
}

int cmm_batch_check(cmm_plan* p, int new_slot, int init) {
  if (!p || !p->b_active) return fail(CMM_E_ARG, "batch not initialised");
  hipLaunchKernelGGL(k_batch_check, dim3((p->nb + 63) / 64), dim3(64), 0, p->stream, p->bd, red_of(p), new_slot, init);
  LAUNCHCHK("k_batch_check");
  return CMM_OK;
}

int cmm_batch_state(cmm_plan* p, int* active, int* iters, int* status, int* final_slot, double* omega) {
  if (!p || !p->b_active) return fail(CMM_E_ARG, "batch not initialised");
  const int nb = p->nb;
  if (active) HIPCHK(hipMemcpyAsync(active, p->b_active, nb * sizeof(int), hipMemcpyDeviceToHost, p->stream));
  if (iters) HIPCHK(hipMemcpyAsync(iters, p->b_iters, nb * sizeof(int), hipMemcpyDeviceToHost, p->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, p->b_status, nb * sizeof(int), hipMemcpyDeviceToHost, p->stream));
  if (final_slot) HIPCHK(hipMemcpyAsync(final_slot, p->b_final, nb * sizeof(int), hipMemcpyDeviceToHost, p->stream));
  if (omega) HIPCHK(hipMemcpyAsync(omega, p->b_omega, nb * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CMM_OK;
}

int cmm_batch_release(cmm_plan* p) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (p->active == p->b_active) p->active = nullptr;
  return CMM_OK;
}

int cmm_graph_begin(cmm_plan* p) {
  if (!p) return fail(CMM_E_ARG, "null plan");
  if (p->stream == 0) return fail(CMM_E_ARG, "graph capture needs a non-default stream");
  HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  return CMM_OK;
}

int cmm_graph_end(cmm_plan* p, void** exec_out) {
  if (!p || !exec_out) return fail(CMM_E_ARG, "null argument");
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(p->stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(CMM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  *exec_out = (void*)ex;
  return CMM_OK;
}

int cmm_graph_launch(cmm_plan* p, void* exec) {
  if (!p || !exec) return fail(CMM_E_ARG, "null argument");
  HIPCHK(hipGraphLaunch((hipGraphExec_t)exec, p->stream));
  return CMM_OK;
}

int cmm_graph_destroy(void* exec) {
  if (exec) HIPCHK(hipGraphExecDestroy((hipGraphExec_t)exec));
  return CMM_OK;
}
