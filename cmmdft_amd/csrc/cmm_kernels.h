// Device-side parameter blocks shared by the kernels and their launchers.
#pragma once

#include "cmm_common.h"

#define CMM_MMAX 8   // longest Anderson history supported on the device

namespace cmm {

enum { MODE_C1 = 0, MODE_PICARD = 1, MODE_ANDRES = 2, MODE_CHECK = 3 };
enum { MODE_MIX = 0, MODE_TRIAL = 1, MODE_SEED = 2 };

// Per-instance reduction results (doubles, on the device), NSCAL per instance.
constexpr int NSCAL = 32;
constexpr int S_PHI = 0;    // sum Phi, sum f_ex(rho_bar), max n3            (functional pass)
constexpr int S_NCUR = 3;   // N, ideal sum, ext sum of rho, then per mode:  (final pass)
                            //   picard: N', sum|d|, max|d|, nonfinite
                            //   anderson: nonfinite, Gram row[CMM_MMAX]
                            //     (and 0 for the ideal and ext sums)
                            //   check: nonfinite, |G - rho|^2, DER sum, DER sum
                            //          without the tolerance weights
constexpr int S_GRAM = S_NCUR + 4;
constexpr int S_FMFA = 28;  // (1/2) sum rho c_MFA (mean-field energy / dr)   (x weights pass)
constexpr int S_ADAMP = 30; // the damping the last Anderson coefficients used  (k_anderson_coef)
constexpr int S_PN = 16;    // N', sum|d|, max|d|, nonfinite, ideal', ext'  (density producers)

// A z-pass tile: P pencils (B = P/2 complex sequences per field), S = nf*B
// sequences in all, element row stride ES (odd, >= S) in LDS.
struct ZTile {
  int P, B, S, ES;
  FastDiv fn;     // divide by nz
  FastDiv fnzh;   // divide by nz/2 + 1
  FastDiv fS;     // divide by S
};

struct GridDev {
  int nx, ny, nz, nzh;
  double inv_npts;
  Fft1D fx, fy, fz;
  KAxis kx, ky, kz;
};

// Functional: FMT hard spheres (base 0 Rosenfeld, 1 modified, 2 White-Bear
// II; anti = the 'a' variant) and WDA with a sum of EOS terms (kind 0 mean
// field, 1 Carnahan-Starling, 2 MBWR).
struct FuncDev {
  int use_hs, use_wda, base, anti, neos;
  int tensor;       // the 't' variant: six tensor weighted densities (fields 6..11)
  int eos_kind[4];
  double eos_fac[4];
  double eos_par[4][16];
  double R, D, kT;
  double n3_series;  // MFMT / White-Bear II: series forms of phi2, phi3 for n3 <= this
  double c_w2, c_w3, inv4piR2, inv4piR;
  double c_tor;     // 3 / R: tensor weight B = w2 - (3/R) w3
  const double2* kmfa;   // mean-field (MFA) kernel, half spectrum [nx][ny][nz/2+1]; null when off
  double wl2;       // linear WDA (every EOS term mean field, f_ex = A rho_bar^2): 2A, its terms applied
                    // in k space; 0 when the WDA takes the pointwise route
  const double* kwt;   // the weights w2, w3, w_D of every half-spectrum element [3][kwt_n], or null
  long long kwt_n;
};

struct InstDev {
  double fug;     // fugacity exp(beta mu) / beta / Lambda^3
  double alpha;   // mixing parameter (times correction factor) / trial step
};

struct RedDev {
  double* partials;   // per instance (stride pstride): [quantity][block]
  long long pstride;
  double* scal;
};

// Density-update passes.  Real fields are [inst][npts] with instance stride
// fstride; history rings are [slot][inst][npts] with slot stride hstride
// (residuals) / hstride2 (densities).
struct ElDev {
  const double* rho;
  const double* vext;
  int vext_batched;
  long long fstride;
  const InstDev* inst;
  double beta, lam3, zero_thr;
  double* c1out;
  double* rho_out;
  double2* rz;
  long long rzstride;
  double* rhist;
  long long hstride;
  int slot_new;
  int m;
  const double* rhohist;
  long long hstride2;
  const double* coef;
  int mk;
  int order[CMM_MMAX];
  int order_rho[CMM_MMAX];
  int clip_trial;
  double a_tol, r_tol;   // MODE_CHECK: tolerances of the derivative criterion
};

struct AndDev {
  double* gram;
  int m, mk;
  int order[CMM_MMAX];
  int slot_new, slot_prev;
  double* damping;
  const double* cf;
  int adaptive;
  double fac_up, fac_down, dmin, dmax;
  double* coef;
  double* alpha_out;
  int store_only;   // store this step's Gram row into the ring and stop
};

// Batched solve bookkeeping (one entry per instance).
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXSTEPS = 2, ST_FALLBACK = 3, ST_NONFINITE = 4 };

struct BatchDev {
  int* active;
  int* iters;
  int* status;
  int* final_slot;
  double* omega_prev;
  double* omega;
  const double* mu;
  double beta, kT, lam3, dr, tol;
  int nsteps;
  int nb;
  int mfa;        // the functional has the mean-field part (energy in S_FMFA)
};

}  // namespace cmm
