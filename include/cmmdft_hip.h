/*
 * C ABI of libcmmdft_hip.so: the MI355X (gfx950) implementation of the cDFT
 * hot path of CmmDFT (FMT/WDA weighted densities and their back-convolution,
 * the Picard/Anderson Euler-Lagrange update, the LJ external-potential grid).
 *
 * Conventions
 *  - All arrays are device pointers (e.g. torch.Tensor.data_ptr()), FP64,
 *    C-contiguous.  A real field is double[nbatch][nx][ny][nz]; a half
 *    spectrum is double2 (interleaved re, im) [nbatch][nx][ny][nz/2+1].
 *  - Lengths in bohr, energies in hartree (atomic units).
 *  - Every call returns a status: CMM_OK, or an error whose text is given by
 *    cmm_last_error() (thread-local).  CMM_E_ARG maps to ValueError,
 *    CMM_E_NONFINITE to FloatingPointError, CMM_E_HIP to RuntimeError.
 *  - Work is enqueued on the plan's stream (cmm_plan_set_stream); only
 *    cmm_get_scalars and cmm_synchronize wait for it.  The library never
 *    frees caller memory.  A plan is not thread-safe: one per stream.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *  cmm_plan_create       Grid.__init__                 cmmdft/system.py:231-295
 *  cmm_fft_r2c/_c2r      Grid.fft / Grid.ifft          cmmdft/system.py:310-332
 *  cmm_set_functional    HardSphereFunctional.__init__/set_temperature,
 *                        WDAVFunctional.set_temperature
 *                                                      cmmdft/functionals.py:384-487, 1133-1155
 *  cmm_derive            HardSphereFunctional.derive + WDAVFunctional.derive
 *                        (+ get_phi / EOS f_ex for value())
 *                                                      cmmdft/functionals.py:489-618, 620-767, 1157-1184
 *  cmm_weighted_densities HardSphereFunctional._get_density_functions,
 *                        WDAVFunctional._get_weighted_density
 *                                                      cmmdft/functionals.py:489-532, 1157-1158
 *  cmm_finish_c1         Solver._get_C1                cmmdft/solver.py:137-144
 *  cmm_finish_picard     Picard.update_rho_static      cmmdft/solver.py:392-401 (+ get_new_rho :128-129)
 *  cmm_finish_anderson   Anderson.update_rho (residual + history)  cmmdft/solver.py:527-535, 546-557
 *  cmm_anderson_coef     Anderson.update_rho_Anderson (SLSQP simplex LSQ) and
 *                        _get_damping_coefficient      cmmdft/solver.py:537-544, 581-597
 *  cmm_anderson_mix      Anderson.update_rho_Anderson (mixing, clip)  cmmdft/solver.py:594-604
 *  cmm_trial             Picard.update_rho_hybrid trial densities     cmmdft/solver.py:414-441, 451-452
 *  cmm_seed              Grid.fft of a density + Solver._get_Omega ideal/N terms
 *                                                      cmmdft/solver.py:103-119
 *  cmm_lj_grid/_points   get_external_potential / lennard_jones
 *                                                      cmmdft/extpot_calculator.py:156-222
 *  cmm_effpot            generate_effective_potential / effective_potential_vectorized
 *                        (orientational average for non-spherical guests)
 *                                                      cmmdft/extpot_calculator.py:572-646
 *  cmm_lj_derivs         get_external_potential_derivatives  cmmdft/extpot_calculator.py:264-346
 *  cmm_interp_table/_points, cmm_effpot_interp
 *                        Interpolator (tricubic, trilinear, estimated
 *                        derivatives) and the interpolated effective potential
 *                                                      cmmdft/extpot_calculator.py:391-548, 572-646
 *  cmm_lj_grid_grad      lennard_jones(derivative=True) (the reference's
 *                        radial derivatives; here with respect to the
 *                        host-type sigma/epsilon)     cmmdft/extpot_calculator.py:156-183
 *  cmm_mfa_lj_potential  MFAFunctional.generate_potential_lj (on the grid
 *                        or, folded, on its supercell)  cmmdft/functionals.py:878-910
 *  cmm_fold              the supercell sampling kpotential[::r0,::r1,::r2]
 *                        of MFAFunctional.derive, done in real space
 *                                                      cmmdft/functionals.py:917-923
 *  cmm_mfa_table         MFAFunctional kpotential (fft x Lanczos, or fft of a
 *                        loaded potential)             cmmdft/functionals.py:809-812, 910
 *  cmm_plan_set_mfa      MFAFunctional.derive / value  cmmdft/functionals.py:912-939
 *  cmm_min_distance      Program.calc_distance          cmmdft/program.py:145-177
 *  cmm_field_reduce      Calculator integrals: loading, Henry coefficient,
 *                        helium void fraction, ideal-gas free energy
 *                                                      cmmdft/calculator.py:60-88, 152-184, 387-397
 *  cmm_hessian_apply, cmm_adjoint_*
 *                        no reference counterpart: the reference has no
 *                        derivatives of the converged density; these give
 *                        dN/dmu and dN/d(host LJ parameters) by the implicit
 *                        function theorem at the fixed point of
 *                        Solver.get_new_rho                cmmdft/solver.py:128-129
 */
#ifndef CMMDFT_HIP_H
#define CMMDFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CMM_OK 0
#define CMM_E_ARG 1
#define CMM_E_HIP 2
#define CMM_E_NONFINITE 3
#define CMM_E_UNSUPPORTED 4

#define CMM_NSCAL 32      /* doubles of reduction results per instance */
#define CMM_MMAX_HIST 8   /* longest Anderson history */
#define CMM_MAX_RADIX_STAGES 8

typedef struct cmm_plan cmm_plan;

/* Functional definition.  hs_base: -1 none, 0 Rosenfeld FMT, 1 MFMT,
 * 2 White-Bear II; hs_antisym: the 'a' variant; hs_tensor: the 't' variant
 * (twelve weighted-density fields; nz <= 170).  WDA EOS terms: kind 0 mean
 * field (par[0] = a), 1 Carnahan-Starling (par = {eta, kT}), 2 MBWR
 * (par = {a1..a8, b1..b6, sigma^3, epsilon}); factor multiplies the term. */
typedef struct {
  int hs_base;
  int hs_antisym;
  int hs_tensor;
  int use_wda;
  int n_eos;
  int eos_kind[4];
  double eos_factor[4];
  double eos_par[4][16];
  double R;    /* hard-sphere radius */
  double D;    /* WDA weight diameter */
  double kT;   /* 1 / beta */
  double n3_series;  /* MFMT / White-Bear II: phi2, phi3 and their n3
                        derivatives by their second-order series where
                        n3 <= n3_series; 1e-8 is the reference's
                        (functionals.py:726-767), at most 1e-3 */
  int wda_pointwise; /* 0: a WDA whose EOS terms are all mean field (f_ex =
                        A rho_bar^2, linear derivative) is applied in k space
                        by the x passes, its weighted density never formed;
                        1: the pointwise route of every other WDA */
} cmm_functional;

int cmm_version(void);
const char* cmm_last_error(void);

/* Grid plan for nbatch instances of an nx*ny*nz grid in the cell whose rows
 * are rvecs[0..2], rvecs[3..5], rvecs[6..8].  Owns the FFT tables, the
 * k-space tables and the spectral workspace (7 half spectra per instance). */
int cmm_plan_create(int nx, int ny, int nz, const double* rvecs, int nbatch, cmm_plan** out);
int cmm_plan_destroy(cmm_plan* plan);
int cmm_plan_set_stream(cmm_plan* plan, void* stream);
/* tile shapes: x-pass columns, y-pass columns, z-pass pencils (simple),
 * z-pass pencils (6-field functional pass); 0 keeps the current value */
int cmm_plan_set_tiles(cmm_plan* plan, int tb_x, int tb_y, int p_z, int p_phi);
/* FFT radix sequence of one axis (0 x, 1 y, 2 z): radices from {2,3,4,5,8}
 * whose product is the axis length, at most CMM_MAX_RADIX_STAGES stages,
 * applied in order (the autotuner's search space together with the tiles;
 * the default is 5s, 3s, 8s, 4s, 2s) */
int cmm_plan_set_radix(cmm_plan* plan, int axis, int nstages, const int* radices);
int cmm_plan_get_radix(cmm_plan* plan, int axis, int* nstages, int* radices);
/* info[0..11] = nx, ny, nz, nbatch, tb_x, tb_y, p_z, p_phi, workspace bytes, nzh,
 * nf (weighted-density fields of the functional set: 6 or 12), 1 when the
 * functional's WDA is applied in k space (see wda_pointwise) */
int cmm_plan_get_info(cmm_plan* plan, long long* info, int n);
/* device int[nbatch] mask of instances to process (NULL = all) */
int cmm_plan_set_active(cmm_plan* plan, const int* dev_mask);
/* sets the hard-sphere / WDA functional; clears the mean-field table */
int cmm_set_functional(cmm_plan* plan, const cmm_functional* f);
/* Deterministic reduction of a device field x[n] (points with mask[i] = 0
 * skipped when mask is not NULL) into device out[0]: op 0 sum x, 1 sum
 * exp(-beta x), 2 sum min(max(exp(-beta x), 0), 1), 3 min x, 4 max x, 5 sum
 * x^2, 6 sum c (ln(lam3 c) - 1) with c = max(x, 1e-20), 7 the same with
 * c = 1e-30 where x <= 1e-8 (the floor of FreeEnergy.track,
 * cmmdft/functionals.py:151-155).  work: device scratch of at least 1024
 * doubles. */
int cmm_field_reduce(const double* x, const unsigned char* mask, long long n, int op, double beta, double lam3,
                     double* work, double* out, void* stream);
/* Mean-field (MFA) part: kd is a caller-owned complex half spectrum
 * double2[nx][ny][nz/2+1] made by cmm_mfa_table (NULL switches the part
 * off); it must stay valid while set.  The derivative chain then adds
 * c_MFA = IFFT(K rho^) to c_ex, and the forward chain reduces
 * (1/2) sum_x rho c_MFA into scalar slot 28 (times dr: the MFA energy).
 * Call after cmm_set_functional. */
int cmm_plan_set_mfa(cmm_plan* plan, const double* kd);
/* LJ pair potential of the mean-field part on the grid of ns[0] x ns[1] x
 * ns[2] points of the cell rvecs (rows are the cell vectors; points
 * frac . rvecs with frac_a = i_a / ns_a - 1/2), folded onto the grid of
 * ns / reps points (reps = {1,1,1}: the plain potential):
 * U = 4 eps ((sigma/r)^12 - (sigma/r)^6) for r > rmin, limit otherwise;
 * with use_cutoff, U - U(cutoff) for r <= cutoff and 0 beyond.
 * out: device double[ns/reps], on stream. */
int cmm_mfa_lj_potential(const int ns[3], const double rvecs[9], const int reps[3], double sigma, double epsilon,
                         double rmin, double limit, double cutoff, int use_cutoff, double* out, void* stream);
/* out[i] = sum over image offsets j of in[i + j n], n = ns / reps (device
 * arrays, in: ns points, out: ns / reps points) */
int cmm_fold(const int ns[3], const int reps[3], const double* in, double* out, void* stream);
/* Mean-field kernel table from uhat = cmm_fft_r2c of the folded potential
 * (one instance): kd = scale uhat L P_h with L the Lanczos factor (lanczos =
 * 1; the reference applies it to generated, not to loaded potentials), P_h
 * the Hermitian part of the reference's phase exp(i pi scalprod) of the
 * supercell with vectors rvecs_sc (= rvecs with column a times reps[a]) at
 * the wave vectors reps . m, and scale = dr / npts of the grid. */
int cmm_mfa_table(cmm_plan* plan, const double* uhat, const int reps[3], const double rvecs_sc[9], int lanczos,
                  double scale, double* kd);
/* per-instance fugacity and mixing parameter (host arrays of nbatch) */
int cmm_set_instances(cmm_plan* plan, const double* fug, const double* alpha);

/* unnormalised 3-D transforms over the batch (X: interleaved complex) */
int cmm_fft_r2c(cmm_plan* plan, const double* x, double* X);
int cmm_fft_c2r(cmm_plan* plan, const double* X, double* x);

/* R2C of a density into the plan workspace; reduces N, ideal-gas and
 * external sums of rho (scalars S_PN..).  vext may be NULL; vext_batched
 * says whether vext has one field per instance. */
int cmm_seed(cmm_plan* plan, const double* rho, const double* vext, int vext_batched, double lam3);
/* functional chain on the seeded density: weighted densities, pointwise
 * Phi/WDA derivatives (reduces sum Phi, sum f_ex, max n3), back-convolution;
 * leaves the x,y-transformed derivative sum in the workspace */
int cmm_derive(cmm_plan* plan);
/* value-only chain (no back-convolution): reduces sum Phi, sum f_ex, max n3 */
int cmm_value(cmm_plan* plan);
/* weighted densities n2, n3, nv2x, nv2y, nv2z, rho_bar (then, for the tensor
 * variant, n_t xx xy xz yy yz zz) of the seeded density into
 * out[nbatch][nf][npts], nf = 6 or 12 (n3 unclipped) */
int cmm_weighted_densities(cmm_plan* plan, double* out);

/* Pointwise free-energy densities (per instance, on the plan's stream):
 * phi[npts] = kT Phi of the weighted densities nfld[nbatch][nf][npts] (as
 * cmm_weighted_densities writes them), fex/dfex[npts] = f_ex(rbar),
 * f_ex'(rbar) of the WDA part at the weighted density rbar[nbatch][npts].
 * NULL inputs or outputs are skipped.
 * Replaces HardSphereFunctional.value(local=True) (get_phi) and the EOS
 * terms of WDAVFunctional.derive / value with a given weighted density
 *                                                      cmmdft/functionals.py:611-643, 1160-1184 */
int cmm_local_density(cmm_plan* plan, const double* nfld, const double* rbar, double* phi, double* fex,
                      double* dfex);
/* The reference's k-space hard-sphere weights on the full nx*ny*nz grid of
 * the plan: out[0][..] = sinc(kR) sigma_Lanczos, out[1][..] =
 * 3 (sin kR - kR cos kR)/(kR)^3 sigma_Lanczos (both sigma_Lanczos where
 * kR <= 1e-8); w0, w1 = R w0, w2 = 4 pi R^2 w0, w3 = 4 pi R^3/3 out[1],
 * w_v2 = -i k w3, w_v1 = w_v2 / (4 pi R) follow.
 * Replaces HardSphereFunctional._init_weight_functions (the attributes
 * scalar_weight_functions, vector_weight_functions)
 *                                                      cmmdft/functionals.py:421-455 */
int cmm_hs_weight_table(cmm_plan* plan, double R, double* out);

/* launch one pass of the functional chain on the current workspace
 * (measurement aid, kernel only, no reduction of its partials): 0 functional
 * z pass, 1 forward x pass with weights, 2 back-convolution x pass, 3 y pass
 * over the 6 fields, 4 y pass over rho */
int cmm_launch_pass(cmm_plan* plan, int which);

/* final passes after cmm_derive; rho is the density the derivative belongs to.
 * Scalars [3..5] = N, ideal-gas sum sum rho_c (ln(lam3 rho_c) - 1) and
 * external sum sum rho V of rho; cmm_finish_anderson forms N only (its ideal
 * and external slots read 0: the producer pass gives them for the same
 * density) */
int cmm_finish_c1(cmm_plan* plan, const double* rho, const double* vext, int vext_batched, double* c1,
                  double lam3);
int cmm_finish_picard(cmm_plan* plan, const double* rho, const double* vext, int vext_batched, double* rho_out,
                      double* c1_out, double beta, double lam3, double zero_thr);
int cmm_finish_anderson(cmm_plan* plan, const double* rho, const double* vext, int vext_batched, double* rhist,
                        int m, int slot_new, double* c1_out, double beta, double lam3);
/* convergence measures of rho (writes no field): scalars [3..9] = N, ideal
 * sum, external sum, non-finite count of G - rho, sum (G - rho)^2, and over
 * |rho| > 1e-8 the sums of (|rho| beta dOmega / (a_tol + r_tol |rho|))^2 and
 * (|rho| beta dOmega)^2 with beta dOmega = ln(lam3 rho_c) + beta C1 - beta mu
 * (rho_c = rho, or 1e-30 below 1e-12).
 * Replaces the RES and DER criteria of Solver._check_convergence and its
 * history columns                              cmmdft/solver.py:131-135, 186-233 */
int cmm_finish_check(cmm_plan* plan, const double* rho, const double* vext, int vext_batched, double beta,
                     double lam3, double a_tol, double r_tol);
/* out[i] = beta exp(-beta c1[i]) fug (device arrays of n), on stream.
 * Replaces Solver.get_new_rho                  cmmdft/solver.py:128-129 */
int cmm_boltzmann(const double* c1, long long n, double beta, double fug, double* out, void* stream);

/* Anderson coefficients on the device for the mk newest residual slots
 * order[0..mk) (oldest first); reads this step's Gram row from the scalars. */
int cmm_anderson_init(cmm_plan* plan, const double* damping, const double* cf);
/* store this step's Gram row (scalars of the last cmm_finish_anderson) into
 * the device Gram ring without mixing: a hybrid (Picard) step of the
 * Anderson solver still saves its residual into the history, as
 * Anderson._save_previous_rhos does on every step  cmmdft/solver.py:527-535, 550-552 */
int cmm_anderson_store_row(cmm_plan* plan, int m, int slot_new);
int cmm_anderson_coef(cmm_plan* plan, int m, int mk, const int* order, int slot_new, int slot_prev, int adaptive,
                      double fac_up, double fac_down, double dmin, double dmax);
int cmm_anderson_mix(cmm_plan* plan, const double* rhohist, const int* order_rho, const double* rhist,
                     const int* order, int mk, const double* rho_cur, double* rho_out, const double* vext,
                     int vext_batched, double lam3);
/* host copies of the per-instance damping and the last coefficients
 * (alpha[nbatch][CMM_MMAX_HIST]); the damping the last coefficients used is
 * also scalar [30] */
int cmm_anderson_state(cmm_plan* plan, double* damping, double* alpha);
/* set the per-instance damping (host array of nbatch): when the Anderson
 * density of a step turns out to have a non-finite Boltzmann factor, that
 * step is taken again as a hybrid Picard step, keeping its damping update
 * and undoing the later one                     cmmdft/solver.py:537-544, 561-574 */
int cmm_anderson_set_damping(cmm_plan* plan, const double* damping);
/* rho_out = rho + alpha_inst * r_slot (clip to 1e-30 below 1e-12 if clip) */
int cmm_trial(cmm_plan* plan, const double* rho, const double* rhist, int slot, double* rho_out,
              const double* vext, int vext_batched, double lam3, int clip);

/* Batched solves (no reference counterpart: the reference loops over state
 * points serially, examples/methane_ZIF8.ipynb).  cmm_batch_reset uploads the
 * per-instance chemical potentials, marks every instance active and makes the
 * batch's device mask the plan's active mask; cmm_batch_check tests the
 * density just produced and derived (RIUE < tol, the grand-potential
 * safeguard of Anderson.update_rho, solver.py:567-574, non-finite values,
 * nsteps) and deactivates finished instances; status codes: 0 running,
 * 1 converged, 2 step limit, 3 needs the hybrid fallback, 4 non-finite. */
int cmm_batch_reset(cmm_plan* plan, const double* mu, double beta, double lam3, double dr, double tol, int nsteps);
int cmm_batch_check(cmm_plan* plan, int new_slot, int init);
int cmm_batch_state(cmm_plan* plan, int* active, int* iters, int* status, int* final_slot, double* omega);
int cmm_batch_release(cmm_plan* plan);
/* capture the plan stream's launches into a HIP graph (kernels only) */
int cmm_graph_begin(cmm_plan* plan);
int cmm_graph_end(cmm_plan* plan, void** exec_out);
int cmm_graph_launch(cmm_plan* plan, void* exec);
int cmm_graph_destroy(void* exec);

/* Linear response of a converged density (adjoint solve).
 * cmm_hessian_apply: M v for the half spectrum of v in the plan workspace
 * (left there by cmm_seed(v)), M = d c_ex / d rho at the density whose
 * weighted densities nfld[nbatch][6][npts] were written by
 * cmm_weighted_densities; cmm_finish_c1 with vext = NULL then writes M v.
 * The adjoint solve J lambda = dr (J = diag(kT/rho) + M) runs as conjugate
 * gradients on (I + S M S) y = S dr, S = sqrt(beta rho), lambda = S y, on
 * caller vectors y, r, p, q [nbatch][npts]:
 *   cmm_adjoint_seed(init=1, parity 0); then per iteration k (parity k & 1):
 *   cmm_hessian_apply -> cmm_adjoint_final -> cmm_adjoint_update(parity) ->
 *   cmm_adjoint_seed(init=0, parity of k+1).
 * Scalars per instance: [24] p.q, [25 + parity] r.r, [27] sum lambda. */
int cmm_hessian_apply(cmm_plan* plan, const double* nfld);
int cmm_adjoint_seed(cmm_plan* plan, const double* rho, double beta, double dr, int init, int parity, double* y,
                     double* r, double* p);
int cmm_adjoint_final(cmm_plan* plan, const double* rho, double beta, const double* p, double* q);
int cmm_adjoint_update(cmm_plan* plan, int parity, double* y, double* r, const double* p, const double* q);
/* lambda = S y; sum lambda = dN/dmu into scalar [27] */
int cmm_adjoint_lambda(cmm_plan* plan, const double* rho, double beta, const double* y, double* lambda);

/* copy scalars [nbatch][CMM_NSCAL] to host (waits for the stream) */
int cmm_get_scalars(cmm_plan* plan, double* host_out);
int cmm_synchronize(cmm_plan* plan);

/* LJ external potential on the nx*ny*nz grid of the cell rvecs (points
 * frac . rvecs, frac in [-1/2, 1/2)).  atoms: natom x {x, y, z, sigma_i,
 * epsilon_i} (host parameters before mixing); guest sigma_f, epsilon_f;
 * Lorentz-Berthelot mixing on the device.  images = 0: minimum image with row
 * norms (the reference rule); 1: every periodic image within the cutoff
 * (orthorhombic cells).  limit > 0 clamps non-finite and larger values to
 * limit.  Writes V[nx*ny*nz]. */
int cmm_lj_grid(int nx, int ny, int nz, const double* rvecs, const double* atoms, int natom, double sigma_f,
                double epsilon_f, double cutoff, int images, double limit, double* V, void* stream);
/* the same sum at M explicit points (points: M x 3, device), V[M] */
int cmm_lj_points(const double* points, long long M, const double* rvecs, const double* atoms, int natom,
                  double sigma_f, double epsilon_f, double cutoff, int images, double limit, double* V, void* stream);

/* Orientationally averaged potential of a rigid guest of ng atoms (types
 * gtype[i] < nt, LJ parameters gsig/geps per type):
 *   V[p] = -(1/beta) log sum_r weights[r] exp(-beta sum_i V_gtype[i](p + offsets[r][i]))
 * with V_t the LJ field of the host atoms (x, y, z, sigma, epsilon rows) for
 * a guest atom of type t, as in cmm_lj_grid; a NaN result becomes nanval.
 * points == NULL: the nx x ny x nz grid of the cell; else the M points
 * (M x 3).  offsets: device [nrot][ng][3] (COM + R_r (x_i - COM)); weights:
 * device [nrot]; dmax >= max |offsets|.  ng <= 16, nt <= 4.
 * Replaces generate_effective_potential with get_external_potential
 * generators (effective_potential_vectorized)
 *                                 cmmdft/extpot_calculator.py:572-646, 186-222 */
int cmm_effpot(int nx, int ny, int nz, const double* points, long long M, const double* rvecs, const double* atoms,
               int natom, int ng, const int* gtype, int nt, const double* gsig, const double* geps,
               const double* offsets, const double* weights, int nrot, double dmax, double beta, double cutoff,
               int images, double nanval, double* V, void* stream);

/* Field and derivatives of the LJ sum for a guest atom (sigma_f, epsilon_f),
 * on the nx x ny x nz grid of the cell (points == NULL) or at M points:
 * out[8][npts] = V, dV/dx h0, dV/dy h1, dV/dz h2, d2V/dxdy h0 h1,
 * d2V/dxdz h0 h2, d2V/dydz h1 h2, d3V/dxdydz h0 h1 h2 (h = spacings), V and
 * the first derivatives clipped to +-vmax, the others zeroed where V > vmax.
 * Replaces get_external_potential_derivatives
 *                                 cmmdft/extpot_calculator.py:264-346 */
int cmm_lj_derivs(int nx, int ny, int nz, const double* points, long long M, const double* rvecs,
                  const double* atoms, int natom, double sigma_f, double epsilon_f, double cutoff, int images,
                  const double* spacings, double vmax, double* out, void* stream);

/* Interpolation table [nx][ny][nz][8] (channel fastest) from grid values
 * [8][nx][ny][nz] (estimate = 0) or with the seven derivative channels
 * estimated from values[0] by periodic central differences (estimate = 1).
 * Replaces Interpolator.__init__ / estimate_all_derivatives
 *                                 cmmdft/extpot_calculator.py:391-470 */
int cmm_interp_table(const double* values, int nx, int ny, int nz, int estimate, double* table, void* stream);

/* Tricubic (mode 0) or trilinear (mode 1) interpolation of a table at M
 * points: fractional index (x - origin) / spacing per axis, periodic wrap.
 * n, origin, spacing: host arrays of 3.
 * Replaces Interpolator.tricubic_interpolation / trilinear_interpolation
 *                                 cmmdft/extpot_calculator.py:472-548 */
int cmm_interp_points(const double* table, const int* n, const double* origin, const double* spacing, int mode,
                      const double* points, long long M, double* out, void* stream);

/* cmm_effpot with the site potentials interpolated from one table per guest
 * atom type (tables: host array of nt device pointers; tn, torigin,
 * tspacing: host arrays nt x 3; tmode: host array nt).  rvecs is read for
 * the grid form only.
 * Replaces generate_effective_potential with interpolator generators
 *                                 cmmdft/extpot_calculator.py:572-646 */
int cmm_effpot_interp(int nx, int ny, int nz, const double* points, long long M, const double* rvecs, int ng,
                      const int* gtype, int nt, const double* const* tables, const int* tn, const double* torigin,
                      const double* tspacing, const int* tmode, const double* offsets, const double* weights,
                      int nrot, double beta, double nanval, double* V, void* stream);

/* Distance from each point of the nx x ny x nz grid of the cell to the
 * nearest of the natom host atoms (device pos[natom][3]), minimum image in
 * fractional coordinates.  Replaces Program.calc_distance
 *                                 cmmdft/program.py:145-177 */
int cmm_min_distance(int nx, int ny, int nz, const double* rvecs, const double* pos, int natom, double* out,
                     void* stream);

/* Backward of cmm_lj_grid: grad[2t] = sum_i w_i dV_i/dsigma_t,
 * grad[2t+1] = sum_i w_i dV_i/depsilon_t for host atom types t < ntypes
 * (types: device int[natom]), with w_i the sum of the nw weight fields
 * w[nw][nx*ny*nz]; clamped points contribute nothing.  work is a device
 * buffer of cmm_lj_grad_workspace() doubles; grad is a device
 * double[2 ntypes]. */
int cmm_lj_grad_workspace(int nx, int ny, int nz, int ntypes, long long* ndoubles);
int cmm_lj_grid_grad(int nx, int ny, int nz, const double* rvecs, const double* atoms, const int* types, int natom,
                     int ntypes, double sigma_f, double epsilon_f, double cutoff, int images, double limit,
                     const double* w, int nw, double* work, double* grad, void* stream);

/* cmm_lj_grid_grad without the workspace argument and without any device
 * allocation: its per-block partials go to a fixed device array of
 * CMM_NOALLOC_SCRATCH_DOUBLES doubles (4 MiB) that the library holds on each
 * GPU it runs on until the process ends.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Grids
 * and type counts whose workspace (cmm_lj_grad_workspace) exceeds the array
 * are rejected with CMM_E_ARG.  Calls that overlap on one GPU share the
 * array and race; the caller still owns w and grad. */
#define CMM_NOALLOC_SCRATCH_DOUBLES (1 << 19)
int cmm_lj_grid_grad_noalloc(int nx, int ny, int nz, const double* rvecs, const double* atoms, const int* types,
                             int natom, int ntypes, double sigma_f, double epsilon_f, double cutoff, int images,
                             double limit, const double* w, int nw, double* grad, void* stream);

#ifdef __cplusplus
}
#endif
#endif
