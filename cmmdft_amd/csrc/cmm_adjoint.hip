// Linear response of the converged Euler-Lagrange density (adjoint solve).
//
// At a solution rho of  R(rho) = kT ln(Lambda^3 rho) + c(rho) + V - mu = 0
// (c = the excess functional derivative), dR/drho = J = diag(kT/rho) + M with
// M = dc/drho, which is symmetric because the back-convolution applies the
// transposed weights of the forward convolution.  The adjoint equation
// J lambda = g (g = dN/drho = dr) is solved in the symmetric form
//     (I + S M S) y = S g,   S = diag(sqrt(beta rho)),   lambda = S y,
// by conjugate gradients.  Points with rho = 0 (frozen by the solver's zero
// floor) get S = 0 and drop out.  Then dN/dmu = sum lambda and
// dN/dtheta = -sum lambda dV/dtheta.
//
// M v is the functional chain of derive() with the pointwise stage replaced by
// its directional derivative: forward-mode dual numbers through fmt_point /
// wda_point at the weighted densities of rho (kept in HBM) along the weighted
// densities of v.
//
// CG scalars live in the per-instance scalar block: S_ADJ_PQ = p.q,
// S_ADJ_RR + parity = r.r of the current and the previous residual, S_ADJ_LAM
// = sum lambda.
#include "cmm_kernels.h"

namespace cmm {

constexpr int S_ADJ_PQ = 24;
constexpr int S_ADJ_RR = 25;    // two slots, alternating by iteration parity
constexpr int S_ADJ_LAM = 27;

struct AdjDev {
  const double* rho;    // converged density [inst][npts]
  const double* nfld;   // weighted densities of rho [inst][NF][npts] (n3 unclipped)
  long long fstride;    // npts
  double beta, dr;
  double* y;
  double* r;
  double* p;
  double* q;
  double* lam;
  int parity;
};

// z pass of the Hessian product: NF weighted densities of v (x,y inverse-
// transformed) -> C2R -> directional derivative of the pointwise derivative
// fields at the weighted densities of rho -> R2C -> F[0..NF) in place.
// NF = 5: a linear WDA, whose Hessian term the x passes apply in k space.
template <int NF>
__global__ void __launch_bounds__(NT) k_zhess(double2* __restrict__ F, long long Fs, long long Ffs,
                                             const double* __restrict__ nfld, long long npts, GridDev g, FuncDev fp,
                                             ZTile t, const int* __restrict__ active) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const int n = g.nz;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  const FastDiv fBn = make_fastdiv((unsigned)(t.B * g.nzh));
  constexpr bool VZ = (NF == 5 || NF == 6);
  z_load_tile<VZ, NF>(sm, F + inst * Fs, Ffs, p0, t, npen, n, g.nzh, fBn, g.kz.k, g.kz.nyq);
  __syncthreads();
  z_fft<1>(sm, t, g.fz, s_tw);
  const double* nb = nfld + (long long)inst * (NF == 5 ? 6 : NF) * npts;   // six fields when NF = 5
  const int tot = t.P * n;
  const FastDiv fP = make_fastdiv((unsigned)t.P);
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int e = (int)fdiv((unsigned)i, fP), pl = i - e * t.P;
    if (p0 + pl >= npen) continue;
    const long long idx = (p0 + pl) * n + e;
    double& r0 = z_val(sm, t, 0, pl, e);
    double& r1 = z_val(sm, t, 1, pl, e);
    double& r2 = z_val(sm, t, 2, pl, e);
    double& r3 = z_val(sm, t, 3, pl, e);
    double& r4 = z_val(sm, t, 4, pl, e);
    double g0 = 0.0, g1 = 0.0, gv0 = 0.0, gv1 = 0.0, gv2 = 0.0, g5 = 0.0;
    double gt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (fp.use_hs) {
      const Dual a0(nb[idx], r0), a1(nb[npts + idx], r1), a2(nb[2 * npts + idx], r2), a3(nb[3 * npts + idx], r3),
          a4(nb[4 * npts + idx], r4);
      PhiOut<Dual> o;
      if constexpr (NF == 12) {
        Dual tn[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) tn[c] = Dual(nb[(6 + c) * npts + idx], z_val(sm, t, 6 + c, pl, e));
        o = fmt_point<Dual, true>(fp, a0, a1, a2, a3, a4, tn);
#pragma unroll
        for (int c = 0; c < 6; ++c) gt[c] = o.gt[c].d;
      } else {
        o = fmt_point<Dual>(fp, a0, a1, a2, a3, a4);
      }
      g0 = o.g0.d;
      g1 = o.g1.d;
      gv0 = o.gv[0].d;
      gv1 = o.gv[1].d;
      gv2 = o.gv[2].d;
    }
    if constexpr (NF != 5) {
      if (fp.use_wda) {
        Dual fw, dfw;
        wda_point<Dual>(fp, Dual(nb[5 * npts + idx], z_val(sm, t, 5, pl, e)), fw, dfw);
        g5 = dfw.d;
      }
    }
    r0 = g0;
    r1 = g1;
    r2 = gv0;
    r3 = gv1;
    r4 = gv2;
    if constexpr (NF != 5) z_val(sm, t, 5, pl, e) = g5;
    if constexpr (NF == 12) {
#pragma unroll
      for (int c = 0; c < 6; ++c) z_val(sm, t, 6 + c, pl, e) = gt[c];
    }
  }
  __syncthreads();
  z_fft<-1>(sm, t, g.fz, s_tw);
  z_store_tile<VZ, NF>(sm, F + inst * Fs, Ffs, p0, t, npen, n, g.nzh, fBn, g.kz.k, g.kz.nyq);
}

// CG direction and its transform:
//   INIT: r = p = S g, y = 0 (reduces r.r into slot parity 0)
//   STEP: p = r + (rr_cur / rr_prev) p
// then v = S p, R2C along z -> rz.
enum { ADJ_INIT = 0, ADJ_STEP = 1 };

template <int MODE>
__global__ void __launch_bounds__(NT) k_adj_seed(GridDev g, AdjDev a, ZTile t, double2* __restrict__ rz,
                                                long long rzs, const int* __restrict__ active, RedDev red) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  __shared__ double sh[CMM_RED_SH];
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const int n = g.nz;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  const long long off = inst * a.fstride;
  const double* sc = red.scal + inst * NSCAL;
  double bcg = 0.0;
  if (MODE == ADJ_STEP) {
    const double rc = sc[S_ADJ_RR + a.parity], rp = sc[S_ADJ_RR + (a.parity ^ 1)];
    bcg = (rp != 0.0) ? rc / rp : 0.0;
  }
  double acc[1] = {0.0};
  const int tot = t.P * n;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
    const long long pp = p0 + pl;
    double v = 0.0;
    if (pp < npen) {
      const long long idx = off + pp * n + e;
      const double S = sqrt(a.beta * fmax(a.rho[idx], 0.0));
      double pv;
      if (MODE == ADJ_INIT) {
        pv = S * a.dr;
        a.r[idx] = pv;
        a.y[idx] = 0.0;
        acc[0] += pv * pv;
      } else {
        pv = a.r[idx] + bcg * a.p[idx];
      }
      a.p[idx] = pv;
      v = S * pv;
    }
    z_val(sm, t, 0, pl, e) = v;
  }
  __syncthreads();
  z_fft<-1>(sm, t, g.fz, s_tw);
  z_store_half(sm, rz + inst * rzs, p0, t, npen, n, g.nzh, 0);
  if (MODE == ADJ_INIT) {
    const int ops[1] = {0};
    block_reduce<1>(acc, ops, sh);
    grid_reduce_store<1>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
  }
}

// C2R of the back-convolved M v -> q = p + S (M v); reduces p.q.
__global__ void __launch_bounds__(NT) k_adj_final(const double2* __restrict__ F, long long Fs, GridDev g, AdjDev a,
                                                 ZTile t, const int* __restrict__ active, RedDev red) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  __shared__ double sh[CMM_RED_SH];
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const int n = g.nz;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  z_load_half(sm, F + inst * Fs, p0, t, npen, n, g.nzh, 0);
  __syncthreads();
  z_fft<1>(sm, t, g.fz, s_tw);
  const long long off = inst * a.fstride;
  double acc[1] = {0.0};
  const int tot = t.P * n;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
    const long long pp = p0 + pl;
    if (pp >= npen) continue;
    const long long idx = off + pp * n + e;
    const double S = sqrt(a.beta * fmax(a.rho[idx], 0.0));
    const double pv = a.p[idx];
    const double qv = pv + S * z_val(sm, t, 0, pl, e);
    a.q[idx] = qv;
    acc[0] += pv * qv;
  }
  const int ops[1] = {0};
  block_reduce<1>(acc, ops, sh);
  grid_reduce_store<1>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// y += alpha p, r -= alpha q with alpha = rr_cur / p.q; reduces the new r.r
// into the other parity slot.
__global__ void __launch_bounds__(NT) k_adj_update(AdjDev a, const int* __restrict__ active, RedDev red) {
  __shared__ double sh[CMM_RED_SH];
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const double* sc = red.scal + inst * NSCAL;
  const double pq = sc[S_ADJ_PQ];
  const double alpha = (pq != 0.0) ? sc[S_ADJ_RR + a.parity] / pq : 0.0;
  const long long off = inst * a.fstride;
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.fstride;
       i += (long long)gridDim.x * blockDim.x) {
    const long long idx = off + i;
    a.y[idx] += alpha * a.p[idx];
    const double rn = a.r[idx] - alpha * a.q[idx];
    a.r[idx] = rn;
    acc[0] += rn * rn;
  }
  const int ops[1] = {0};
  block_reduce<1>(acc, ops, sh);
  grid_reduce_store<1>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// lambda = S y; reduces sum lambda (= dN/dmu).
__global__ void __launch_bounds__(NT) k_adj_lambda(AdjDev a, const int* __restrict__ active, RedDev red) {
  __shared__ double sh[CMM_RED_SH];
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const long long off = inst * a.fstride;
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.fstride;
       i += (long long)gridDim.x * blockDim.x) {
    const long long idx = off + i;
    const double l = sqrt(a.beta * fmax(a.rho[idx], 0.0)) * a.y[idx];
    a.lam[idx] = l;
    acc[0] += l;
  }
  const int ops[1] = {0};
  block_reduce<1>(acc, ops, sh);
  grid_reduce_store<1>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

}  // namespace cmm
