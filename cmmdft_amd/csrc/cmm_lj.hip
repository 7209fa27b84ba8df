// Lennard-Jones host-guest external potential on the grid, FP64.
//
// One block owns a brick of 2 x 4 x 32 grid points (z contiguous, so the
// writes of a wavefront are coalesced).  The host atoms are streamed through
// LDS in chunks; each chunk is first culled against the brick (an atom whose
// minimum-image distance from the brick centre is at least r_c + the brick's
// circumradius cannot reach any point of the brick: the minimum-image
// distance is 1-Lipschitz in the point), then compacted in atom order, so the
// per-point sum runs over the same atoms in the same order as a plain loop.
#include "cmm_common.h"

namespace cmm {

constexpr int LJ_NT = 256;
constexpr int LJ_BX = 2, LJ_BY = 4, LJ_BZ = 32;

struct LJDev {
  int nx, ny, nz;
  double rv[9];
  double L[3];
  double inv_n[3];
  double rc;
  double sigf, epsf;
  int images;
  int nimg[3];
  double limit;
};

__device__ __forceinline__ void lj_point(const LJDev& d, int ix, int iy, int iz, double* p) {
  const double f0 = ix * d.inv_n[0] + (-0.5);
  const double f1 = iy * d.inv_n[1] + (-0.5);
  const double f2 = iz * d.inv_n[2] + (-0.5);
#pragma unroll
  for (int b = 0; b < 3; ++b) p[b] = d.rv[0 * 3 + b] * f0 + d.rv[1 * 3 + b] * f1 + d.rv[2 * 3 + b] * f2;
}

__device__ __forceinline__ double min_image(double x, double L) { return x - L * rint(x / L); }

// Sum over the host atoms for one point p, streaming atom chunks through
// LDS.  (c, reach): the block's bounding sphere centre and r_c + radius.
struct LJShared {
  double x[LJ_NT], y[LJ_NT], z[LJ_NT], sig[LJ_NT], eps[LJ_NT], sh[LJ_NT];
  int wcnt[LJ_NT / 64];
  int total;
};

__device__ double lj_sum(const double* __restrict__ atoms, int natom, const LJDev& d, const double* p, const double* c,
                         double reach, LJShared& S) {
  const int tid = threadIdx.x;
  const int nimg_tot = d.images ? (2 * d.nimg[0] + 1) * (2 * d.nimg[1] + 1) * (2 * d.nimg[2] + 1) : 1;
  const long long ntot = (long long)natom * nimg_tot;
  double acc = 0.0;
  for (long long c0 = 0; c0 < ntot; c0 += LJ_NT) {
    const long long j = c0 + tid;
    bool keep = false;
    double ax = 0, ay = 0, az = 0, sm = 0, em = 0, shift = 0;
    if (j < ntot) {
      const int ia = (int)(j / nimg_tot);
      const int im = (int)(j - (long long)ia * nimg_tot);
      const double* A = atoms + 5 * ia;
      ax = A[0];
      ay = A[1];
      az = A[2];
      double dx, dy, dz;
      if (d.images) {
        const int n0 = 2 * d.nimg[0] + 1, n1 = 2 * d.nimg[1] + 1;
        const int i0 = im % n0 - d.nimg[0];
        const int i1 = (im / n0) % n1 - d.nimg[1];
        const int i2 = im / (n0 * n1) - d.nimg[2];
        ax += i0 * d.rv[0] + i1 * d.rv[3] + i2 * d.rv[6];
        ay += i0 * d.rv[1] + i1 * d.rv[4] + i2 * d.rv[7];
        az += i0 * d.rv[2] + i1 * d.rv[5] + i2 * d.rv[8];
        dx = c[0] - ax;
        dy = c[1] - ay;
        dz = c[2] - az;
      } else {
        dx = min_image(c[0] - ax, d.L[0]);
        dy = min_image(c[1] - ay, d.L[1]);
        dz = min_image(c[2] - az, d.L[2]);
      }
      keep = (dx * dx + dy * dy + dz * dz) < reach * reach;
      // Lorentz-Berthelot mixing and the cut-off shift
      sm = 0.5 * (A[3] + d.sigf);
      em = sqrt(A[4] * d.epsf);
      const double q = sm / d.rc;
      const double q2 = q * q;
      const double rc6 = q2 * q2 * q2;
      shift = 4.0 * em * (rc6 * rc6 - rc6);
    }
    // order-preserving compaction
    const unsigned long long mask = __ballot(keep);
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) S.wcnt[wid] = __popcll(mask);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wid; ++w) off += S.wcnt[w];
    if (keep) {
      const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
      S.x[pos] = ax;
      S.y[pos] = ay;
      S.z[pos] = az;
      S.sig[pos] = sm;
      S.eps[pos] = em;
      S.sh[pos] = shift;
    }
    if (tid == 0) {
      int t = 0;
      for (int w = 0; w < LJ_NT / 64; ++w) t += S.wcnt[w];
      S.total = t;
    }
    __syncthreads();
    const int cnt = S.total;
    for (int e = 0; e < cnt; ++e) {
      double rx = p[0] - S.x[e], ry = p[1] - S.y[e], rz = p[2] - S.z[e];
      if (!d.images) {
        rx = min_image(rx, d.L[0]);
        ry = min_image(ry, d.L[1]);
        rz = min_image(rz, d.L[2]);
      }
      const double R = sqrt(rx * rx + ry * ry + rz * rz + 1e-16);
      if (R < d.rc) {
        const double s = S.sig[e] / R;
        const double s2 = s * s;
        const double s6 = s2 * s2 * s2;
        acc += 4.0 * S.eps[e] * (s6 * s6 - s6) - S.sh[e];
      }
    }
    __syncthreads();
  }
  return acc;
}

__device__ __forceinline__ double lj_finish(const LJDev& d, double acc) {
  if (d.limit > 0.0 && (!isfinite(acc) || acc > d.limit)) acc = d.limit;
  return acc;
}

// Regular-grid form: one block per brick of LJ_BX x LJ_BY x LJ_BZ points.
__global__ void __launch_bounds__(LJ_NT) k_lj(const double* __restrict__ atoms, int natom, LJDev d,
                                             double* __restrict__ V) {
  __shared__ LJShared S;
  const int tid = threadIdx.x;
  const int bx0 = blockIdx.z * LJ_BX, by0 = blockIdx.y * LJ_BY, bz0 = blockIdx.x * LJ_BZ;
  const int ix = bx0 + tid / (LJ_BY * LJ_BZ);
  const int iy = by0 + (tid / LJ_BZ) % LJ_BY;
  const int iz = bz0 + tid % LJ_BZ;
  const bool valid = (ix < d.nx) && (iy < d.ny) && (iz < d.nz);
  double p[3];
  lj_point(d, ix, iy, iz, p);
  const int ex = min(LJ_BX, d.nx - bx0), ey = min(LJ_BY, d.ny - by0), ez = min(LJ_BZ, d.nz - bz0);
  double c[3];
  {
    const double f0 = (bx0 + 0.5 * (ex - 1)) * d.inv_n[0] - 0.5;
    const double f1 = (by0 + 0.5 * (ey - 1)) * d.inv_n[1] - 0.5;
    const double f2 = (bz0 + 0.5 * (ez - 1)) * d.inv_n[2] - 0.5;
#pragma unroll
    for (int b = 0; b < 3; ++b) c[b] = d.rv[b] * f0 + d.rv[3 + b] * f1 + d.rv[6 + b] * f2;
  }
  double hd = 0.0;
  {
    const double e3[3] = {0.5 * (ex - 1) * d.inv_n[0], 0.5 * (ey - 1) * d.inv_n[1], 0.5 * (ez - 1) * d.inv_n[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      hd += e3[a] * sqrt(d.rv[3 * a] * d.rv[3 * a] + d.rv[3 * a + 1] * d.rv[3 * a + 1] + d.rv[3 * a + 2] * d.rv[3 * a + 2]);
  }
  const double acc = lj_sum(atoms, natom, d, p, c, d.rc + hd + 1e-6, S);
  if (valid) V[((long long)ix * d.ny + iy) * d.nz + iz] = lj_finish(d, acc);
}

// Explicit point list (M x 3): one block per 256 consecutive points, bounded
// by the block's axis-aligned box.
__global__ void __launch_bounds__(LJ_NT) k_lj_points(const double* __restrict__ atoms, int natom, LJDev d,
                                                    const double* __restrict__ pts, long long M,
                                                    double* __restrict__ V) {
  __shared__ LJShared S;
  __shared__ double bb[6][LJ_NT / 64];
  const int tid = threadIdx.x;
  const long long i = (long long)blockIdx.x * LJ_NT + tid;
  const bool valid = i < M;
  const long long ii = valid ? i : (M - 1);
  double p[3] = {pts[3 * ii], pts[3 * ii + 1], pts[3 * ii + 2]};
  double lo[3], hi[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double mn = p[b], mx = p[b];
    for (int o = 32; o > 0; o >>= 1) {
      mn = fmin(mn, __shfl_xor(mn, o, 64));
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((tid & 63) == 0) { bb[b][tid >> 6] = mn; bb[3 + b][tid >> 6] = mx; }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    lo[b] = bb[b][0];
    hi[b] = bb[3 + b][0];
    for (int w = 1; w < LJ_NT / 64; ++w) { lo[b] = fmin(lo[b], bb[b][w]); hi[b] = fmax(hi[b], bb[3 + b][w]); }
  }
  double c[3], r2 = 0.0;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    c[b] = 0.5 * (lo[b] + hi[b]);
    r2 += 0.25 * (hi[b] - lo[b]) * (hi[b] - lo[b]);
  }
  const double acc = lj_sum(atoms, natom, d, p, c, d.rc + sqrt(r2) + 1e-6, S);
  if (valid) V[i] = lj_finish(d, acc);
}

// ---------------------------------------------------------------------------
// Backward of the grid sum with respect to the host atom-type parameters:
//   G[2t]   = sum_i w_i dV_i/dsigma_t,   G[2t+1] = sum_i w_i dV_i/depsilon_t
// for the Lorentz-Berthelot pair sigma_m = (sigma_a + sigma_f)/2,
// epsilon_m = sqrt(epsilon_a epsilon_f):
//   dV_pair/dsigma_a   = 2 eps_m [(12 s^12 - 6 s^6) - (12 q^12 - 6 q^6)] / sigma_m
//   dV_pair/depsilon_a = V_pair / (2 epsilon_a)
// with s = sigma_m/R, q = sigma_m/r_c.  Points whose total V was clamped (or
// is not finite) have zero derivative.  One pass over the atoms per type,
// with the same brick culling as the forward sum; per-block partial sums,
// reduced in a fixed order by k_lj_bwd_reduce.
// ---------------------------------------------------------------------------
struct LJGradShared {
  double x[LJ_NT], y[LJ_NT], z[LJ_NT], sig[LJ_NT], eps[LJ_NT], sh[LJ_NT], fe[LJ_NT], dsh[LJ_NT];
  int wcnt[LJ_NT / 64];
  int total;
};

__device__ void lj_grad_type(const double* __restrict__ atoms, const int* __restrict__ types, int natom, int tsel,
                             const LJDev& d, const double* p, const double* c, double reach, LJGradShared& S,
                             double& gsig, double& geps) {
  const int tid = threadIdx.x;
  const int nimg_tot = d.images ? (2 * d.nimg[0] + 1) * (2 * d.nimg[1] + 1) * (2 * d.nimg[2] + 1) : 1;
  const long long ntot = (long long)natom * nimg_tot;
  gsig = 0.0;
  geps = 0.0;
  for (long long c0 = 0; c0 < ntot; c0 += LJ_NT) {
    const long long j = c0 + tid;
    bool keep = false;
    double ax = 0, ay = 0, az = 0, sm = 0, em = 0, shift = 0, fe = 0, dsh = 0;
    if (j < ntot) {
      const int ia = (int)(j / nimg_tot);
      const int im = (int)(j - (long long)ia * nimg_tot);
      const double* A = atoms + 5 * ia;
      ax = A[0];
      ay = A[1];
      az = A[2];
      double dx, dy, dz;
      if (d.images) {
        const int n0 = 2 * d.nimg[0] + 1, n1 = 2 * d.nimg[1] + 1;
        const int i0 = im % n0 - d.nimg[0];
        const int i1 = (im / n0) % n1 - d.nimg[1];
        const int i2 = im / (n0 * n1) - d.nimg[2];
        ax += i0 * d.rv[0] + i1 * d.rv[3] + i2 * d.rv[6];
        ay += i0 * d.rv[1] + i1 * d.rv[4] + i2 * d.rv[7];
        az += i0 * d.rv[2] + i1 * d.rv[5] + i2 * d.rv[8];
        dx = c[0] - ax;
        dy = c[1] - ay;
        dz = c[2] - az;
      } else {
        dx = min_image(c[0] - ax, d.L[0]);
        dy = min_image(c[1] - ay, d.L[1]);
        dz = min_image(c[2] - az, d.L[2]);
      }
      keep = types[ia] == tsel && (dx * dx + dy * dy + dz * dz) < reach * reach;
      sm = 0.5 * (A[3] + d.sigf);
      em = sqrt(A[4] * d.epsf);
      const double q = sm / d.rc;
      const double q2 = q * q;
      const double rc6 = q2 * q2 * q2;
      shift = 4.0 * em * (rc6 * rc6 - rc6);
      fe = 0.5 / A[4];
      dsh = 2.0 * em * (12.0 * rc6 * rc6 - 6.0 * rc6) / sm;
    }
    const unsigned long long mask = __ballot(keep);
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) S.wcnt[wid] = __popcll(mask);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wid; ++w) off += S.wcnt[w];
    if (keep) {
      const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
      S.x[pos] = ax;
      S.y[pos] = ay;
      S.z[pos] = az;
      S.sig[pos] = sm;
      S.eps[pos] = em;
      S.sh[pos] = shift;
      S.fe[pos] = fe;
      S.dsh[pos] = dsh;
    }
    if (tid == 0) {
      int t = 0;
      for (int w = 0; w < LJ_NT / 64; ++w) t += S.wcnt[w];
      S.total = t;
    }
    __syncthreads();
    const int cnt = S.total;
    for (int e = 0; e < cnt; ++e) {
      double rx = p[0] - S.x[e], ry = p[1] - S.y[e], rz = p[2] - S.z[e];
      if (!d.images) {
        rx = min_image(rx, d.L[0]);
        ry = min_image(ry, d.L[1]);
        rz = min_image(rz, d.L[2]);
      }
      const double R = sqrt(rx * rx + ry * ry + rz * rz + 1e-16);
      if (R < d.rc) {
        const double s = S.sig[e] / R;
        const double s2 = s * s;
        const double s6 = s2 * s2 * s2;
        const double vp = 4.0 * S.eps[e] * (s6 * s6 - s6) - S.sh[e];
        geps += vp * S.fe[e];
        gsig += 2.0 * S.eps[e] * (12.0 * s6 * s6 - 6.0 * s6) / S.sig[e] - S.dsh[e];
      }
    }
    __syncthreads();
  }
}

// partials: [2 ntypes][nblocks] (nblocks = gridDim.x*gridDim.y*gridDim.z);
// w: nw weight fields of nx*ny*nz points
__global__ void __launch_bounds__(LJ_NT) k_lj_bwd(const double* __restrict__ atoms, const int* __restrict__ types,
                                                 int natom, int ntypes, LJDev d, const double* __restrict__ w, int nw,
                                                 double* __restrict__ partials) {
  __shared__ LJShared S;
  __shared__ LJGradShared SG;
  __shared__ double sh[CMM_RED_SH];
  const int tid = threadIdx.x;
  const int bx0 = blockIdx.z * LJ_BX, by0 = blockIdx.y * LJ_BY, bz0 = blockIdx.x * LJ_BZ;
  const long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const long long nblocks = (long long)gridDim.x * gridDim.y * gridDim.z;
  const int ix = bx0 + tid / (LJ_BY * LJ_BZ);
  const int iy = by0 + (tid / LJ_BZ) % LJ_BY;
  const int iz = bz0 + tid % LJ_BZ;
  const bool valid = (ix < d.nx) && (iy < d.ny) && (iz < d.nz);
  double p[3];
  lj_point(d, ix, iy, iz, p);
  const int ex = min(LJ_BX, d.nx - bx0), ey = min(LJ_BY, d.ny - by0), ez = min(LJ_BZ, d.nz - bz0);
  double c[3];
  {
    const double f0 = (bx0 + 0.5 * (ex - 1)) * d.inv_n[0] - 0.5;
    const double f1 = (by0 + 0.5 * (ey - 1)) * d.inv_n[1] - 0.5;
    const double f2 = (bz0 + 0.5 * (ez - 1)) * d.inv_n[2] - 0.5;
#pragma unroll
    for (int b = 0; b < 3; ++b) c[b] = d.rv[b] * f0 + d.rv[3 + b] * f1 + d.rv[6 + b] * f2;
  }
  double hd = 0.0;
  {
    const double e3[3] = {0.5 * (ex - 1) * d.inv_n[0], 0.5 * (ey - 1) * d.inv_n[1], 0.5 * (ez - 1) * d.inv_n[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      hd += e3[a] * sqrt(d.rv[3 * a] * d.rv[3 * a] + d.rv[3 * a + 1] * d.rv[3 * a + 1] + d.rv[3 * a + 2] * d.rv[3 * a + 2]);
  }
  const double reach = d.rc + hd + 1e-6;
  const double V = lj_sum(atoms, natom, d, p, c, reach, S);
  const bool clamped = d.limit > 0.0 && (!isfinite(V) || V > d.limit);
  // weight of the point: the sum of the nw weight fields (e.g. the adjoints of
  // the state points of one framework), in field order
  double wi = 0.0;
  if (valid && !clamped) {
    const long long idx = ((long long)ix * d.ny + iy) * d.nz + iz;
    const long long np = (long long)d.nx * d.ny * d.nz;
    for (int k = 0; k < nw; ++k) wi += w[k * np + idx];
  }
  for (int t = 0; t < ntypes; ++t) {
    double gs, ge;
    lj_grad_type(atoms, types, natom, t, d, p, c, reach, SG, gs, ge);
    double acc[2] = {wi * gs, wi * ge};
    if (wi == 0.0) acc[0] = acc[1] = 0.0;
    const int ops[2] = {0, 0};
    block_reduce<2>(acc, ops, sh);
    if (tid == 0) {
      partials[(2 * t) * nblocks + bid] = acc[0];
      partials[(2 * t + 1) * nblocks + bid] = acc[1];
    }
  }
}

// out[q] = sum over blocks of partials[q][.], fixed order; one block per q.
__global__ void __launch_bounds__(LJ_NT) k_lj_bwd_reduce(const double* __restrict__ partials, long long nblocks,
                                                        double* __restrict__ out) {
  __shared__ double sh[LJ_NT / 64];
  const int q = blockIdx.x;
  double a = 0.0;
  for (long long i = threadIdx.x; i < nblocks; i += blockDim.x) a += partials[q * nblocks + i];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int w = 1; w < LJ_NT / 64; ++w) t += sh[w];
    out[q] = t;
  }
}

}  // namespace cmm
