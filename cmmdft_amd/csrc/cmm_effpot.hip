// Orientationally averaged external potential of a rigid, non-spherical guest,
// FP64 (the direct form of the reference's generate_effective_potential,
// cmmdft/extpot_calculator.py:572-646, with per-atom generators
// get_external_potential, :186-222):
//
//   V_eff(p) = -(1/beta) log sum_r w_r exp(-beta U_r(p)),
//   U_r(p)   = sum_i V_{t(i)}(p + d_ri),   d_ri = COM + R_r (x_i - COM),
//
// where V_t is the host's truncated-and-shifted LJ field seen by a guest atom
// of type t (Lorentz-Berthelot mixing, minimum image with the row norms of
// the cell, strict cut-off), R_r the orientation matrices and w_r their
// weights.  A result that is NaN becomes `nanval` (the reference's
// np.where(isnan) -> limit_potential).
//
// One block owns 256 points (a 4 x 8 x 8 brick of the grid, or 256
// consecutive points of a list).  The orientations are taken EP_G at a time:
// each thread keeps U_r of EP_G orientations in registers while the host
// atoms stream through LDS in chunks, culled against the block's bounding
// sphere grown by the guest's extent max |d_ri| (an atom at least
// r_c + radius + max|d| from the centre reaches no guest atom of any point
// of the block: the minimum-image distance is 1-Lipschitz), with the mixed
// pair parameters of every guest type stored beside the compacted atoms.
// One LDS read of an atom then serves EP_G pair evaluations.  The sum over
// orientations is an online log-sum-exp (running maximum and rescaled sum).
// The work is FP64 VALU bound (about 30 operations per pair).
// Needs cmm_lj.hip (LJDev, lj_point, min_image), included before it.

namespace cmm {

constexpr int EP_NT = 256;
constexpr int EP_BX = 4, EP_BY = 8, EP_BZ = 8;
constexpr int EP_MAXG = 16;   // guest atoms
constexpr int EP_MAXT = 4;    // guest atom types
constexpr int EP_G = 8;       // orientations per register group

struct EffDev {
  LJDev d;       // cell, grid, cut-off, image mode (sigf/epsf unused)
  int nrot, ng, nt;
  int gtype[EP_MAXG];
  double gsig[EP_MAXT], geps[EP_MAXT];
  double dmax;
  double beta;
  double nanval;
};

struct EffShared {
  double x[EP_NT], y[EP_NT], z[EP_NT];
  double s2[EP_MAXT][EP_NT], e4[EP_MAXT][EP_NT], sh[EP_MAXT][EP_NT];
  int wcnt[EP_NT / 64];
  int total;
};

// Cull the atoms [c0, c0 + EP_NT) of the (atom x image) list against the
// sphere (c, reach), compact the kept ones in order into S with the mixed
// parameters of every guest type: sigma_m^2, 4 eps_m, and the shift
// 4 eps_m ((sigma_m/r_c)^12 - (sigma_m/r_c)^6).  Returns the count; ends with
// a barrier.
__device__ int ep_cull(const double* __restrict__ atoms, long long c0, long long ntot, int nimg_tot, const EffDev& e,
                       const double* c, double reach, EffShared& S) {
  const LJDev& d = e.d;
  const int tid = threadIdx.x;
  const long long j = c0 + tid;
  bool keep = false;
  double ax = 0, ay = 0, az = 0, as = 0, ae = 0;
  if (j < ntot) {
    const int ia = (int)(j / nimg_tot);
    const int im = (int)(j - (long long)ia * nimg_tot);
    const double* A = atoms + 5 * ia;
    ax = A[0];
    ay = A[1];
    az = A[2];
    as = A[3];
    ae = A[4];
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
    for (int t = 0; t < e.nt; ++t) {
      const double sm = 0.5 * (as + e.gsig[t]);
      const double em = sqrt(ae * e.geps[t]);
      const double q = sm / d.rc;
      const double q2 = q * q;
      const double rc6 = q2 * q2 * q2;
      S.s2[t][pos] = sm * sm;
      S.e4[t][pos] = 4.0 * em;
      S.sh[t][pos] = 4.0 * em * (rc6 * rc6 - rc6);
    }
  }
  if (tid == 0) {
    int t = 0;
    for (int w = 0; w < EP_NT / 64; ++w) t += S.wcnt[w];
    S.total = t;
  }
  __syncthreads();
  return S.total;
}

// 1/x to within an ulp: hardware reciprocal and one Newton step.
__device__ __forceinline__ double ep_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  const double err = fma(-x, r, 1.0);
  return fma(r, err, r);
}

// Online log-sum-exp: fold a = -beta U with weight w into (mx, sum, nan).
__device__ __forceinline__ void ep_fold(double a, double w, double& mx, double& sum, bool& nan) {
  if (a != a) {
    nan = true;
  } else if (a > mx) {
    sum = sum * exp(mx - a) + w;   // exp(-inf) = 0 on the first fold
    mx = a;
  } else if (a != -INFINITY) {
    sum += w * exp(a - mx);
  }
}

// Point coordinates and the block's bounding sphere (centre c, radius hd).
template <bool GRID>
__device__ __forceinline__ bool ep_point(const EffDev& e, const double* __restrict__ pts, long long M, double* p,
                                         double* c, double& hd, long long& out_idx) {
  const LJDev& d = e.d;
  const int tid = threadIdx.x;
  if constexpr (GRID) {
    const int bx0 = blockIdx.z * EP_BX, by0 = blockIdx.y * EP_BY, bz0 = blockIdx.x * EP_BZ;
    const int ix = bx0 + tid / (EP_BY * EP_BZ);
    const int iy = by0 + (tid / EP_BZ) % EP_BY;
    const int iz = bz0 + tid % EP_BZ;
    lj_point(d, ix, iy, iz, p);
    const int ex = min(EP_BX, d.nx - bx0), ey = min(EP_BY, d.ny - by0), ez = min(EP_BZ, d.nz - bz0);
    const double f0 = (bx0 + 0.5 * (ex - 1)) * d.inv_n[0] - 0.5;
    const double f1 = (by0 + 0.5 * (ey - 1)) * d.inv_n[1] - 0.5;
    const double f2 = (bz0 + 0.5 * (ez - 1)) * d.inv_n[2] - 0.5;
#pragma unroll
    for (int b = 0; b < 3; ++b) c[b] = d.rv[b] * f0 + d.rv[3 + b] * f1 + d.rv[6 + b] * f2;
    const double e3[3] = {0.5 * (ex - 1) * d.inv_n[0], 0.5 * (ey - 1) * d.inv_n[1], 0.5 * (ez - 1) * d.inv_n[2]};
    hd = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      hd += e3[a] * sqrt(d.rv[3 * a] * d.rv[3 * a] + d.rv[3 * a + 1] * d.rv[3 * a + 1] + d.rv[3 * a + 2] * d.rv[3 * a + 2]);
    out_idx = ((long long)ix * d.ny + iy) * d.nz + iz;
    return (ix < d.nx) && (iy < d.ny) && (iz < d.nz);
  } else {
    __shared__ double bb[6][EP_NT / 64];
    const long long i = (long long)blockIdx.x * EP_NT + tid;
    const bool valid = i < M;
    const long long ii = valid ? i : (M - 1);
#pragma unroll
    for (int b = 0; b < 3; ++b) p[b] = pts[3 * ii + b];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double mn = p[b], mx = p[b];
      for (int o = 32; o > 0; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
      }
      if ((tid & 63) == 0) {
        bb[b][tid >> 6] = mn;
        bb[3 + b][tid >> 6] = mx;
      }
    }
    __syncthreads();
    double r2 = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double lo = bb[b][0], hi = bb[3 + b][0];
      for (int w = 1; w < EP_NT / 64; ++w) {
        lo = fmin(lo, bb[b][w]);
        hi = fmax(hi, bb[3 + b][w]);
      }
      c[b] = 0.5 * (lo + hi);
      r2 += 0.25 * (hi - lo) * (hi - lo);
    }
    hd = sqrt(r2);
    out_idx = i;
    return valid;
  }
}

// off: [nrot][ng][3] guest-atom offsets d_ri; w: [nrot] weights.
template <bool GRID, bool IMG>
__global__ void __launch_bounds__(EP_NT) k_effpot(const double* __restrict__ atoms, int natom, EffDev e,
                                                 const double* __restrict__ off, const double* __restrict__ w,
                                                 const double* __restrict__ pts, long long M,
                                                 double* __restrict__ V) {
  __shared__ EffShared S;
  const LJDev& d = e.d;
  double p[3], c[3], hd;
  long long oi;
  const bool valid = ep_point<GRID>(e, pts, M, p, c, hd, oi);
  const double reach = d.rc + hd + e.dmax + 1e-6;
  const double rc2 = d.rc * d.rc;
  const double iL0 = 1.0 / d.L[0], iL1 = 1.0 / d.L[1], iL2 = 1.0 / d.L[2];
  const int nimg_tot = d.images ? (2 * d.nimg[0] + 1) * (2 * d.nimg[1] + 1) * (2 * d.nimg[2] + 1) : 1;
  const long long ntot = (long long)natom * nimg_tot;
  double mx = -INFINITY, sum = 0.0;
  bool nan = false;
  for (int r0 = 0; r0 < e.nrot; r0 += EP_G) {
    double acc[EP_G];
#pragma unroll
    for (int g = 0; g < EP_G; ++g) acc[g] = 0.0;
    for (long long c0 = 0; c0 < ntot; c0 += EP_NT) {
      const int cnt = ep_cull(atoms, c0, ntot, nimg_tot, e, c, reach, S);
      for (int i = 0; i < e.ng; ++i) {
        const int t = e.gtype[i];
        double qx[EP_G], qy[EP_G], qz[EP_G];
#pragma unroll
        for (int g = 0; g < EP_G; ++g) {
          const int r = min(r0 + g, e.nrot - 1);
          const double* o = off + ((long long)r * e.ng + i) * 3;
          qx[g] = p[0] + o[0];
          qy[g] = p[1] + o[1];
          qz[g] = p[2] + o[2];
        }
        const double* s2t = S.s2[t];
        const double* e4t = S.e4[t];
        const double* sht = S.sh[t];
        for (int k = 0; k < cnt; ++k) {
          const double ax = S.x[k], ay = S.y[k], az = S.z[k];
          const double s2m = s2t[k], e4 = e4t[k], sh = sht[k];
#pragma unroll
          for (int g = 0; g < EP_G; ++g) {
            double rx = qx[g] - ax, ry = qy[g] - ay, rz = qz[g] - az;
            if constexpr (!IMG) {
              rx = fma(-d.L[0], rint(rx * iL0), rx);
              ry = fma(-d.L[1], rint(ry * iL1), ry);
              rz = fma(-d.L[2], rint(rz * iL2), rz);
            }
            const double r2 = rx * rx + ry * ry + rz * rz + 1e-16;
            if (r2 < rc2) {
              const double s2 = s2m * ep_rcp(r2);
              const double s6 = s2 * s2 * s2;
              acc[g] += fma(e4, s6 * s6 - s6, -sh);
            }
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int g = 0; g < EP_G; ++g)
      if (r0 + g < e.nrot) ep_fold(-e.beta * acc[g], w[r0 + g], mx, sum, nan);
  }
  if (valid) {
    double v = (sum > 0.0) ? -(mx + log(sum)) / e.beta : INFINITY;
    if (nan || v != v) v = e.nanval;
    V[oi] = v;
  }
}

}  // namespace cmm

// ---------------------------------------------------------------------------
// Interpolated form (the reference's interpolate=True path,
// cmmdft/functionals.py:1035-1055): the LJ field of each guest-atom type and
// its derivatives on a fine grid (get_external_potential_derivatives,
// extpot_calculator.py:264-346), interpolated at the rotated sites
// (Interpolator.tricubic_interpolation / trilinear_interpolation, :391-548).
// ---------------------------------------------------------------------------
namespace cmm {

constexpr int LJD_NT = 256;

// Field and derivative channels at one point: V, dV/dx, dV/dy, dV/dz,
// d2V/dxdy, d2V/dxdz, d2V/dydz, d3V/dxdydz of the truncated-and-shifted LJ
// sum (minimum image; derivatives of the unshifted pair potential inside the
// cut-off, zero beyond).
__device__ void ljd_sum(const double* __restrict__ atoms, int natom, const LJDev& d, const double* p,
                        const double* c, double reach, LJShared& S, double (&acc)[8]) {
  const int tid = threadIdx.x;
  const int nimg_tot = d.images ? (2 * d.nimg[0] + 1) * (2 * d.nimg[1] + 1) * (2 * d.nimg[2] + 1) : 1;
  const long long ntot = (long long)natom * nimg_tot;
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = 0.0;
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
      sm = 0.5 * (A[3] + d.sigf);
      em = sqrt(A[4] * d.epsf);
      const double q = sm / d.rc;
      const double q2 = q * q;
      const double rc6 = q2 * q2 * q2;
      shift = 4.0 * em * (rc6 * rc6 - rc6);
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
      const double r2 = rx * rx + ry * ry + rz * rz + 1e-16;
      const double R = sqrt(r2);
      if (R < d.rc) {
        const double s = S.sig[e] / R;
        const double s2 = s * s;
        const double s6 = s2 * s2 * s2;
        const double s12 = s6 * s6;
        const double e4 = 4.0 * S.eps[e];
        const double ir2 = 1.0 / r2;
        const double d1 = 6.0 * e4 * (s6 - 2.0 * s12) * ir2;
        const double d2 = 24.0 * e4 * (7.0 * s12 - 2.0 * s6) * (ir2 * ir2);
        const double d3 = 96.0 * e4 * (5.0 * s6 - 28.0 * s12) * (ir2 * ir2 * ir2);
        acc[0] += e4 * (s12 - s6) - S.sh[e];
        acc[1] += d1 * rx;
        acc[2] += d1 * ry;
        acc[3] += d1 * rz;
        acc[4] += d2 * rx * ry;
        acc[5] += d2 * rx * rz;
        acc[6] += d2 * ry * rz;
        acc[7] += d3 * rx * ry * rz;
      }
    }
    __syncthreads();
  }
}

// Clamps and the unit-cube scaling of get_external_potential_derivatives
// (:326-346): V and the first derivatives clipped to +-vmax; the mixed
// derivatives zeroed where V > vmax; derivative k times the product of the
// spacings of its axes.  out[q * stride + idx].
__device__ __forceinline__ void ljd_store(double (&a)[8], const double* h, double vmax, double* out,
                                          long long stride, long long idx) {
  const bool big = a[0] > vmax;
  const double v = fmin(fmax(a[0], -vmax), vmax);
  out[idx] = v;
  out[stride + idx] = fmin(fmax(a[1], -vmax), vmax) * h[0];
  out[2 * stride + idx] = fmin(fmax(a[2], -vmax), vmax) * h[1];
  out[3 * stride + idx] = fmin(fmax(a[3], -vmax), vmax) * h[2];
  out[4 * stride + idx] = big ? 0.0 : a[4] * (h[0] * h[1]);
  out[5 * stride + idx] = big ? 0.0 : a[5] * (h[0] * h[2]);
  out[6 * stride + idx] = big ? 0.0 : a[6] * (h[1] * h[2]);
  out[7 * stride + idx] = big ? 0.0 : a[7] * (h[0] * h[1] * h[2]);
}

struct LJDerivArgs {
  double h[3];
  double vmax;
};

// grid form (LJ_BX x LJ_BY x LJ_BZ bricks, as k_lj) or point-list form
template <bool GRID>
__global__ void __launch_bounds__(LJ_NT) k_lj_derivs(const double* __restrict__ atoms, int natom, LJDev d,
                                                    LJDerivArgs da, const double* __restrict__ pts, long long M,
                                                    double* __restrict__ out) {
  __shared__ LJShared S;
  const int tid = threadIdx.x;
  double p[3], c[3], reach;
  long long idx, stride;
  bool valid;
  if constexpr (GRID) {
    const int bx0 = blockIdx.z * LJ_BX, by0 = blockIdx.y * LJ_BY, bz0 = blockIdx.x * LJ_BZ;
    const int ix = bx0 + tid / (LJ_BY * LJ_BZ);
    const int iy = by0 + (tid / LJ_BZ) % LJ_BY;
    const int iz = bz0 + tid % LJ_BZ;
    valid = (ix < d.nx) && (iy < d.ny) && (iz < d.nz);
    lj_point(d, ix, iy, iz, p);
    const int ex = min(LJ_BX, d.nx - bx0), ey = min(LJ_BY, d.ny - by0), ez = min(LJ_BZ, d.nz - bz0);
    const double f0 = (bx0 + 0.5 * (ex - 1)) * d.inv_n[0] - 0.5;
    const double f1 = (by0 + 0.5 * (ey - 1)) * d.inv_n[1] - 0.5;
    const double f2 = (bz0 + 0.5 * (ez - 1)) * d.inv_n[2] - 0.5;
#pragma unroll
    for (int b = 0; b < 3; ++b) c[b] = d.rv[b] * f0 + d.rv[3 + b] * f1 + d.rv[6 + b] * f2;
    const double e3[3] = {0.5 * (ex - 1) * d.inv_n[0], 0.5 * (ey - 1) * d.inv_n[1], 0.5 * (ez - 1) * d.inv_n[2]};
    double hd = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      hd += e3[a] * sqrt(d.rv[3 * a] * d.rv[3 * a] + d.rv[3 * a + 1] * d.rv[3 * a + 1] + d.rv[3 * a + 2] * d.rv[3 * a + 2]);
    reach = d.rc + hd + 1e-6;
    idx = ((long long)ix * d.ny + iy) * d.nz + iz;
    stride = (long long)d.nx * d.ny * d.nz;
  } else {
    __shared__ double bb[6][LJD_NT / 64];
    const long long i = (long long)blockIdx.x * LJD_NT + tid;
    valid = i < M;
    const long long ii = valid ? i : (M - 1);
#pragma unroll
    for (int b = 0; b < 3; ++b) p[b] = pts[3 * ii + b];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double mn = p[b], mx = p[b];
      for (int o = 32; o > 0; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
      }
      if ((tid & 63) == 0) {
        bb[b][tid >> 6] = mn;
        bb[3 + b][tid >> 6] = mx;
      }
    }
    __syncthreads();
    double r2 = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double lo = bb[b][0], hi = bb[3 + b][0];
      for (int w = 1; w < LJD_NT / 64; ++w) {
        lo = fmin(lo, bb[b][w]);
        hi = fmax(hi, bb[3 + b][w]);
      }
      c[b] = 0.5 * (lo + hi);
      r2 += 0.25 * (hi - lo) * (hi - lo);
    }
    reach = d.rc + sqrt(r2) + 1e-6;
    idx = i;
    stride = M;
  }
  double acc[8];
  ljd_sum(atoms, natom, d, p, c, reach, S, acc);
  if (valid) ljd_store(acc, da.h, da.vmax, out, stride, idx);
}

// Channel-fastest interpolation table [nx][ny][nz][8] from the reference's
// channel-major grid values [8][nx][ny][nz] (EST = 0), or with the
// derivatives estimated from channel 0 by periodic central differences
// (EST = 1, Interpolator.estimate_all_derivatives, :454-470: each mixed
// derivative is the difference of the lower one at the two neighbours, the
// same operations in the same order as the reference's rolls).
__device__ __forceinline__ long long wrap3(int i, int j, int k, int nx, int ny, int nz) {
  i = (i + nx) % nx;
  j = (j + ny) % ny;
  k = (k + nz) % nz;
  return ((long long)i * ny + j) * nz + k;
}

template <int EST>
__global__ void k_interp_table(const double* __restrict__ vals, int nx, int ny, int nz, double* __restrict__ tab) {
  const long long n = (long long)nx * ny * nz;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double o[8];
  if (EST == 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = vals[q * n + idx];
  } else {
    const int k = (int)(idx % nz);
    const int j = (int)((idx / nz) % ny);
    const int i = (int)(idx / ((long long)ny * nz));
    auto V = [&](int a, int b, int c) { return vals[wrap3(a, b, c, nx, ny, nz)]; };
    auto Vx = [&](int a, int b, int c) { return (V(a + 1, b, c) - V(a - 1, b, c)) / 2; };
    auto Vy = [&](int a, int b, int c) { return (V(a, b + 1, c) - V(a, b - 1, c)) / 2; };
    auto Vz = [&](int a, int b, int c) { return (V(a, b, c + 1) - V(a, b, c - 1)) / 2; };
    auto Vxy = [&](int a, int b, int c) { return (Vx(a, b + 1, c) - Vx(a, b - 1, c)) / 2; };
    o[0] = V(i, j, k);
    o[1] = Vx(i, j, k);
    o[2] = Vy(i, j, k);
    o[3] = Vz(i, j, k);
    o[4] = Vxy(i, j, k);
    o[5] = (Vx(i, j, k + 1) - Vx(i, j, k - 1)) / 2;
    o[6] = (Vy(i, j, k + 1) - Vy(i, j, k - 1)) / 2;
    o[7] = (Vxy(i, j, k + 1) - Vxy(i, j, k - 1)) / 2;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) tab[idx * 8 + q] = o[q];
}

struct InterpTab {
  const double* tab;   // [nx][ny][nz][8]
  int n[3];
  double origin[3], h[3];
  int mode;            // 0 tricubic, 1 trilinear
};

// cubic Hermite basis on [0, 1]: value weights at 0 and 1, slope weights at 0 and 1
__device__ __forceinline__ void hermite(double t, double (&w)[2][2]) {
  const double t2 = t * t, t3 = t2 * t;
  w[0][0] = 2.0 * t3 - 3.0 * t2 + 1.0;
  w[1][0] = -2.0 * t3 + 3.0 * t2;
  w[0][1] = t3 - 2.0 * t2 + t;
  w[1][1] = t3 - t2;
}

// Interpolated value at q.  Tricubic: the unique polynomial of degree <= 3 in
// each coordinate matching V and its seven derivative channels at the cell's
// eight corners (the interpolant that the reference's 64 x 64 coefficient
// matrix, Lekien & Marsden 2005, gives), evaluated as the tensor product of
// one-dimensional cubic Hermite bases.  Indices wrap periodically with the
// table's own period; fractional index (q - origin) / h per Cartesian axis.
__device__ double interp_at(const InterpTab& T, double qx, double qy, double qz) {
  const double sx = (qx - T.origin[0]) / T.h[0];
  const double sy = (qy - T.origin[1]) / T.h[1];
  const double sz = (qz - T.origin[2]) / T.h[2];
  const double fx = floor(sx), fy = floor(sy), fz = floor(sz);
  const double rx = sx - fx, ry = sy - fy, rz = sz - fz;
  const int nx = T.n[0], ny = T.n[1], nz = T.n[2];
  const int ix = (int)(((long long)fx % nx + nx) % nx);
  const int iy = (int)(((long long)fy % ny + ny) % ny);
  const int iz = (int)(((long long)fz % nz + nz) % nz);
  const int jx = ix + 1 == nx ? 0 : ix + 1;
  const int jy = iy + 1 == ny ? 0 : iy + 1;
  const int jz = iz + 1 == nz ? 0 : iz + 1;
  const int xs[2] = {ix, jx}, ys[2] = {iy, jy}, zs[2] = {iz, jz};
  if (T.mode == 1) {
    const double wx[2] = {1.0 - rx, rx}, wy[2] = {1.0 - ry, ry}, wz[2] = {1.0 - rz, rz};
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
      const double f = T.tab[(((long long)xs[cx] * ny + ys[cy]) * nz + zs[cz]) * 8];
      v += f * (wx[cx] * wy[cy] * wz[cz]);
    }
    return v;
  }
  double wx[2][2], wy[2][2], wz[2][2];
  hermite(rx, wx);
  hermite(ry, wy);
  hermite(rz, wz);
  double v = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
    const double2* g = reinterpret_cast<const double2*>(T.tab + (((long long)xs[cx] * ny + ys[cy]) * nz + zs[cz]) * 8);
    const double2 a = g[0], b = g[1], e = g[2], f = g[3];
    // channels V, Vx, Vy, Vz, Vxy, Vxz, Vyz, Vxyz
    const double x0 = wx[cx][0], x1 = wx[cx][1], y0 = wy[cy][0], y1 = wy[cy][1], z0 = wz[cz][0], z1 = wz[cz][1];
    v += z0 * (y0 * (x0 * a.x + x1 * a.y) + y1 * (x0 * b.x + x1 * e.x)) +
         z1 * (y0 * (x0 * b.y + x1 * e.y) + y1 * (x0 * f.x + x1 * f.y));
  }
  return v;
}

// Interpolation at M listed points (one thread each).
__global__ void k_interp_points(InterpTab T, const double* __restrict__ pts, long long M, double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  out[i] = interp_at(T, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

struct EffInterpDev {
  InterpTab tab[EP_MAXT];
  int nrot, ng;
  int gtype[EP_MAXG];
  double beta, nanval;
  int nx, ny, nz;
  double rv[9];
  double inv_n[3];
};

// Orientational average with interpolated site potentials: one thread per
// point, every orientation and site interpolated from the site type's table.
template <bool GRID>
__global__ void __launch_bounds__(EP_NT) k_effpot_interp(EffInterpDev e, const double* __restrict__ off,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ pts, long long M,
                                                        double* __restrict__ V) {
  const int tid = threadIdx.x;
  double p[3];
  long long oi;
  bool valid;
  if constexpr (GRID) {
    const int bx0 = blockIdx.z * EP_BX, by0 = blockIdx.y * EP_BY, bz0 = blockIdx.x * EP_BZ;
    const int ix = bx0 + tid / (EP_BY * EP_BZ);
    const int iy = by0 + (tid / EP_BZ) % EP_BY;
    const int iz = bz0 + tid % EP_BZ;
    valid = (ix < e.nx) && (iy < e.ny) && (iz < e.nz);
    const double f0 = ix * e.inv_n[0] + (-0.5);
    const double f1 = iy * e.inv_n[1] + (-0.5);
    const double f2 = iz * e.inv_n[2] + (-0.5);
#pragma unroll
    for (int b = 0; b < 3; ++b) p[b] = e.rv[b] * f0 + e.rv[3 + b] * f1 + e.rv[6 + b] * f2;
    oi = ((long long)ix * e.ny + iy) * e.nz + iz;
  } else {
    const long long i = (long long)blockIdx.x * EP_NT + tid;
    valid = i < M;
    if (!valid) return;
#pragma unroll
    for (int b = 0; b < 3; ++b) p[b] = pts[3 * i + b];
    oi = i;
  }
  if (!valid) return;
  double mx = -INFINITY, sum = 0.0;
  bool nan = false;
  for (int r = 0; r < e.nrot; ++r) {
    double u = 0.0;
    for (int i = 0; i < e.ng; ++i) {
      const double* o = off + ((long long)r * e.ng + i) * 3;
      u += interp_at(e.tab[e.gtype[i]], p[0] + o[0], p[1] + o[1], p[2] + o[2]);
    }
    ep_fold(-e.beta * u, w[r], mx, sum, nan);
  }
  double v = (sum > 0.0) ? -(mx + log(sum)) / e.beta : INFINITY;
  if (nan || v != v) v = e.nanval;
  V[oi] = v;
}

}  // namespace cmm
