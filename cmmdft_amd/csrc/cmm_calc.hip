// Reductions of one grid field for the post-processing of converged
// densities and potentials (cmmdft/calculator.py: loadings, Henry
// coefficient, helium void fraction, ideal-gas free energy).  Deterministic:
// a fixed grid of blocks with grid-stride loops stores block partials, one
// block sums them in order.
#include "cmm_common.h"

namespace cmm {

constexpr int RED_BLOCKS = 1024;
constexpr int RED_NT = 256;

enum FieldOp {
  FOP_SUM = 0,        // x
  FOP_BOLTZ = 1,      // exp(-beta x)
  FOP_BOLTZ_CLIP = 2, // min(max(exp(-beta x), 0), 1)
  FOP_MIN = 3,
  FOP_MAX = 4,
  FOP_SUMSQ = 5,      // x^2
  FOP_FID = 6,        // c (ln(lam3 c) - 1) with c = max(x, 1e-20): ideal-gas free energy density / kT
  FOP_FID_TRACK = 7,  // the same with c = 1e-30 where x <= 1e-8 (FreeEnergy.track, functionals.py:151-155)
};

__device__ __forceinline__ double fop_value(int op, double x, double beta, double lam3) {
  switch (op) {
    case FOP_BOLTZ: return exp(-beta * x);
    case FOP_BOLTZ_CLIP: return fmin(fmax(exp(-beta * x), 0.0), 1.0);
    case FOP_SUMSQ: return x * x;
    case FOP_FID: {
      const double c = fmax(x, 1e-20);
      return c * (log(c * lam3) - 1.0);
    }
    case FOP_FID_TRACK: {
      const double c = (x <= 1e-8) ? 1e-30 : x;
      return c * (log(lam3 * c) - 1.0);
    }
    default: return x;
  }
}

__global__ void __launch_bounds__(RED_NT) k_field_reduce(const double* __restrict__ x,
                                                        const unsigned char* __restrict__ mask, long long n, int op,
                                                        double beta, double lam3, double* __restrict__ partials) {
  __shared__ double sh[CMM_RED_SH];
  const bool mx = op == FOP_MIN || op == FOP_MAX;
  double acc[1] = {op == FOP_MIN ? INFINITY : (op == FOP_MAX ? -INFINITY : 0.0)};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (mask && !mask[i]) continue;
    const double v = fop_value(op, x[i], beta, lam3);
    if (op == FOP_MIN) acc[0] = fmin(acc[0], v);
    else if (op == FOP_MAX) acc[0] = fmax(acc[0], v);
    else acc[0] += v;
  }
  if (op == FOP_MIN) acc[0] = -acc[0];     // block_reduce's max of the negated values
  const int ops[1] = {mx ? 1 : 0};
  block_reduce<1>(acc, ops, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = (op == FOP_MIN) ? -acc[0] : acc[0];
}

__global__ void __launch_bounds__(RED_NT) k_field_reduce_final(const double* __restrict__ partials, int nb, int op,
                                                              double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = op == FOP_MIN ? INFINITY : (op == FOP_MAX ? -INFINITY : 0.0);
  for (int b = 0; b < nb; ++b) {
    if (op == FOP_MIN) a = fmin(a, partials[b]);
    else if (op == FOP_MAX) a = fmax(a, partials[b]);
    else a += partials[b];
  }
  out[0] = a;
}

}  // namespace cmm

namespace cmm {

// Distance from every grid point to the nearest host atom
// (Program.calc_distance, cmmdft/program.py:145-177): the minimum image of
// the difference vector taken in fractional coordinates, f = G d,
// f -= rint(f), d = f R (yaff's Cell.mic with G the reciprocal vectors as
// rows), then the minimum over the atoms.  Atoms stream through LDS in
// chunks; one thread per point.
constexpr int MD_NT = 256;

struct MinDistDev {
  int nx, ny, nz;
  double rv[9];   // cell vectors as rows
  double gv[9];   // reciprocal vectors as rows (G R^T = I)
  double inv_n[3];
};

__global__ void __launch_bounds__(MD_NT) k_min_dist(const double* __restrict__ pos, int natom, MinDistDev d,
                                                   double* __restrict__ out) {
  __shared__ double sx[MD_NT], sy[MD_NT], sz[MD_NT];
  const long long n = (long long)d.nx * d.ny * d.nz;
  const long long idx = (long long)blockIdx.x * MD_NT + threadIdx.x;
  const long long ii = idx < n ? idx : n - 1;
  const int iz = (int)(ii % d.nz);
  const int iy = (int)((ii / d.nz) % d.ny);
  const int ix = (int)(ii / ((long long)d.ny * d.nz));
  const double f0 = ix * d.inv_n[0] + (-0.5), f1 = iy * d.inv_n[1] + (-0.5), f2 = iz * d.inv_n[2] + (-0.5);
  double p[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) p[b] = d.rv[b] * f0 + d.rv[3 + b] * f1 + d.rv[6 + b] * f2;
  double best = INFINITY;
  for (int c0 = 0; c0 < natom; c0 += MD_NT) {
    const int j = c0 + threadIdx.x;
    if (j < natom) {
      sx[threadIdx.x] = pos[3 * j];
      sy[threadIdx.x] = pos[3 * j + 1];
      sz[threadIdx.x] = pos[3 * j + 2];
    }
    __syncthreads();
    const int cnt = min(MD_NT, natom - c0);
    for (int e = 0; e < cnt; ++e) {
      const double v[3] = {p[0] - sx[e], p[1] - sy[e], p[2] - sz[e]};
      double f[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        f[a] = d.gv[3 * a] * v[0] + d.gv[3 * a + 1] * v[1] + d.gv[3 * a + 2] * v[2];
        f[a] -= rint(f[a]);
      }
      double r2 = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double w = f[0] * d.rv[b] + f[1] * d.rv[3 + b] + f[2] * d.rv[6 + b];
        r2 += w * w;
      }
      best = fmin(best, r2);
    }
    __syncthreads();
  }
  if (idx < n) out[idx] = sqrt(best);
}

}  // namespace cmm
