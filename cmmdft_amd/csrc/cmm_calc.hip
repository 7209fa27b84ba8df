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
