// Mean-field (MFA) attraction: the guest-guest pair potential on the grid and
// the kernel table of its convolution (cmmdft/functionals.py:770-939).
//
// The reference evaluates c_MFA = ifft(fft(rho) fft(U) L) V with its phased
// transforms (system.py:310-323): the phase P = exp(i pi scalprod) of the two
// forward transforms leaves one factor P in the product, so
//     c_MFA = dr Re(IFFT[rho^ U^ L P])          (unnormalised transforms).
// With tail corrections U lives on a supercell of reps_a times the grid and
// the convolution uses every reps_a-th supercell wave vector, i.e. the
// transform of U folded onto the grid, with the supercell's phase at those
// wave vectors.  The real part of the inverse equals the inverse of the
// Hermitian part of the spectrum; rho^ and U^ are Hermitian, so the table
// holds
//     K(k) = (dr / N) U^(k) L(k) (P(k) + conj(P(-k))) / 2,
// and the back-convolution adds K rho^ to its k-space sum.
#include "cmm_kernels.h"

namespace cmm {

struct MfaLJ {
  double sigma, eps4, rmin, limit, cutoff, shift;
  int use_cutoff;
};

struct Cell9 {
  double rv[9];   // rows are the cell vectors
};

// LJ pair potential at the points of an ns0 x ns1 x ns2 grid of the cell c
// (frac_a = i_a / ns_a - 1/2, x = sum_a frac_a rv[a]; system.py:276-282),
// folded onto the grid of n_a = ns_a / reps_a points:
//   out[i] = sum over the image offsets j of U(point(i + j n)).
// U (functionals.py:878-910): 4 eps ((s/r)^12 - (s/r)^6) for r > rmin, the
// limit value otherwise; with a cutoff, minus U(cutoff) for r <= cutoff and 0
// beyond.
__global__ void k_mfa_lj(double* __restrict__ out, int n0, int n1, int n2, int r0, int r1, int r2, Cell9 c, MfaLJ lj) {
  const long long npts = (long long)n0 * n1 * n2;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const int i2 = (int)(i % n2);
  const int i1 = (int)((i / n2) % n1);
  const int i0 = (int)(i / ((long long)n1 * n2));
  const int s0 = n0 * r0, s1 = n1 * r1, s2 = n2 * r2;
  const double d0 = 1.0 / s0, d1 = 1.0 / s1, d2 = 1.0 / s2;
  double acc = 0.0;
  for (int j0 = 0; j0 < r0; ++j0) {
    const double f0 = (double)(i0 + j0 * n0) * d0 + (-0.5);
    for (int j1 = 0; j1 < r1; ++j1) {
      const double f1 = (double)(i1 + j1 * n1) * d1 + (-0.5);
      for (int j2 = 0; j2 < r2; ++j2) {
        const double f2 = (double)(i2 + j2 * n2) * d2 + (-0.5);
        const double x = c.rv[0] * f0 + c.rv[3] * f1 + c.rv[6] * f2;
        const double y = c.rv[1] * f0 + c.rv[4] * f1 + c.rv[7] * f2;
        const double z = c.rv[2] * f0 + c.rv[5] * f1 + c.rv[8] * f2;
        const double r = sqrt(x * x + y * y + z * z);
        double u = lj.limit;
        if (r > lj.rmin) {
          const double q = lj.sigma / r;
          const double q2 = q * q;
          const double q6 = q2 * q2 * q2;
          u = lj.eps4 * (q6 * q6 - q6);
        }
        if (lj.use_cutoff) u = (r <= lj.cutoff) ? u - lj.shift : 0.0;
        acc += u;
      }
    }
  }
  out[i] = acc;
}

// out[i] = sum over the image offsets j of in[i + j n] (in: ns0 x ns1 x ns2).
__global__ void k_fold(const double* __restrict__ in, double* __restrict__ out, int n0, int n1, int n2, int r0, int r1,
                       int r2) {
  const long long npts = (long long)n0 * n1 * n2;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const int i2 = (int)(i % n2);
  const int i1 = (int)((i / n2) % n1);
  const int i0 = (int)(i / ((long long)n1 * n2));
  const long long s1 = (long long)n1 * r1, s2 = (long long)n2 * r2;
  double acc = 0.0;
  for (int j0 = 0; j0 < r0; ++j0)
    for (int j1 = 0; j1 < r1; ++j1)
      for (int j2 = 0; j2 < r2; ++j2)
        acc += in[((long long)(i0 + j0 * n0) * s1 + (i1 + j1 * n1)) * s2 + (i2 + j2 * n2)];
  out[i] = acc;
}

// K over the half spectrum [nx][ny][nz/2+1] from the unnormalised R2C
// transform U^ of the (folded) potential; ph[a] holds the per-axis phase
// factors p_a(m) (P = p_x p_y p_z).  On an axis' Nyquist plane m maps to
// itself under k -> -k, so conj(P(-k)) takes conj(p_a) there and p_a
// elsewhere.
__global__ void k_mfa_table(const double2* __restrict__ uhat, double2* __restrict__ kd, GridDev g,
                            const double2* __restrict__ ph, int lanczos, double scale) {
  const long long nhalf = (long long)g.nx * g.ny * g.nzh;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nhalf) return;
  const int iz = (int)(i % g.nzh);
  const int iy = (int)((i / g.nzh) % g.ny);
  const int ix = (int)(i / ((long long)g.ny * g.nzh));
  const double2 px = ph[ix], py = ph[g.nx + iy], pz = ph[g.nx + g.ny + iz];
  const double2 qx = (ix == g.kx.nyq) ? make_double2(px.x, -px.y) : px;
  const double2 qy = (iy == g.ky.nyq) ? make_double2(py.x, -py.y) : py;
  const double2 qz = (iz == g.kz.nyq) ? make_double2(pz.x, -pz.y) : pz;
  const double2 P = cmul(cmul(px, py), pz);
  const double2 Q = cmul(cmul(qx, qy), qz);
  double s = 0.5 * scale;
  if (lanczos) s *= g.kx.lanczos[ix] * g.ky.lanczos[iy] * g.kz.lanczos[iz];
  kd[i] = cscale(cmul(uhat[i], make_double2(P.x + Q.x, P.y + Q.y)), s);
}

}  // namespace cmm
