// HIP kernels of the cDFT hot path for CDNA4 (gfx950), all FP64.
//
// Data layout in HBM (C order, z fastest, per instance of a batch):
//   real field      double [nx][ny][nz]
//   half spectrum   double2[nx][ny][nz/2+1]   (R2C along z, then complex y, x)
// A 3-D transform is three pencil passes (z, y, x).  The z pass packs two real
// pencils into one complex sequence (two-for-one), so odd lengths need no
// special case.  Every pass stages a tile of pencils in LDS and runs an
// in-place Stockham FFT there; the functional's k-space weights, the
// pointwise derivatives of the FMT/WDA free-energy density and the
// Euler-Lagrange update are fused into the passes that own the data.
#include "cmm_common.h"
#include "cmm_kernels.h"

#include <math.h>

namespace cmm {

constexpr int NT = 256;   // threads per block of the FFT passes
constexpr int VPT = 8;    // complex values per thread per Stockham stage

// ---------------------------------------------------------------------------
// z-pass tiles.  A block owns P consecutive z pencils (two real pencils per
// complex sequence: pair q = pencils p0+2q (real part) and p0+2q+1 (imag
// part)) of nf fields.  LDS layout is element-major with an odd row stride:
// element e of sequence s = f*B + q sits at lds[e*ES + s], B = P/2, ES >= nf*B
// and odd, so both the transposing global<->LDS copies and every Stockham
// stage are free of LDS bank conflicts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double& z_val(double2* lds, const ZTile& t, int f, int pl, int e) {
  return reinterpret_cast<double*>(lds)[2 * (e * t.ES + f * t.B + (pl >> 1)) + (pl & 1)];
}

// Loaders issue the global loads of LU elements per thread before any of
// their LDS stores, with addresses clamped into range and out-of-range values
// replaced afterwards: no branch separates the loads, so they are all in
// flight together.  Offsets inside one field of one instance are 32-bit
// (pencils nx ny <= 65536, nz/2+1 <= 129 elements each).
constexpr int LU = 4;

__device__ __forceinline__ void z_load_real(double2* lds, const double* src, long long p0, const ZTile& t, long long npen,
                                            int n, int f) {
  const int tot = t.P * n;
  const int ip0 = (int)p0, ilast = (int)npen - 1;
  for (int i0 = threadIdx.x; i0 < tot; i0 += LU * blockDim.x) {
    double v[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, tot - 1);
      const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
      v[u] = src[min(ip0 + pl, ilast) * n + e];
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < tot) {
        const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
        z_val(lds, t, f, pl, e) = (p0 + pl < npen) ? v[u] : 0.0;
      }
    }
  }
}

// forward-transformed pairs -> two half spectra
__device__ __forceinline__ void z_store_half(const double2* lds, double2* dst, long long p0, const ZTile& t, long long npen,
                                             int n, int nzh, int f) {
  const int tot = t.B * nzh;
  const int ip0 = (int)p0, inpen = (int)npen;
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int q = (int)fdiv((unsigned)i, t.fnzh), k = i - q * nzh;
    const double2 zk = lds[k * t.ES + f * t.B + q];
    const double2 zm = lds[(k == 0 ? 0 : n - k) * t.ES + f * t.B + q];
    const double2 A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
    const double2 Bv = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
    const int pa = ip0 + 2 * q, o = pa * nzh + k;
    if (pa < inpen) dst[o] = A;
    if (pa + 1 < inpen) dst[o + nzh] = Bv;
  }
}

// two half spectra -> complex sequence whose inverse FFT is (a + i b).
// The k = 0 and k = n/2 coefficients are projected onto their real parts,
// i.e. the Hermitian part is taken, which is what the real part of a full
// complex inverse transform does.
__device__ __forceinline__ void z_load_half(double2* lds, const double2* src, long long p0, const ZTile& t, long long npen,
                                            int n, int nzh, int f) {
  const int tot = t.B * nzh;
  const int ip0 = (int)p0, ilast = (int)npen - 1;
  for (int i0 = threadIdx.x; i0 < tot; i0 += LU * blockDim.x) {
    double2 A[LU], B[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, tot - 1);
      const int q = (int)fdiv((unsigned)i, t.fnzh), k = i - q * nzh;
      const int pa = ip0 + 2 * q;
      A[u] = src[min(pa, ilast) * nzh + k];
      B[u] = src[min(pa + 1, ilast) * nzh + k];
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < tot) {
        const int q = (int)fdiv((unsigned)i, t.fnzh), k = i - q * nzh;
        const int pa = ip0 + 2 * q;
        double2 a = (pa <= ilast) ? A[u] : make_double2(0.0, 0.0);
        double2 b = (pa + 1 <= ilast) ? B[u] : make_double2(0.0, 0.0);
        const bool selfconj = (k == 0) || (2 * k == n);
        if (selfconj) { a.y = 0.0; b.y = 0.0; }
        lds[k * t.ES + f * t.B + q] = make_double2(a.x - b.y, a.y + b.x);
        if (!selfconj) lds[(n - k) * t.ES + f * t.B + q] = make_double2(a.x + b.y, b.x - a.y);
      }
    }
  }
}

// kz of a z-spectrum index (table kzk, Nyquist index kzn), zero on the Nyquist plane
__device__ __forceinline__ double kz_of(const double* kzk, int kzn, int k) { return (k == kzn) ? 0.0 : kzk[k]; }

// nf half-spectrum fields (field stride fs) into the tile in one loop.
// VZ (six-field functional): field 4, the z component of n_v2, is not stored;
// it is -i kz times field 1's (n3 = w3 rho^) spectrum, zero on the Nyquist
// plane, formed here from field 1 per pencil.
template <bool VZ = false, int NU = LU>
__device__ __forceinline__ void z_load_half_multi(double2* lds, const double2* src, long long fs, int nf, long long p0,
                                                  const ZTile& t, long long npen, int n, int nzh, const FastDiv& fBn,
                                                  const double* kzk = nullptr, int kzn = -1) {
  const int per = t.B * nzh;
  const int tot = nf * per;
  const int ip0 = (int)p0, ilast = (int)npen - 1;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NU * blockDim.x) {
    double2 A[NU], B[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, tot - 1);
      const int f = (int)fdiv((unsigned)i, fBn);
      const int r = i - f * per;
      const int q = (int)fdiv((unsigned)r, t.fnzh), k = r - q * nzh;
      const int pa = ip0 + 2 * q;
      const double2* sf = src + ((VZ && f == 4) ? 1 : f) * fs;
      A[u] = sf[min(pa, ilast) * nzh + k];
      B[u] = sf[min(pa + 1, ilast) * nzh + k];
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < tot) {
        const int f = (int)fdiv((unsigned)i, fBn);
        const int r = i - f * per;
        const int q = (int)fdiv((unsigned)r, t.fnzh), k = r - q * nzh;
        const int pa = ip0 + 2 * q;
        double2 a = (pa <= ilast) ? A[u] : make_double2(0.0, 0.0);
        double2 b = (pa + 1 <= ilast) ? B[u] : make_double2(0.0, 0.0);
        if (VZ && f == 4) {
          const double kk = kz_of(kzk, kzn, k);
          a = cmul_i(cscale(a, kk), -1.0);
          b = cmul_i(cscale(b, kk), -1.0);
        }
        const bool selfconj = (k == 0) || (2 * k == n);
        if (selfconj) { a.y = 0.0; b.y = 0.0; }
        lds[k * t.ES + f * t.B + q] = make_double2(a.x - b.y, a.y + b.x);
        if (!selfconj) lds[(n - k) * t.ES + f * t.B + q] = make_double2(a.x + b.y, b.x - a.y);
      }
    }
  }
}

// The same fields of a runtime-shaped tile by columns: one thread per (pencil
// pair q, kz k) column holding that column of all NF fields, the index
// arithmetic once per column and field 4 (VZ) formed from field 1's loaded
// values; FULL: every pencil of the tile exists.  Per element the operations
// of z_load_half_multi.  NF <= 6 (a column of twelve fields would hold 48
// complex values in registers).
template <bool VZ, int NF, bool FULL>
__device__ __forceinline__ void z_load_cols(double2* lds, const double2* src, long long fs, long long p0,
                                            const ZTile& t, long long npen, int n, int nzh, const double* kzk,
                                            int kzn) {
  static_assert(NF <= 6, "column loads hold every field of a column in registers");
  const int ncol = t.B * nzh;
  const int ip0 = (int)p0, ilast = (int)npen - 1;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    const int q = (int)fdiv((unsigned)c, t.fnzh), k = c - q * nzh;
    const int pa = ip0 + 2 * q;
    const int oa = (FULL ? pa : min(pa, ilast)) * nzh + k;
    const int ob = (FULL ? pa + 1 : min(pa + 1, ilast)) * nzh + k;
    double2 A[NF], B[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (VZ && f == 4) continue;
      A[f] = src[f * fs + oa];
      B[f] = src[f * fs + ob];
    }
    const double kk = VZ ? kz_of(kzk, kzn, k) : 0.0;
    const bool selfconj = (k == 0) || (2 * k == n);
    double2 hi[NF];   // the values of kz slot n - k, stored after the loop
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int fl = (VZ && f == 4) ? 1 : f;
      double2 a = (FULL || pa <= ilast) ? A[fl] : make_double2(0.0, 0.0);
      double2 b = (FULL || pa + 1 <= ilast) ? B[fl] : make_double2(0.0, 0.0);
      if (VZ && f == 4) {
        a = cmul_i(cscale(a, kk), -1.0);
        b = cmul_i(cscale(b, kk), -1.0);
      }
      if (selfconj) { a.y = 0.0; b.y = 0.0; }
      lds[k * t.ES + f * t.B + q] = make_double2(a.x - b.y, a.y + b.x);
      hi[f] = make_double2(a.x + b.y, b.x - a.y);
    }
    if (!selfconj) {
#pragma unroll
      for (int f = 0; f < NF; ++f) lds[(n - k) * t.ES + f * t.B + q] = hi[f];
    }
  }
}

// the two half spectra of pair q of field f at index k of a transformed tile
__device__ __forceinline__ void z_unpack(const double2* lds, const ZTile& t, int n, int f, int q, int k, double2& A,
                                         double2& Bv) {
  const double2 zk = lds[k * t.ES + f * t.B + q];
  const double2 zm = lds[(k == 0 ? 0 : n - k) * t.ES + f * t.B + q];
  A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
  Bv = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
}

// VZ (six-field functional): the derivative field of the z component of n_v2
// is not stored; its back-convolution term i kz w3 G_vz joins field 1's
// (w3 G_n3) as G_n3 + i kz G_vz, formed here per pencil.
template <bool VZ = false>
__device__ __forceinline__ void z_store_half_multi(const double2* lds, double2* dst, long long fs, int nf, long long p0,
                                                   const ZTile& t, long long npen, int n, int nzh, const FastDiv& fBn,
                                                   const double* kzk = nullptr, int kzn = -1) {
  const int per = t.B * nzh;
  const int tot = (VZ ? nf - 1 : nf) * per;
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int fs0 = (int)fdiv((unsigned)i, fBn);
    const int r = i - fs0 * per;
    const int f = (VZ && fs0 >= 4) ? fs0 + 1 : fs0;
    const int q = (int)fdiv((unsigned)r, t.fnzh), k = r - q * nzh;
    double2 A, Bv;
    z_unpack(lds, t, n, f, q, k, A, Bv);
    if (VZ && f == 1) {
      double2 A4, B4;
      z_unpack(lds, t, n, 4, q, k, A4, B4);
      const double kk = kz_of(kzk, kzn, k);
      A = cadd(A, cmul_i(cscale(A4, kk), 1.0));
      Bv = cadd(Bv, cmul_i(cscale(B4, kk), 1.0));
    }
    const int pa = (int)p0 + 2 * q, o = pa * nzh + k;
    double2* df = dst + f * fs;
    if (pa < (int)npen) df[o] = A;
    if (pa + 1 < (int)npen) df[o + nzh] = Bv;
  }
}

// The transformed tile's fields by columns (see z_load_cols): per element
// the operations of z_store_half_multi.
template <bool VZ, int NF, bool FULL>
__device__ __forceinline__ void z_store_cols(const double2* lds, double2* dst, long long fs, long long p0,
                                             const ZTile& t, long long npen, int n, int nzh, const double* kzk,
                                             int kzn) {
  static_assert(NF <= 6, "column stores are built for six fields at most");
  const int ncol = t.B * nzh;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    const int q = (int)fdiv((unsigned)c, t.fnzh), k = c - q * nzh;
    const int pa = (int)p0 + 2 * q, o = pa * nzh + k;
    const double kk = VZ ? kz_of(kzk, kzn, k) : 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (VZ && f == 4) continue;
      double2 A, Bv;
      z_unpack(lds, t, n, f, q, k, A, Bv);
      if (VZ && f == 1) {
        double2 A4, B4;
        z_unpack(lds, t, n, 4, q, k, A4, B4);
        A = cadd(A, cmul_i(cscale(A4, kk), 1.0));
        Bv = cadd(Bv, cmul_i(cscale(B4, kk), 1.0));
      }
      double2* df = dst + f * fs;
      if (FULL || pa < (int)npen) df[o] = A;
      if (FULL || pa + 1 < (int)npen) df[o + nzh] = Bv;
    }
  }
}

// A z tile's NF fields in and out: by columns for up to six fields (the
// full-tile form on every tile but the grid's last), else per element.
template <bool VZ, int NF>
__device__ __forceinline__ void z_load_tile(double2* lds, const double2* src, long long fs, long long p0,
                                            const ZTile& t, long long npen, int n, int nzh, const FastDiv& fBn,
                                            const double* kzk, int kzn) {
  if constexpr (NF <= 6) {
    if (p0 + t.P <= npen) z_load_cols<VZ, NF, true>(lds, src, fs, p0, t, npen, n, nzh, kzk, kzn);
    else z_load_cols<VZ, NF, false>(lds, src, fs, p0, t, npen, n, nzh, kzk, kzn);
  } else {
    z_load_half_multi<VZ>(lds, src, fs, NF, p0, t, npen, n, nzh, fBn, kzk, kzn);
  }
}

template <bool VZ, int NF>
__device__ __forceinline__ void z_store_tile(const double2* lds, double2* dst, long long fs, long long p0,
                                             const ZTile& t, long long npen, int n, int nzh, const FastDiv& fBn,
                                             const double* kzk, int kzn) {
  if constexpr (NF <= 6) {
    if (p0 + t.P <= npen) z_store_cols<VZ, NF, true>(lds, dst, fs, p0, t, npen, n, nzh, kzk, kzn);
    else z_store_cols<VZ, NF, false>(lds, dst, fs, p0, t, npen, n, nzh, kzk, kzn);
  } else {
    z_store_half_multi<VZ>(lds, dst, fs, NF, p0, t, npen, n, nzh, fBn, kzk, kzn);
  }
}

// all nf*B sequences of a z tile, in place (R16: the two-stage 16 x 8 form
// of length 128, for the LDS-bound functional z pass)
template <int SGN, bool R16 = false>
__device__ __forceinline__ void z_fft(double2* lds, const ZTile& t, const Fft1D& fz, const double2* tws) {
  lds_fft<SGN, true, VPT, R16>(lds, t.S, t.fS, 1, t.ES, fz, tws);
}

// ---------------------------------------------------------------------------
// The same z tile with its shape fixed at compile time (P pencils, NF fields,
// length N): the loaders, the storer and the transforms below do the
// operations of their runtime-tile counterparts above in the same order, with
// LDS offsets that are immediates and index splits by constant divisions.
// Blocks of NT threads.
// ---------------------------------------------------------------------------
template <int P_, int NF_, int N_>
struct ZTileC {
  static constexpr int P = P_, NF = NF_, N = N_;
  static constexpr int B = P_ / 2, S = NF_ * (P_ / 2), ES = (NF_ * (P_ / 2)) | 1, NZH = N_ / 2 + 1;
};

template <class TC>
__device__ __forceinline__ double& z_val_c(double2* lds, int f, int pl, int e) {
  return reinterpret_cast<double*>(lds)[2 * (e * TC::ES + f * TC::B + (pl >> 1)) + (pl & 1)];
}

template <class TC>
__device__ __forceinline__ void z_unpack_c(const double2* lds, int f, int q, int k, double2& A, double2& Bv) {
  const double2 zk = lds[k * TC::ES + f * TC::B + q];
  const double2 zm = lds[(k == 0 ? 0 : TC::N - k) * TC::ES + f * TC::B + q];
  A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
  Bv = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
}

// The tile's half-spectrum fields by columns: one thread per (pencil pair q,
// kz k) column, holding that column of every field, so the index arithmetic
// is done once per column rather than once per field, and field 4 (VZ: the z
// component of n_v2, -i kz times field 1's spectrum, zero on the Nyquist
// plane) is formed from field 1's loaded values instead of loading them
// again.  FULL: every pencil of the tile exists (all tiles but the grid's
// last), so no address is clamped and no value zero-filled.  Per element the
// same operations as the runtime tile's z_load_half_multi / z_store_half_multi.
template <bool VZ, class TC, bool FULL>
__device__ __forceinline__ void z_load_cols_c(double2* lds, const double2* src, long long fs, long long p0,
                                              long long npen, const double* kzk, int kzn) {
  constexpr int n = TC::N, nzh = TC::NZH, NF = TC::NF;
  constexpr int ncol = TC::B * nzh;
  const int ip0 = (int)p0, ilast = (int)npen - 1;
  for (int c = threadIdx.x; c < ncol; c += NT) {
    const int q = (int)((unsigned)c / (unsigned)nzh), k = c - q * nzh;
    const int pa = ip0 + 2 * q;
    const int oa = (FULL ? pa : min(pa, ilast)) * nzh + k;
    const int ob = (FULL ? pa + 1 : min(pa + 1, ilast)) * nzh + k;
    double2 A[NF], B[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (VZ && f == 4) continue;
      A[f] = src[f * fs + oa];
      B[f] = src[f * fs + ob];
    }
    const double kk = VZ ? kz_of(kzk, kzn, k) : 0.0;
    const bool selfconj = (k & (n / 2 - 1)) == 0;   // k = 0 or n/2 (k <= n/2)
    double2 hi[NF];   // the values of kz slot n - k, stored after the loop
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int fl = (VZ && f == 4) ? 1 : f;
      double2 a = (FULL || pa <= ilast) ? A[fl] : make_double2(0.0, 0.0);
      double2 b = (FULL || pa + 1 <= ilast) ? B[fl] : make_double2(0.0, 0.0);
      if (VZ && f == 4) {
        a = cmul_i(cscale(a, kk), -1.0);
        b = cmul_i(cscale(b, kk), -1.0);
      }
      if (selfconj) { a.y = 0.0; b.y = 0.0; }
      lds[k * TC::ES + f * TC::B + q] = make_double2(a.x - b.y, a.y + b.x);
      hi[f] = make_double2(a.x + b.y, b.x - a.y);
    }
    if (!selfconj) {
#pragma unroll
      for (int f = 0; f < NF; ++f) lds[(n - k) * TC::ES + f * TC::B + q] = hi[f];
    }
  }
}

template <bool VZ, class TC, bool FULL>
__device__ __forceinline__ void z_store_cols_c(const double2* lds, double2* dst, long long fs, long long p0,
                                               long long npen, const double* kzk, int kzn) {
  constexpr int nzh = TC::NZH, NF = TC::NF;
  constexpr int ncol = TC::B * nzh;
  for (int c = threadIdx.x; c < ncol; c += NT) {
    const int q = (int)((unsigned)c / (unsigned)nzh), k = c - q * nzh;
    const int pa = (int)p0 + 2 * q, o = pa * nzh + k;
    const double kk = VZ ? kz_of(kzk, kzn, k) : 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (VZ && f == 4) continue;
      double2 A, Bv;
      z_unpack_c<TC>(lds, f, q, k, A, Bv);
      if (VZ && f == 1) {
        double2 A4, B4;
        z_unpack_c<TC>(lds, 4, q, k, A4, B4);
        A = cadd(A, cmul_i(cscale(A4, kk), 1.0));
        Bv = cadd(Bv, cmul_i(cscale(B4, kk), 1.0));
      }
      double2* df = dst + f * fs;
      if (FULL || pa < (int)npen) df[o] = A;
      if (FULL || pa + 1 < (int)npen) df[o + nzh] = Bv;
    }
  }
}

// all S sequences of the tile, length 128 as radix 16 then 8 (z_fft<SGN, true>)
template <int SGN, class TC>
__device__ __forceinline__ void z_fft_c(double2* lds, const double2* tws) {
  static_assert(TC::N == 128, "compile-time z tiles are built for length 128");
  lds_fft_t<SGN, 128, TC::S, TC::ES, NT, 1, 16, 8>(lds, tws);
}

// ---------------------------------------------------------------------------
// Plain transform passes (3-D FFT entry points and building blocks)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_zr2c(const double* __restrict__ x, long long xs, double2* __restrict__ X, long long Xs,
                                            GridDev g, ZTile t, const int* __restrict__ active) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  z_load_real(sm, x + inst * xs, p0, t, npen, g.nz, 0);
  __syncthreads();
  z_fft<-1>(sm, t, g.fz, s_tw);
  z_store_half(sm, X + inst * Xs, p0, t, npen, g.nz, g.nzh, 0);
}

// field blockIdx.y of a batch of nf-field spectra -> real field
__global__ void __launch_bounds__(NT) k_zc2r(const double2* __restrict__ X, long long Xs, long long Xfs, double* __restrict__ x,
                                            long long xs, long long xfs, GridDev g, ZTile t, const int* __restrict__ active) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  const int inst = blockIdx.z;
  const int f = blockIdx.y;
  if (active && !active[inst]) return;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  z_load_half(sm, X + inst * Xs + f * Xfs, p0, t, npen, g.nz, g.nzh, 0);
  __syncthreads();
  z_fft<1>(sm, t, g.fz, s_tw);
  double* out = x + inst * xs + f * xfs;
  const int tot = t.P * g.nz;
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * g.nz;
    const long long p = p0 + pl;
    if (p < npen) out[p * g.nz + e] = z_val(sm, t, 0, pl, e);
  }
}

// Tile order of the y and x passes, XCD-aware: the hardware deals
// consecutive workgroups to the 8 XCDs in turn, so workgroup w takes tile
// (w % 8) * per + w / 8 (the grid is padded to 8 * per): runs of consecutive
// tiles, which share the cache lines their column ranges straddle, stay on
// one XCD's L2.
__device__ __forceinline__ long long xcd_tile(int per) {
  return (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// y pass: one block per tile of TB consecutive kz columns of one x plane (all
// y) of one field; tiles t = x * ntz + kz-tile.
// fmap != 0: the block row blockIdx.y works on field (fmap >> 8 blockIdx.y) &
// 0xff (the chain's travelling fields), else on field blockIdx.y.
// PAIR (six-field functionals): the block of field 1 (n3) also does field 3
// (the y component of n_v2), which is i ky' times field 1 on either side of
// the y transform (ky' = ky, zero on its Nyquist plane): for SGN = +1 (the
// forward chain) n_v2,y = -i ky' n3 is formed from n3's tile and transformed
// and stored beside it; for SGN = -1 (the back chain) both derivative tiles
// are transformed and G_n3 + i ky' G_vy is stored in field 1 (field 3 is not
// stored).  The x passes then move neither, and the values are those they
// formed on the same elements.
template <int SGN, bool PAIR = false>
__global__ void __launch_bounds__(NT) k_ypass(double2* __restrict__ F, long long is, long long fs, GridDev g, int TB,
                                             FastDiv fTB, int ntz, int per, const int* __restrict__ active,
                                             unsigned fmap) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fy);
  const int inst = blockIdx.z;
  if (active && !active[inst]) return;
  const long long t = xcd_tile(per);
  if (t >= (long long)g.nx * ntz) return;
  const int x = (int)(t / ntz);
  const int k0 = (int)(t - (long long)x * ntz) * TB;
  const int tbv = min(TB, g.nzh - k0);
  const int fld = fmap ? (int)((fmap >> (8 * blockIdx.y)) & 0xffu) : (int)blockIdx.y;
  double2* base = F + inst * is + fld * fs + (long long)x * g.ny * g.nzh + k0;
  const int tot = g.ny * TB;
  if (PAIR && fld == 1) {
    // two rounds of cw <= ceil(TB / 2) columns: the columns' field-1 and
    // field-3 sequences side by side in LDS (2 cw sequences, element-major;
    // the launch sizes LDS for 2 ceil(TB / 2) columns), transformed together;
    // no value is held in registers across a transform
    double2* base3 = base + 2 * fs;
    const int c = (TB + 1) >> 1;
    for (int j0 = 0; j0 < TB; j0 += c) {
      const int cw = min(c, TB - j0), ns = 2 * cw;
      const FastDiv fC = make_fastdiv((unsigned)cw), fS = make_fastdiv((unsigned)ns);
      const int tot2 = g.ny * cw;
      for (int i = threadIdx.x; i < tot2; i += blockDim.x) {
        const int y = (int)fdiv((unsigned)i, fC), j = i - y * cw;
        const bool in = j0 + j < tbv;
        const long long o = (long long)y * g.nzh + min(j0 + j, tbv - 1);
        const double2 v = base[o];
        double2 w;
        if (SGN > 0) {
          const double kyp = (y == g.ky.nyq) ? 0.0 : g.ky.k[y];
          w = cmul_i(cscale(v, kyp), -1.0);
        } else {
          w = base3[o];
        }
        sm[y * ns + j] = make_double2(in ? v.x : 0.0, in ? v.y : 0.0);
        sm[y * ns + cw + j] = make_double2(in ? w.x : 0.0, in ? w.y : 0.0);
      }
      __syncthreads();
      lds_fft<SGN, true, VPT>(sm, ns, fS, 1, ns, g.fy, s_tw);
      for (int i = threadIdx.x; i < tot2; i += blockDim.x) {
        const int y = (int)fdiv((unsigned)i, fC), j = i - y * cw;
        if (j0 + j < tbv) {
          const long long o = (long long)y * g.nzh + j0 + j;
          if (SGN > 0) {
            base[o] = sm[y * ns + j];
            base3[o] = sm[y * ns + cw + j];
          } else {
            const double kyp = (y == g.ky.nyq) ? 0.0 : g.ky.k[y];
            base[o] = cadd(sm[y * ns + j], cmul_i(cscale(sm[y * ns + cw + j], kyp), 1.0));
          }
        }
      }
      __syncthreads();
    }
    return;
  }
  for (int i0 = threadIdx.x; i0 < tot; i0 += LU * blockDim.x) {

    double2 v[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, tot - 1);
      const int y = (int)fdiv((unsigned)i, fTB), b = i - y * TB;
      v[u] = base[(long long)y * g.nzh + min(b, tbv - 1)];
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < tot) {
        const int y = (int)fdiv((unsigned)i, fTB), b = i - y * TB;
        sm[i] = (b < tbv) ? v[u] : make_double2(0.0, 0.0);
      }
    }
  }
  __syncthreads();
  lds_fft<SGN, true, VPT>(sm, TB, fTB, 1, TB, g.fy, s_tw);
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int y = (int)fdiv((unsigned)i, fTB), b = i - y * TB;
    if (b < tbv) base[(long long)y * g.nzh + b] = sm[i];
  }
}

// k_ypass at ny = 128 with the radix sequence 8, 4, 4 and the tile of TB
// columns fixed at compile time: the same operations in the same order as
// k_ypass with that sequence and tile, so the same results bit for bit.
// LDS offsets and index splits are constants, and the kernel holds no
// generic transform, so the paired block (PAIR, field 1: see k_ypass) can
// keep its tile's values in registers across a transform within the
// five-waves-per-SIMD budget that LDS sets: for SGN = +1 it transforms
// -i ky' n3 from the loaded values, then n3; for SGN = -1 it transforms
// G_n3, keeps the result, transforms G_vy (loaded with G_n3) and stores
// G_n3 + i ky' G_vy.
template <int SGN, bool PAIR, int TB>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PAIR ? 5 : 1))) k_ypass128(double2* __restrict__ F, long long is, long long fs, GridDev g,
                                                int ntz, int per, const int* __restrict__ active, unsigned fmap) {
  constexpr int TOT = 128 * TB;
  constexpr int E = (TOT + NT - 1) / NT;   // tile elements per thread: tid + k NT, k < E
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[128];
  load_twiddles(s_tw, g.fy);
  const int inst = blockIdx.z;
  if (active && !active[inst]) return;
  const long long t = xcd_tile(per);
  if (t >= (long long)g.nx * ntz) return;
  const int x = (int)(t / ntz);
  const int k0 = (int)(t - (long long)x * ntz) * TB;
  const int tbv = min(TB, g.nzh - k0);
  const int fld = fmap ? (int)((fmap >> (8 * blockIdx.y)) & 0xffu) : (int)blockIdx.y;
  double2* base = F + inst * is + fld * fs + (long long)x * 128 * g.nzh + k0;
  double2* base3 = base + 2 * fs;
  const bool pair = PAIR && fld == 1;
  double2 v[E], w[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = min((int)threadIdx.x + k * NT, TOT - 1);
    const int y = i / TB, b = i - y * TB;
    v[k] = base[(long long)y * g.nzh + min(b, tbv - 1)];
  }
  if (PAIR && SGN < 0 && pair) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = min((int)threadIdx.x + k * NT, TOT - 1);
      const int y = i / TB, b = i - y * TB;
      w[k] = base3[(long long)y * g.nzh + min(b, tbv - 1)];
    }
  }
  if (PAIR && SGN > 0 && pair) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < TOT) {
        const int y = i / TB, b = i - y * TB;
        const double kyp = (y == g.ky.nyq) ? 0.0 : g.ky.k[y];
        const double2 u = cmul_i(cscale(v[k], kyp), -1.0);
        sm[i] = make_double2(b < tbv ? u.x : 0.0, b < tbv ? u.y : 0.0);
      }
    }
    __syncthreads();
    lds_fft_t<SGN, 128, TB, TB, NT, 1, 8, 4, 4>(sm, s_tw);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < TOT) {
        const int y = i / TB, b = i - y * TB;
        if (b < tbv) base3[(long long)y * g.nzh + b] = sm[i];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = (int)threadIdx.x + k * NT;
    if (i < TOT) {
      const int b = i % TB;
      sm[i] = make_double2(b < tbv ? v[k].x : 0.0, b < tbv ? v[k].y : 0.0);
    }
  }
  __syncthreads();
  lds_fft_t<SGN, 128, TB, TB, NT, 1, 8, 4, 4>(sm, s_tw);
  if (PAIR && SGN < 0 && pair) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      v[k] = (i < TOT) ? sm[i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < TOT) {
        const int b = i % TB;
        sm[i] = make_double2(b < tbv ? w[k].x : 0.0, b < tbv ? w[k].y : 0.0);
      }
    }
    __syncthreads();
    lds_fft_t<SGN, 128, TB, TB, NT, 1, 8, 4, 4>(sm, s_tw);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < TOT) {
        const int y = i / TB, b = i - y * TB;
        const double kyp = (y == g.ky.nyq) ? 0.0 : g.ky.k[y];
        if (b < tbv) base[(long long)y * g.nzh + b] = cadd(v[k], cmul_i(cscale(sm[i], kyp), 1.0));
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = (int)threadIdx.x + k * NT;
    if (i < TOT) {
      const int y = i / TB, b = i - y * TB;
      if (b < tbv) base[(long long)y * g.nzh + b] = sm[i];
    }
  }
}

// x pass: tile = TB consecutive (ky, kz) columns of the flattened plane, all x.
template <int SGN>
__global__ void __launch_bounds__(NT) k_xpass(double2* __restrict__ F, long long is, long long fs, GridDev g, int TB,
                                             FastDiv fTB, const int* __restrict__ active) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fx);
  const int inst = blockIdx.z;
  if (active && !active[inst]) return;
  const long long nb = (long long)g.ny * g.nzh;
  const long long b0 = (long long)blockIdx.x * TB;
  const int tbv = (int)min((long long)TB, nb - b0);
  double2* base = F + inst * is + blockIdx.y * fs + b0;
  const int tot = g.nx * TB;
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int x = (int)fdiv((unsigned)i, fTB), b = i - x * TB;
    sm[i] = (b < tbv) ? base[(long long)x * nb + b] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<SGN, true, VPT>(sm, TB, fTB, 1, TB, g.fx, s_tw);
#pragma unroll 4
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int x = (int)fdiv((unsigned)i, fTB), b = i - x * TB;
    if (b < tbv) base[(long long)x * nb + b] = sm[i];
  }
}

// ---------------------------------------------------------------------------
// Analytic k-space weights (FMT: Rosenfeld weights with the Lanczos sigma
// factor; WDA: normalised ball of diameter D).  For one k point, returns
// w2 (= 4 pi R^2 sinc(kR)), w3 (= 4 pi R^3/3 * 3(sin-w cos)/w^3), wD and the
// k components.  omega <= 1e-8 uses the k -> 0 limits.
// ---------------------------------------------------------------------------
struct KW {
  double w2, w3, wD;
  double wt;  // tensor weight amplitude B = -4 pi R^2 (j_ball - j0) (Lanczos included) = w2 - (3/R) w3
  bool k0;    // omega = |k| R <= 1e-8: the k -> 0 limits apply, vector and tensor weights vanish
};

__device__ __forceinline__ double j_ball(double w, double s, double c) {
  return 3.0 * (s - w * c) / (w * w * w);
}

__device__ __forceinline__ KW kweights(const GridDev& g, const FuncDev& fp, int ix, int iy, int iz) {
  KW r;
  const double kx = g.kx.k[ix], ky = g.ky.k[iy], kz = g.kz.k[iz];
  const double sl = g.kx.lanczos[ix] * g.ky.lanczos[iy] * g.kz.lanczos[iz];
  const double kn = sqrt(kx * kx + ky * ky + kz * kz);
  const double w = kn * fp.R;
  double s, c;
  r.k0 = (w <= 1e-8);
  if (r.k0) {
    r.w2 = fp.c_w2 * sl;
    r.w3 = fp.c_w3 * sl;
  } else {
    sincos(w, &s, &c);
    r.w2 = fp.c_w2 * (s / w) * sl;
    r.w3 = fp.c_w3 * j_ball(w, s, c) * sl;
  }
  const double wd = kn * fp.D;
  if (wd <= 1e-8) {
    r.wD = sl;
  } else if (fp.D == 2.0 * fp.R && !r.k0) {
    // WDA diameter D = 2R (the hard-sphere radius): wd = 2w exactly, and the
    // double-angle identities reuse sin w, cos w
    const double s2 = 2.0 * s * c, c2 = (c - s) * (c + s);
    r.wD = j_ball(wd, s2, c2) * sl;
  } else {
    sincos(wd, &s, &c);
    r.wD = j_ball(wd, s, c) * sl;
  }
  r.wt = r.w2 - fp.c_tor * r.w3;
  return r;
}

// The weights w2, w3, w_D of the half-spectrum elements with x <= nx/2
// (flat index x (ny nzh) + y nzh + z) into out[3][n]: the table the x passes
// read.  Elements x and nx - x have the same weights bit for bit (their kx
// differ in sign only, and the Lanczos factor is even), so the passes read
// row min(x, nx - x), and a block, which holds every x of its columns, reads
// each row twice at nearly the same time: the second read is served by the
// cache instead of HBM.
__global__ void __launch_bounds__(NT) k_kw_table(GridDev g, FuncDev fp, double* __restrict__ out) {
  const long long nb = (long long)g.ny * g.nzh;
  const long long n = (long long)(g.nx / 2 + 1) * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i / nb);
    const long long bb = i - x * nb;
    const int y = (int)(bb / g.nzh), z = (int)(bb - (long long)y * g.nzh);
    const KW w = kweights(g, fp, x, y, z);
    out[i] = w.w2;
    out[n + i] = w.w3;
    out[2 * n + i] = w.wD;
  }
}

// Weight of field f at grid indices (ix, iy, iz).  Forward (DIR = -1): the
// weights that produce n2, n3, n_v2 (x,y,z) and the WDA density from rho^.
// Back (DIR = +1): the weights that convolve the derivative fields (vector
// weights change sign).  Each vector component is zero on the Nyquist plane of
// its own axis: there its spectrum has no Hermitian part, so the real part of
// the inverse transform does not see it.  At k = 0 the vector weights vanish.
//
// Tensor fields 6..11 (xx, xy, xz, yy, yz, zz; the 't' variant): real, even
// weights B (k_i k_j / k^2 - delta_ij / 3), zero at k = 0, the same in both
// directions.  An off-diagonal component is odd in k_i and in k_j, so where
// exactly one of the two indices is on its axis' Nyquist plane its spectrum
// has no Hermitian part and it is zero there (on both planes it is even).
__device__ __forceinline__ double2 tensor_w(const GridDev& g, const KW& w, int ix, int iy, int iz, int c, double2 v) {
  if (w.k0) return make_double2(0.0, 0.0);
  const double k[3] = {g.kx.k[ix], g.ky.k[iy], g.kz.k[iz]};
  const bool nq[3] = {ix == g.kx.nyq, iy == g.ky.nyq, iz == g.kz.nyq};
  const int a = (c < 3) ? 0 : (c < 5 ? 1 : 2);
  const int b = (c < 3) ? c : (c < 5 ? c - 2 : 2);
  if (a != b && nq[a] != nq[b]) return make_double2(0.0, 0.0);
  const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  const double h = k[a] * k[b] / k2 - (a == b ? 1.0 / 3.0 : 0.0);
  return cscale(v, w.wt * h);
}

__device__ __forceinline__ double2 apply_w(const GridDev& g, const KW& w, int ix, int iy, int iz, int f, double2 v,
                                           int dir) {
  const bool k0 = w.k0;
  switch (f) {
    case 0: return cscale(v, w.w2);
    case 1: return cscale(v, w.w3);
    case 2: return (k0 || ix == g.kx.nyq) ? make_double2(0.0, 0.0) : cmul_i(cscale(v, w.w3 * g.kx.k[ix]), (double)dir);
    case 3: return (k0 || iy == g.ky.nyq) ? make_double2(0.0, 0.0) : cmul_i(cscale(v, w.w3 * g.ky.k[iy]), (double)dir);
    case 4: return (k0 || iz == g.kz.nyq) ? make_double2(0.0, 0.0) : cmul_i(cscale(v, w.w3 * g.kz.k[iz]), (double)dir);
    case 5: return cscale(v, w.wD);
    default: return tensor_w(g, w, ix, iy, iz, f - 6, v);
  }
}

// ---------------------------------------------------------------------------
// x passes of the functional chain.  A block (NTX threads) owns a tile of TB
// consecutive (ky, kz) columns of the flattened (y, z) plane, all nx, with
// nx * TB <= NTX, and holds all NF fields of the tile in LDS at once
// (element-major, [x][NF TB]), so that its global loads are all in flight
// together and one Stockham pass transforms the NF TB sequences.
// Tiles are taken in the XCD-aware order of xcd_tile.
// ---------------------------------------------------------------------------
constexpr int NTX = 512;

// Six-field functionals (NF = 6) weight and transform four fields only: on a
// column of the x pass ky and kz are constant, so the y and z components of
// n_v2, i ky w3 rho^ and i kz w3 rho^ (zero on their axis' Nyquist plane),
// are i ky' and i kz' times the n3 column before and after the x transform.
// ky is the y pass's own frequency and kz the z pass's: the paired y pass
// forms the y component from n3 and folds i ky' G_vy into G_n3 (k_ypass,
// PAIR), the z pass forms the z component and folds i kz' G_vz
// (z_load_half_multi), so neither is stored by the x passes (vz = 1 stores
// both, for callers that read the fields).
// Slots 0..3 hold the fields n2, n3, nv2_x, rho_bar (weights 0, 1, 2, 5).
// NF = 5: the six-field functional whose WDA is linear (every EOS term mean
// field, f_ex = A rho_bar^2): rho_bar is not formed; its energy
// A sum rho_bar^2 = (A/N) sum_k |w_D rho^|^2 and its derivative term
// w_D * (2A rho_bar) = IFFT(2A w_D^2 rho^) are taken in k space by the x
// passes, so slots 0..2 hold n2, n3, nv2_x and four fields travel through
// the z pass instead of five.
//
// TBC > 0: the tile fixed at compile time -- nx = 128 with the radix sequence
// 8, 4, 4 and TB = TBC columns, in blocks of 128 TBC threads -- so LDS
// offsets are immediates and index splits constant divisions; the same
// operations in the same order as the runtime tile (TBC = 0) with that radix
// sequence and TB, so the same results bit for bit.
template <int NF>
struct XFields {
  static constexpr bool red = (NF == 6 || NF == 5);
  static constexpr bool wlin = (NF == 5);
  static constexpr int nc = NF == 6 ? 4 : (NF == 5 ? 3 : NF);   // fields weighted and transformed
  __device__ static constexpr int field(int c) { return (NF == 6 && c == 3) ? 5 : c; }
};

// Tile shape of an x pass: runtime (TBC = 0) or compile-time (TBC columns
// of nx = 128).
template <int TBC>
struct XTile {
  static constexpr bool CT = TBC > 0;
  int TB, nx;
  FastDiv fTB;
  __device__ XTile(int tb, const FastDiv& f, int n) : TB(CT ? TBC : tb), nx(CT ? 128 : n), fTB(f) {}
  __device__ __forceinline__ int div(int i) const {
    if constexpr (CT) return (int)((unsigned)i / (unsigned)TBC);
    else return (int)fdiv((unsigned)i, fTB);
  }
};

// in-LDS x transform of `batch` sequences (NS of them, element stride ES):
// the compile-time stages for TBC > 0, else the plan's radix sequence
template <int SGN, int TBC, int NS, int ES>
__device__ __forceinline__ void x_fft(double2* buf, int batch, const FastDiv& fb, int es, const GridDev& g,
                                      const double2* tws) {
  if constexpr (TBC > 0) lds_fft_t<SGN, 128, NS, ES, 128 * TBC, 1, 8, 4, 4>(buf, tws);
  else lds_fft<SGN, true, VPT>(buf, batch, fb, 1, es, g.fx, tws);
}

// Weights of x-pass element (x, wy, wz), flat half-spectrum index idx: from
// the plan's table (filled by k_kw_table with kweights' values) when there is
// one, else computed.  The k = 0 flag and wt are formed as in kweights.
__device__ __forceinline__ KW x_kw(const GridDev& g, const FuncDev& fp, int x, int wy, int wz, long long bb) {
  if (fp.kwt) {
    const long long idx = (long long)min(x, g.nx - x) * ((long long)g.ny * g.nzh) + bb;
    KW w;
    w.w2 = fp.kwt[idx];
    w.w3 = fp.kwt[fp.kwt_n + idx];
    w.wD = fp.kwt[2 * fp.kwt_n + idx];
    const double kx = g.kx.k[x], ky = g.ky.k[wy], kz = g.kz.k[wz];
    w.k0 = (sqrt(kx * kx + ky * ky + kz * kz) * fp.R <= 1e-8);
    w.wt = w.w2 - fp.c_tor * w.w3;
    return w;
  }
  return kweights(g, fp, x, wy, wz);
}

// ky and kz of a column, zero on their axis' Nyquist plane
__device__ __forceinline__ void column_k(const GridDev& g, int wy, int wz, double& kyp, double& kzp) {
  kyp = (wy == g.ky.nyq) ? 0.0 : g.ky.k[wy];
  kzp = (wz == g.kz.nyq) ? 0.0 : g.kz.k[wz];
}

// LDS row stride (in complex values) of a compile-time x tile of n values
// per row: rows 256 B apart would put all the rows a wave touches on the
// same half of the LDS banks, so a multiple of 8 is padded by 4 (64 B).
__host__ __device__ constexpr int x_row(int n) { return (n % 8 == 0) ? n + 4 : n; }

// forward chain: rho (z,y transformed) -> x FFT -> weighted copies ->
// inverse x FFT of all of them -> F[0..NF).  The 1/N of the inverse 3-D
// transform is folded into the weights.  NF = 6 (fields n2, n3, nv2 x/y/z,
// rho_bar) or 12 (plus the six tensor components).
// MFA: also the mean-field energy (1/2) sum_x rho c_MFA = (1/2) sum_k
// |rho^(k)|^2 K(k) over the full spectrum (K Hermitian, so the sum is real):
// on the half spectrum the planes kz = 0 and kz = Nyquist count once, the
// others twice.  NF = 5 (linear WDA): also sum_x f_ex(rho_bar) =
// (2A/N) sum_half |rho^|^2 w_D^2 with the same plane weights.  Partials, one
// per wave (quantities: the MFA energy, then the WDA energy), are indexed by
// tile (ntiles of them, wave_reduce_store).
template <int NF, bool MFA, int OCC, int TBC = 0>
__global__ void __launch_bounds__(NTX) __attribute__((amdgpu_waves_per_eu(OCC)))
k_xweights(const double2* __restrict__ rz, long long rzs, double2* __restrict__ F, long long Fs, long long Ffs,
           GridDev g, FuncDev fp, int TBr, FastDiv fTB, FastDiv fS6, FastDiv fnzh, int per, long long ntiles,
           const int* __restrict__ active, RedDev red, int vz) {
  using XF = XFields<NF>;
  constexpr int NC = XF::nc;
  const XTile<TBC> xt(TBr, fTB, g.nx);
  const int TB = xt.TB;
  extern __shared__ double2 sm[];    // [nx][SR] fields, then the nx twiddles
  constexpr int SRC = TBC > 0 ? x_row(NC * TBC) : 0;
  const int SR = TBC > 0 ? SRC : NC * TB;   // LDS row stride
  double2* s_tw = sm + xt.nx * SR;
  load_twiddles(s_tw, g.fx);
  const int inst = blockIdx.z;
  if (active && !active[inst]) return;
  const long long tile = xcd_tile(per);
  if (tile >= ntiles) return;
  const int T = xt.nx * TB;
  const int S6 = NC * TB;
  double2* B = sm;                    // [nx][SR]: NC fields of TB columns per row
  double2* A = sm + (NC - 1) * TB;    // rho^ in the slot of the last field: [x * SR + b]
  const long long nb = (long long)g.ny * g.nzh;
  const long long b0 = tile * TB;
  const int tbv = (int)min((long long)TB, nb - b0);
  const double2* src = rz + inst * rzs + b0;
  // compile-time tile: one element per thread, whose weights are read from
  // the table with the tile's loads (the transform then hides their
  // latency), instead of after the forward transform
  KW wpre{};
  if constexpr (TBC > 0) {
    const int x = xt.div(threadIdx.x), b = (int)threadIdx.x - x * TB;
    const unsigned bb = (unsigned)min(b0 + b, nb - 1);
    const int wy = (int)fdiv(bb, fnzh);
    wpre = x_kw(g, fp, x, wy, (int)bb - wy * g.nzh, bb);
  }
  for (int i0 = threadIdx.x; i0 < T; i0 += LU * blockDim.x) {
    double2 v[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, T - 1);
      const int x = xt.div(i), b = i - x * TB;
      v[u] = src[(long long)x * nb + min(b, tbv - 1)];
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < T) {
        const int x = xt.div(i), b = i - x * TB;
        A[x * SR + b] = (b < tbv) ? v[u] : make_double2(0.0, 0.0);
      }
    }
  }
  __syncthreads();
  x_fft<-1, TBC, TBC, SRC>(A, TB, fTB, SR, g, s_tw);
  constexpr bool E = MFA || XF::wlin;   // energies reduced here
  double emfa = 0.0, ewl = 0.0;
  // each element's weighted copies overwrite only that element's own slots
  // (its rho^ among them), and one thread reads and writes them
  for (int i = threadIdx.x; i < T; i += blockDim.x) {
    const int x = xt.div(i), b = i - x * TB;
    const unsigned bb = (unsigned)min(b0 + b, nb - 1);
    const int wy = (int)fdiv(bb, fnzh);
    const int wz = (int)bb - wy * g.nzh;
    const KW w = TBC > 0 ? wpre : x_kw(g, fp, x, wy, wz, bb);
    const double2 a = A[x * SR + b];
    if (E && b < tbv) {
      const double half = (wz == 0 || wz == g.kz.nyq) ? 0.5 : 1.0;
      const double a2 = half * (a.x * a.x + a.y * a.y);
      if (MFA) emfa += a2 * fp.kmfa[(long long)x * nb + bb].x;
      if (XF::wlin) ewl += a2 * (w.wD * w.wD);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      B[x * SR + c * TB + b] = cscale(apply_w(g, w, x, wy, wz, XF::field(c), a, -1), g.inv_npts);
  }
  if (E) {
    const double acc[2] = {emfa, ewl * fp.wl2 * g.inv_npts};
    const int ops[2] = {0, 0};
    wave_reduce_store<2>(acc, ops, red.partials + (long long)inst * red.pstride, (int)ntiles, (int)tile);
  }
  __syncthreads();
  x_fft<1, TBC, NC * TBC, SRC>(B, S6, fS6, SR, g, s_tw);
  if constexpr (XF::red) {
    // one element's six fields per thread; consecutive threads write
    // consecutive columns of a row of each field
    double2* dst = F + inst * Fs + b0;
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
      const int x = xt.div(i), b = i - x * TB;
      if (b >= tbv) continue;
      const unsigned bb = (unsigned)(b0 + b);
      const int wy = (int)fdiv(bb, fnzh);
      const int wz = (int)bb - wy * g.nzh;
      const long long o = (long long)x * nb + b;
      const double2 n3 = B[x * SR + TB + b];
      dst[o] = B[x * SR + b];
      dst[Ffs + o] = n3;
      dst[2 * Ffs + o] = B[x * SR + 2 * TB + b];
      if (vz) {
        double kyp, kzp;
        column_k(g, wy, wz, kyp, kzp);
        dst[3 * Ffs + o] = cmul_i(cscale(n3, kyp), -1.0);
        dst[4 * Ffs + o] = cmul_i(cscale(n3, kzp), -1.0);
      }
      if constexpr (!XF::wlin) dst[5 * Ffs + o] = B[x * SR + 3 * TB + b];
    }
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double2* dst = F + inst * Fs + f * Ffs + b0;
      for (int i = threadIdx.x; i < T; i += blockDim.x) {
        const int x = xt.div(i), b = i - x * TB;
        if (b < tbv) dst[(long long)x * nb + b] = B[x * SR + f * TB + b];
      }
    }
  }
}

// back-convolution: the NF derivative fields -> x FFT of all of them ->
// weighted sum -> inverse x FFT -> F[0] (in place: the block reads its whole
// tile of every field before it writes field 0's tile).
// MFA: the density's spectrum (rz, z and y transformed) is loaded as one more
// field and transformed with the others; K rho^ joins the weighted sum.
// NF = 5 (linear WDA): rho^ is loaded likewise and (2A/N) w_D^2 rho^ joins
// the sum in place of the back-convolved WDA derivative field.
template <int NF, bool MFA, int OCC, int TBC = 0>
__global__ void __launch_bounds__(NTX) __attribute__((amdgpu_waves_per_eu(OCC)))
k_xback(double2* __restrict__ F, long long Fs, long long Ffs, GridDev g, FuncDev fp, int TBr, FastDiv fTB,
        FastDiv fS6, FastDiv fT, FastDiv fnzh, int per, long long ntiles, const int* __restrict__ active,
        const double2* __restrict__ rz, long long rzs) {
  using XF = XFields<NF>;
  constexpr int NC = XF::nc;
  constexpr bool LR = MFA || XF::wlin;      // rho^ joins the fields
  constexpr int NL = NC + (LR ? 1 : 0);     // fields held in LDS
  const XTile<TBC> xt(TBr, fTB, g.nx);
  const int TB = xt.TB;
  extern __shared__ double2 sm[];    // [nx][SR] fields, then the nx twiddles
  constexpr int SRC = TBC > 0 ? x_row(NL * TBC) : 0;
  const int SR = TBC > 0 ? SRC : NL * TB;   // LDS row stride
  double2* s_tw = sm + xt.nx * SR;
  load_twiddles(s_tw, g.fx);
  const int inst = blockIdx.z;
  if (active && !active[inst]) return;
  const long long tile = xcd_tile(per);
  if (tile >= ntiles) return;
  const int T = xt.nx * TB;
  const int S6 = NL * TB;
  double2* B = sm;    // [nx][SR]: NL fields of TB columns per row; the weighted sum goes to field 0's slot
  const long long nb = (long long)g.ny * g.nzh;
  const long long b0 = tile * TB;
  const int tbv = (int)min((long long)TB, nb - b0);
  const double2* src = F + inst * Fs + b0;
  const double2* rsrc = LR ? rz + inst * rzs + b0 : nullptr;
  if constexpr (XF::red) {
    // one element's fields (and rho^) per thread, loads in flight together;
    // field 1 holds G_n3 + i ky' G_vy + i kz' G_vz (the paired y pass and
    // the z pass have folded the y and z components of G_v into it)
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
      const int x = xt.div(i), b = i - x * TB;
      const long long o = (long long)x * nb + min(b, tbv - 1);
      const double2 v0 = src[o], h = src[Ffs + o], v2 = src[2 * Ffs + o];
      const double2 v5 = XF::wlin ? make_double2(0.0, 0.0) : src[5 * Ffs + o];
      const double2 r = LR ? rsrc[o] : make_double2(0.0, 0.0);
      // columns past the tile's end hold zeros (selected per component: a
      // select of whole double2 values is lowered through scratch memory)
      const bool in = b < tbv;
      B[x * SR + b] = make_double2(in ? v0.x : 0.0, in ? v0.y : 0.0);
      B[x * SR + TB + b] = make_double2(in ? h.x : 0.0, in ? h.y : 0.0);
      B[x * SR + 2 * TB + b] = make_double2(in ? v2.x : 0.0, in ? v2.y : 0.0);
      if constexpr (!XF::wlin) B[x * SR + 3 * TB + b] = make_double2(in ? v5.x : 0.0, in ? v5.y : 0.0);
      if (LR) B[x * SR + NC * TB + b] = make_double2(in ? r.x : 0.0, in ? r.y : 0.0);
    }
  } else {
    for (int i0 = threadIdx.x; i0 < NL * T; i0 += LU * blockDim.x) {
      double2 v[LU];
#pragma unroll
      for (int u = 0; u < LU; ++u) {
        const int i = min(i0 + u * (int)blockDim.x, NL * T - 1);
        const int f = (int)fdiv((unsigned)i, fT), r = i - f * T;
        const int x = (int)fdiv((unsigned)r, fTB), b = r - x * TB;
        const long long o = (long long)x * nb + min(b, tbv - 1);
        v[u] = (MFA && f == NF) ? rsrc[o] : src[f * Ffs + o];
      }
#pragma unroll
      for (int u = 0; u < LU; ++u) {
        const int i = i0 + u * (int)blockDim.x;
        if (i < NL * T) {
          const int f = (int)fdiv((unsigned)i, fT), r = i - f * T;
          const int x = (int)fdiv((unsigned)r, fTB), b = r - x * TB;
          B[x * SR + f * TB + b] = (b < tbv) ? v[u] : make_double2(0.0, 0.0);
        }
      }
    }
  }
  __syncthreads();
  x_fft<-1, TBC, NL * TBC, SRC>(B, S6, fS6, SR, g, s_tw);
  for (int i = threadIdx.x; i < T; i += blockDim.x) {
    const int x = xt.div(i), b = i - x * TB;
    const unsigned bb = (unsigned)min(b0 + b, nb - 1);
    const int wy = (int)fdiv(bb, fnzh);
    const int wz = (int)bb - wy * g.nzh;
    const KW w = x_kw(g, fp, x, wy, wz, bb);
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      acc = cadd(acc, cscale(apply_w(g, w, x, wy, wz, XF::field(c), B[x * SR + c * TB + b], 1), g.inv_npts));
    if (LR) {
      const double2 r = B[x * SR + NC * TB + b];
      if (MFA) acc = cadd(acc, cmul(r, fp.kmfa[(long long)x * nb + bb]));
      if (XF::wlin) acc = cadd(acc, cscale(r, fp.wl2 * g.inv_npts * (w.wD * w.wD)));
    }
    B[x * SR + b] = acc;    // this element's slots are read above by this thread only
  }
  __syncthreads();
  x_fft<1, TBC, TBC, SRC>(B, TB, fTB, SR, g, s_tw);
  double2* dst = F + inst * Fs + b0;
  for (int i = threadIdx.x; i < T; i += blockDim.x) {
    const int x = xt.div(i), b = i - x * TB;
    if (b < tbv) dst[(long long)x * nb + b] = B[x * SR + b];
  }
}

// ---------------------------------------------------------------------------
// Pointwise free-energy density: FMT Phi and its partial derivatives
// (Rosenfeld 'FMT', 'MFMT', White-Bear II 'WBII', optional 'a' variant), and
// the WDA excess free energy per volume of a sum of EOS terms.
// ---------------------------------------------------------------------------
// Forward-mode dual number (value, directional derivative).  The pointwise
// functions below are templates over the scalar type: with double they are the
// functional itself, with Dual they also give the derivative of every output
// along a direction of the weighted densities, i.e. a row of the Jacobian of
// the derivative fields (the Hessian of Phi) applied to that direction.
// Branches and clips select on the value, so the derivative is that of the
// branch the value takes, and a clipped input has zero derivative.
using ::exp;
using ::log;
struct Dual {
  double v, d;
  __device__ __forceinline__ Dual() : v(0.0), d(0.0) {}
  __device__ __forceinline__ Dual(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
__device__ __forceinline__ Dual log(Dual a) { return Dual(::log(a.v), a.d / a.v); }
__device__ __forceinline__ Dual exp(Dual a) {
  const double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
// min(max(x, lo), hi) with fmin/fmax semantics (NaN -> lo)
__device__ __forceinline__ double clip(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
__device__ __forceinline__ Dual clip(Dual x, double lo, double hi) {
  if (x.v >= lo && x.v <= hi) return x;
  return Dual(x.v > hi ? hi : lo, 0.0);
}

template <typename T>
struct PhiOut {
  T phi, g0, g1, gv[3];
  T gt[6];   // tensor derivative fields (the 't' variant)
};

template <typename T>
__device__ __forceinline__ void phi23(int base, bool closed, T n3, T ln3, T r1, T in1, T& p2, T& p3, T& d2,
                                      T& d3) {
  // Same forms as the reference (functionals.py:726-767), whose series
  // branches hold for n3 <= 1e-8 (closed = false; n3_series > 1e-8 keeps the
  // closed forms' cancellation at low n3 out).  Every denominator is a
  // product of powers of (1 - n3) and n3, so the divisions are
  // multiplications by r1 = 1/(1-n3) and (the closed modified-FMT and
  // White-Bear II forms) in1 = 1/n3, and the series' constant divisors
  // multiplications by their reciprocals: q = (1-n3)^2, c = (1-n3)^3,
  // n3 - n3^2 = n3 (1-n3).
  const T n3_2 = n3 * n3, n3_3 = n3_2 * n3;
  const T q = 1.0 - 2.0 * n3 + n3_2;
  const T c = 1.0 - 3.0 * n3 + 3.0 * n3_2 - n3_3;
  const T iq = r1 * r1, ic = iq * r1;
  constexpr double i24 = 1.0 / (24.0 * M_PI), i36 = 1.0 / (36.0 * M_PI);
  constexpr double r9 = 1.0 / 9.0, r18 = 1.0 / 18.0;
  if (base == 0) {                       // Rosenfeld
    p2 = r1;
    d2 = iq;
    p3 = i24 * iq;
    d3 = (1.0 / (12.0 * M_PI)) * ic;
  } else if (base == 1) {                // modified FMT
    p2 = r1;
    d2 = iq;
    if (!closed) {
      p3 = (1.0 - (2.0 * r9) * n3 - r18 * n3_2) * i24 * iq;
      d3 = (8.0 / 3.0 - 0.5 * n3 - 0.1 * n3_2) * i36 * ic;
    } else {
      const T in2 = in1 * in1;
      p3 = (n3 + q * ln3) * i36 * in2 * iq;
      d3 = -(2.0 * n3 - 5.0 * n3_2 + n3_3 + 2.0 * c * ln3) * i36 * (in2 * in1) * ic;
    }
  } else {                               // White-Bear II
    if (!closed) {
      p2 = (1.0 + r9 * n3_2) * r1;
      d2 = (1.0 + (2.0 * r9) * n3 + r18 * n3_2) * iq;
      p3 = (1.0 - (4.0 * r9) * n3 + r18 * n3_2) * i24 * iq;
      d3 = (7.0 / 3.0 - 0.5 * n3 + 0.1 * n3_2) * i36 * ic;
    } else {
      const T in2 = in1 * in1;
      p2 = (5.0 * n3 - n3_2 + 2.0 * (1.0 - n3) * ln3) * ((1.0 / 3.0) * in1 * r1);
      d2 = -2.0 * (n3 - 3.0 * n3_2 + q * ln3) * (1.0 / 3.0) * in2 * iq;
      p3 = -2.0 * (n3 - 3.0 * n3_2 + n3_3 + ln3 * q) * (1.0 / 3.0) * in2 * i24 * iq;
      d3 = (2.0 * n3 - 5.0 * n3_2 + 6.0 * n3_3 - n3_2 * n3_2 + 2.0 * c * ln3) * i36 * (in2 * in1) * ic;
    }
  }
}

// n2, n3 (unclipped), nv2: weighted densities of rho.  Outputs Phi and the
// combined derivative fields (times kT) that the back-convolution weights:
//   g0 = kT (dPhi/dn0 / 4piR^2 + dPhi/dn1 / 4piR + dPhi/dn2)    with w2
//   g1 = kT dPhi/dn3                                           with w3
//   gv = kT (dPhi/dnv1 / 4piR + dPhi/dnv2)                      with -wv2
//   gt = kT dPhi/dn_t (TEN: the 't' variant, tensor tn = xx xy xz yy yz zz;
//        off-diagonal derivatives counted twice, as the symmetric tensor
//        holds each of them twice)                             with w_t
// The 't' terms (functionals.py:636-640, 673-717) add 9/2 (nv2 . T . nv2 -
// tr T^3) to the coefficient of phi3.
// n0, n1 and n_v1 are n2 / 4piR^2, n2 / 4piR and n_v2 / 4piR (their weights
// are those multiples of w2 and w_v2, functionals.py:440-451), so
//   n1 n2 - n_v1 . n_v2 = (n2^2 - |n_v2|^2) / 4piR,
//   g0 = kT (-ln(1-n3) / 4piR^2 + 2 n2 phi2 / 4piR + dq3/dn2 phi3),
//   gv = kT n_v2 (fv - 2 phi2 / 4piR) (+ the tensor terms),
// where fv n_v2 = dq3/dn_v2 phi3.
template <typename T, bool TEN = false>
__device__ __forceinline__ PhiOut<T> fmt_point(const FuncDev& fp, T n2, T n3raw, T vx, T vy, T vz,
                                               const T* tn = nullptr) {
  PhiOut<T> o;
  const double i4pR2 = fp.inv4piR2, i4pR = fp.inv4piR;
  const T n0 = n2 * i4pR2;
  const T n3 = clip(n3raw, 1e-30, 0.99);
  const T om = 1.0 - n3;
  const T ln3 = log(om);
  // the closed modified-FMT / White-Bear II forms need 1/n3 as well: one
  // division gives both reciprocals there
  const bool closed = fp.base != 0 && val(n3) > fp.n3_series;
  T r1, in1;
  if (closed) {
    const T t = 1.0 / (n3 * om);
    r1 = n3 * t;
    in1 = om * t;
  } else {
    r1 = 1.0 / om;
    in1 = r1;
  }
  T p2, p3, d2, d3;
  phi23<T>(fp.base, closed, n3, ln3, r1, in1, p2, p3, d2, d3);
  const T v2 = vx * vx + vy * vy + vz * vz;
  const T s12 = (n2 * n2 - v2) * i4pR;
  T q3, dq3_n2, fv;
  if (fp.anti) {
    const T xi = clip(v2 / (n2 * n2 + 1e-16), 0.0, 1.0);
    const T om3 = 1.0 - xi;
    q3 = n2 * n2 * n2 * om3 * om3 * om3;
    dq3_n2 = 3.0 * n2 * n2 * (1.0 + xi) * om3 * om3;
    fv = -6.0 * n2 * om3 * om3 * p3;
  } else {
    q3 = n2 * n2 * n2 - 3.0 * n2 * v2;
    dq3_n2 = 3.0 * (n2 * n2 - v2);
    fv = -6.0 * n2 * p3;
  }
  if constexpr (TEN) {
    const T xx = tn[0], xy = tn[1], xz = tn[2], yy = tn[3], yz = tn[4], zz = tn[5];
    const T quad = xx * vx * vx + yy * vy * vy + zz * vz * vz + 2.0 * xy * vx * vy + 2.0 * xz * vx * vz +
                   2.0 * yz * vy * vz;
    const T tr3 = xx * xx * xx + yy * yy * yy + zz * zz * zz +
                  3.0 * (xx * xy * xy + xx * xz * xz + yy * xy * xy + yy * yz * yz + zz * xz * xz + zz * yz * yz) +
                  6.0 * xy * xz * yz;
    q3 = q3 + 4.5 * quad;
    q3 = q3 - 4.5 * tr3;
  }
  o.phi = n0 * (-ln3) + s12 * p2 + q3 * p3;
  const double kT = fp.kT;
  o.g0 = kT * (dq3_n2 * p3 + (2.0 * i4pR) * (n2 * p2) - i4pR2 * ln3);
  o.g1 = kT * (n0 * r1 + s12 * d2 + q3 * d3);
  const T cv = fv - (2.0 * i4pR) * p2;
  if constexpr (TEN) {
    const T xx = tn[0], xy = tn[1], xz = tn[2], yy = tn[3], yz = tn[4], zz = tn[5];
    const T p3t = 4.5 * p3;
    o.gv[0] = kT * (vx * cv + p3t * (2.0 * (xx * vx + xy * vy + xz * vz)));
    o.gv[1] = kT * (vy * cv + p3t * (2.0 * (xy * vx + yy * vy + yz * vz)));
    o.gv[2] = kT * (vz * cv + p3t * (2.0 * (xz * vx + yz * vy + zz * vz)));
    o.gt[0] = kT * (p3t * (vx * vx - 3.0 * (xx * xx + xy * xy + xz * xz)));
    o.gt[1] = kT * (p3t * ((vx * vy - 3.0 * (xx * xy + yy * xy + xz * yz)) * 2.0));
    o.gt[2] = kT * (p3t * ((vx * vz - 3.0 * (xx * xz + zz * xz + xy * yz)) * 2.0));
    o.gt[3] = kT * (p3t * (vy * vy - 3.0 * (yy * yy + xy * xy + yz * yz)));
    o.gt[4] = kT * (p3t * ((vy * vz - 3.0 * (yy * yz + zz * yz + xy * xz)) * 2.0));
    o.gt[5] = kT * (p3t * (vz * vz - 3.0 * (zz * zz + xz * xz + yz * yz)));
  } else {
    o.gv[0] = kT * (vx * cv);
    o.gv[1] = kT * (vy * cv);
    o.gv[2] = kT * (vz * cv);
  }
  return o;
}

// EOS term: excess free energy per volume f(rho) and df/drho.
template <typename T>
__device__ __forceinline__ void eos_term(int kind, const double* p, T rho, T& f, T& df) {
  if (kind == 0) {            // mean field: a_ex = a rho
    const double a = p[0];
    f = rho * (a * rho);
    df = rho * a + a * rho;
  } else if (kind == 1) {     // Carnahan-Starling, p = {eta (=4 pi R^3/3), kT}
    const double eta = p[0], kT = p[1];
    const T x = eta * rho;
    const T om = 1.0 - x;
    const T ap = kT * (4.0 * x - 3.0 * x * x) / (om * om);
    const T dap = 2.0 * kT * eta * (2.0 - x) / (om * om * om);
    f = rho * ap;
    df = rho * dap + ap;
  } else {                    // MBWR (Johnson et al. 1993): p = {a1..a8, b1..b6, sigma^3, eps}
    const double s3 = p[14], eps = p[15], gam = 3.0;
    const T rr = rho * s3;
    const T F = exp(-gam * rr * rr);
    const double ig = 1.0 / (2.0 * gam);
    T A = 0.0, dA = 0.0, pw = 1.0;
    for (int i = 0; i < 8; ++i) {      // sum a_i/(i+1) rr^(i+1) ; d: a_i rr^i s3
      dA += p[i] * pw * s3;
      pw *= rr;
      A += p[i] / (i + 1) * pw;
    }
    T G = ig * (1.0 - F);
    A += p[8] * G;
    T rr2 = rr * rr, pw2 = rr2;
    for (int t = 1; t < 6; ++t) {      // G_{t+1} = -ig (F rr^(2t) - 2t G_t)
      G = -ig * (F * pw2 - 2.0 * t * G);
      A += p[8 + t] * G;
      pw2 *= rr2;
    }
    T po = rr;                   // rr^(2t+1)
    for (int t = 0; t < 6; ++t) {
      dA += p[8 + t] * s3 * po * F;
      po *= rr2;
    }
    const T ap = A * eps, dap = dA * eps;
    f = rho * ap;
    df = rho * dap + ap;
  }
}

template <typename T>
__device__ __forceinline__ void wda_point(const FuncDev& fp, T rb, T& f, T& df) {
  f = 0.0;
  df = 0.0;
  for (int t = 0; t < fp.neos; ++t) {
    T ft, dft;
    eos_term<T>(fp.eos_kind[t], fp.eos_par[t], rb, ft, dft);
    f += fp.eos_fac[t] * ft;
    df += fp.eos_fac[t] * dft;
  }
}

// Pointwise stage of the functional z pass at one point: r holds the NF
// weighted densities (n3 unclipped) on entry and the NF derivative fields on
// exit; acc gets Phi, f_ex(rho_bar) and n3 (for the max).  NF = 5: the
// six-field functional with a linear WDA (taken in k space by the x passes),
// without rho_bar.
template <int NF>
__device__ __forceinline__ void zphi_point(const FuncDev& fp, double (&r)[NF], double (&acc)[3]) {
  double gv0 = 0.0, gv1 = 0.0, gv2 = 0.0, g0 = 0.0, g1 = 0.0, g5 = 0.0;
  double gt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (fp.use_hs) {
    PhiOut<double> o;
    if constexpr (NF == 12) {
      o = fmt_point<double, true>(fp, r[0], r[1], r[2], r[3], r[4], &r[6]);
#pragma unroll
      for (int c = 0; c < 6; ++c) gt[c] = o.gt[c];
    } else {
      o = fmt_point<double>(fp, r[0], r[1], r[2], r[3], r[4]);
    }
    acc[0] = o.phi;
    acc[2] = r[1];
    g0 = o.g0;
    g1 = o.g1;
    gv0 = o.gv[0];
    gv1 = o.gv[1];
    gv2 = o.gv[2];
  }
  if constexpr (NF != 5) {
    if (fp.use_wda) {
      double fw, dfw;
      wda_point<double>(fp, r[5], fw, dfw);
      acc[1] = fw;
      g5 = dfw;
    }
  }
  r[0] = g0;
  r[1] = g1;
  r[2] = gv0;
  r[3] = gv1;
  r[4] = gv2;
  if constexpr (NF != 5) r[5] = g5;
  if constexpr (NF == 12) {
#pragma unroll
    for (int c = 0; c < 6; ++c) r[6 + c] = gt[c];
  }
}

// z pass of the functional: NF weighted densities (x,y inverse-transformed)
// -> C2R -> pointwise Phi / WDA derivatives -> R2C -> F[0..NF) in place
// (NF = 12: the tensor variant's six more fields; NF = 5: the six-field
// functional with its linear WDA left to the x passes).
// Reduces sum Phi, sum f_ex(rho_bar), max n3 (unclipped) per instance.
template <int NF>
__global__ void __launch_bounds__(NT) k_zphi(double2* __restrict__ F, long long Fs, long long Ffs, GridDev g, FuncDev fp,
                                            ZTile t, int store, const int* __restrict__ active, RedDev red) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const int n = g.nz;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  const FastDiv fBn = make_fastdiv((unsigned)(t.B * g.nzh));
  constexpr bool VZ = (NF == 5 || NF == 6);   // n_v2's z component formed from n3 here
  z_load_tile<VZ, NF>(sm, F + inst * Fs, Ffs, p0, t, npen, n, g.nzh, fBn, g.kz.k, g.kz.nyq);
  __syncthreads();
  z_fft<1, true>(sm, t, g.fz, s_tw);
  double acc[3] = {0.0, 0.0, -INFINITY};
  const int tot = t.P * n;
  // pencil index fastest: consecutive lanes walk along an LDS row (no bank
  // conflicts); this phase touches no global memory
  const FastDiv fP = make_fastdiv((unsigned)t.P);
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int e = (int)fdiv((unsigned)i, fP), pl = i - e * t.P;
    if (p0 + pl >= npen) continue;
    double r[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) r[f] = z_val(sm, t, f, pl, e);
    double a[3] = {0.0, 0.0, -INFINITY};
    zphi_point<NF>(fp, r, a);
    acc[0] += a[0];
    acc[1] += a[1];
    acc[2] = fmax(acc[2], a[2]);
#pragma unroll
    for (int f = 0; f < NF; ++f) z_val(sm, t, f, pl, e) = r[f];
  }
  if (store) {
    __syncthreads();
    z_fft<-1, true>(sm, t, g.fz, s_tw);
    z_store_tile<VZ, NF>(sm, F + inst * Fs, Ffs, p0, t, npen, n, g.nzh, fBn, g.kz.k, g.kz.nyq);
  }
  const int ops[3] = {0, 0, 1};
  wave_reduce_store<3>(acc, ops, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// k_zphi for nz = 128 with the tile fixed at compile time (P pencils): the
// same operations per element (the results agree with k_zphi's to rounding:
// the compiler contracts the kz products differently).  Its LDS addressing
// and index arithmetic are constants, which cuts the integer work that
// otherwise issues beside the FP64 work.
template <int NF, int P>
__global__ void __launch_bounds__(NT) k_zphi128(double2* __restrict__ F, long long Fs, long long Ffs, GridDev g,
                                               FuncDev fp, int store, const int* __restrict__ active, RedDev red) {
  using TC = ZTileC<P, NF, 128>;
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[128];
  load_twiddles(s_tw, g.fz);
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * P;
  constexpr bool VZ = (NF == 5 || NF == 6);
  const bool full = p0 + P <= npen;   // every pencil of the tile exists
  if (full) z_load_cols_c<VZ, TC, true>(sm, F + inst * Fs, Ffs, p0, npen, g.kz.k, g.kz.nyq);
  else z_load_cols_c<VZ, TC, false>(sm, F + inst * Fs, Ffs, p0, npen, g.kz.k, g.kz.nyq);
  __syncthreads();
  z_fft_c<1, TC>(sm, s_tw);
  double acc[3] = {0.0, 0.0, -INFINITY};
  constexpr int tot = P * 128;
  for (int i = threadIdx.x; i < tot; i += NT) {
    const int e = (int)((unsigned)i / (unsigned)P), pl = i - e * P;
    if (p0 + pl >= npen) continue;
    double r[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) r[f] = z_val_c<TC>(sm, f, pl, e);
    double a[3] = {0.0, 0.0, -INFINITY};
    zphi_point<NF>(fp, r, a);
    acc[0] += a[0];
    acc[1] += a[1];
    acc[2] = fmax(acc[2], a[2]);
#pragma unroll
    for (int f = 0; f < NF; ++f) z_val_c<TC>(sm, f, pl, e) = r[f];
  }
  if (store) {
    __syncthreads();
    z_fft_c<-1, TC>(sm, s_tw);
    if (full) z_store_cols_c<VZ, TC, true>(sm, F + inst * Fs, Ffs, p0, npen, g.kz.k, g.kz.nyq);
    else z_store_cols_c<VZ, TC, false>(sm, F + inst * Fs, Ffs, p0, npen, g.kz.k, g.kz.nyq);
  }
  const int ops[3] = {0, 0, 1};
  wave_reduce_store<3>(acc, ops, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// Pointwise free-energy densities of given weighted densities: kT Phi from
// the NF fields nfld[nf][n] of cmm_weighted_densities (n2, n3 unclipped, nv2
// x/y/z, rho_bar, then the tensor components), and the WDA terms f_ex(rb),
// f_ex'(rb) of a weighted density rb[n].  The local values of
// HardSphereFunctional.value and WDAVFunctional.value / derive with a given
// weighted density (functionals.py:611-618, 1160-1184).
template <int NF>
__global__ void __launch_bounds__(NT) k_local(const double* __restrict__ nfld, long long fs, const double* __restrict__ rb,
                                             long long n, FuncDev fp, double* __restrict__ phi, double* __restrict__ f,
                                             double* __restrict__ df) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (nfld && phi) {
      double r[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) r[c] = nfld[c * fs + i];
      PhiOut<double> o;
      if constexpr (NF == 12) o = fmt_point<double, true>(fp, r[0], r[1], r[2], r[3], r[4], &r[6]);
      else o = fmt_point<double>(fp, r[0], r[1], r[2], r[3], r[4]);
      phi[i] = fp.kT * o.phi;
    }
    if (rb && (f || df)) {
      double fv, dfv;
      wda_point<double>(fp, rb[i], fv, dfv);
      if (f) f[i] = fv;
      if (df) df[i] = dfv;
    }
  }
}

// The reference's k-space hard-sphere weights on the full grid
// (HardSphereFunctional._init_weight_functions, functionals.py:421-455):
// out[0] = w0 = sinc(kR) sigma_L (sigma_L at k R <= 1e-8) and
// out[1] = 3 (sin kR - kR cos kR) / (kR)^3 sigma_L (sigma_L there); w1, w2,
// w3 and the vector weights are multiples of these.
__global__ void __launch_bounds__(NT) k_weight_table(GridDev g, double R, double* __restrict__ out) {
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int iz = (int)(i % g.nz);
    const long long r = i / g.nz;
    const int iy = (int)(r % g.ny), ix = (int)(r / g.ny);
    const double kx = g.kx.k[ix], ky = g.ky.k[iy], kz = g.kz.k[iz];
    const double sl = g.kx.lanczos[ix] * g.ky.lanczos[iy] * g.kz.lanczos[iz];
    const double w = sqrt(kx * kx + ky * ky + kz * kz) * R;
    if (w <= 1e-8) {
      out[i] = sl;
      out[n + i] = sl;
    } else {
      double sw, cw;
      sincos(w, &sw, &cw);
      out[i] = sw / w * sl;
      out[n + i] = j_ball(w, sw, cw) * sl;
    }
  }
}

// ---------------------------------------------------------------------------
// Final z pass: C2R of the back-convolved sum -> C1 = c_ex + V, then one of
//   MODE_C1       write C1
//   MODE_PICARD   rho' = (1-a) rho + a G, G = beta exp(-beta C1) z ;
//                 zero below the threshold; write rho', R2C(rho') -> rz
//   MODE_ANDRES   r = G - rho into the residual ring + Gram row
//   MODE_CHECK    nothing written; |G - rho|^2 and the derivative criterion
//                 of Solver._check_convergence (solver.py:131-135, 201-205):
//                 with beta dOmega = ln(Lambda^3 rho_c) + beta C1 - beta mu
//                 (rho_c: rho < 1e-12 -> 1e-30) and over |rho| > 1e-8 (the
//                 complement of isclose(rho, 0)), the sums of
//                 (|rho| beta dOmega / (a_tol + r_tol |rho|))^2 and of
//                 (|rho| beta dOmega)^2 (the solver history's variant)
// Always reduces (for the density rho entering this pass) N, the ideal-gas
// sum and the external-potential sum.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(NT) k_zfinal(const double2* __restrict__ F, long long Fs, GridDev g, ElDev el,
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
  const double* rho = el.rho + inst * el.fstride;
  const double* V = el.vext ? el.vext + (el.vext_batched ? inst * el.fstride : 0) : nullptr;
  const InstDev ip = el.inst[inst];
  // beta mu from the fugacity exp(beta mu) / (beta Lambda^3)
  const double bmu = (MODE == MODE_CHECK) ? log(ip.fug * el.beta * el.lam3) : 0.0;
  constexpr int NQ = (MODE == MODE_ANDRES) ? (4 + CMM_MMAX) : 8;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  double* rnew_hist = nullptr;
  if (MODE == MODE_ANDRES) rnew_hist = el.rhist + (long long)el.slot_new * el.hstride + inst * el.fstride;
  const int tot = t.P * n;
  // EC elements per thread at a time, loads grouped (addresses clamped into
  // range); the Gram row walks the residual slots in the outer loop so each
  // slot's EC loads are in flight together -- per quantity the sums keep the
  // element order
  constexpr int EC = 8;
  for (int i0 = threadIdx.x; i0 < tot; i0 += EC * blockDim.x) {
    long long idx[EC];
    bool ok[EC];
    double vv[EC], r[EC], rr[EC];
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      const int ic = min(i, tot - 1);
      const int pl = (int)fdiv((unsigned)ic, t.fn), e = ic - pl * n;
      ok[u] = (i < tot) && (p0 + pl < npen);
      idx[u] = min(p0 + pl, npen - 1) * n + e;
    }
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      vv[u] = V ? V[idx[u]] : 0.0;
      r[u] = rho[idx[u]];
    }
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      rr[u] = 0.0;
      if (!ok[u]) continue;
      const int i = i0 + u * (int)blockDim.x;
      const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
      double& cv = z_val(sm, t, 0, pl, e);
      const double C1 = cv + vv[u];
      // ideal-gas and external terms of the grand potential of rho (the
      // Anderson residual pass forms only N: its callers take the other two
      // sums from the producer pass of the same density)
      const double rc = (r[u] < 1e-12) ? 1e-30 : r[u];
      const double lg = (MODE == MODE_ANDRES) ? 0.0 : log(el.lam3 * rc);
      acc[0] += r[u];
      if (MODE != MODE_ANDRES) {
        acc[1] += rc * (lg - 1.0);
        acc[2] += r[u] * vv[u];
      }
      if (MODE == MODE_C1) {
        el.c1out[inst * el.fstride + idx[u]] = C1;
      } else if (MODE == MODE_CHECK) {
        const double G = el.beta * exp(-el.beta * C1) * ip.fug;
        const double d = G - r[u];
        acc[3] += isfinite(d) ? 0.0 : 1.0;
        acc[4] += d * d;
        const double ar = fabs(r[u]);
        if (ar > 1e-8) {
          const double t = ar * (lg + el.beta * C1 - bmu);
          const double q = t / (el.a_tol + el.r_tol * ar);
          acc[5] += q * q;
          acc[6] += t * t;
        }
      } else {
        if (el.c1out) el.c1out[inst * el.fstride + idx[u]] = C1;
        const double G = el.beta * exp(-el.beta * C1) * ip.fug;
        if (MODE == MODE_PICARD) {
          double rn = (1.0 - ip.alpha) * r[u] + ip.alpha * G;
          if (rn < el.zero_thr) rn = 0.0;
          el.rho_out[inst * el.fstride + idx[u]] = rn;
          cv = rn;
          const double d = fabs(rn - r[u]);
          acc[3] += rn;
          acc[4] += d;
          acc[5] = fmax(acc[5], d);
          acc[6] += isfinite(rn) ? 0.0 : 1.0;
        } else {
          rr[u] = G - r[u];
          rnew_hist[idx[u]] = rr[u];
          acc[3] += isfinite(rr[u]) ? 0.0 : 1.0;
        }
      }
    }
    if (MODE == MODE_ANDRES) {
#pragma unroll
      for (int j = 0; j < CMM_MMAX; ++j) {
        if (j < el.m) {
          if (j == el.slot_new) {
#pragma unroll
            for (int u = 0; u < EC; ++u)
              if (ok[u]) acc[4 + j] += rr[u] * rr[u];
          } else {
            const double* hs = el.rhist + (long long)j * el.hstride + inst * el.fstride;
            double o[EC];
#pragma unroll
            for (int u = 0; u < EC; ++u) o[u] = hs[idx[u]];
#pragma unroll
            for (int u = 0; u < EC; ++u)
              if (ok[u]) acc[4 + j] += rr[u] * o[u];
          }
        }
      }
    }
  }
  if (MODE == MODE_PICARD) {
    __syncthreads();
    z_fft<-1>(sm, t, g.fz, s_tw);
    z_store_half(sm, el.rz + inst * el.rzstride, p0, t, npen, n, g.nzh, 0);
  }
  int ops[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) ops[q] = 0;
  if (MODE == MODE_PICARD) ops[5] = 1;
  block_reduce<NQ>(acc, ops, sh);
  grid_reduce_store<NQ>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// Density producer passes (fused with the z R2C of the produced density):
//   MODE_MIX    Anderson mixing  rho' = clip(sum_i c_i rho_i + d_i r_i)
//   MODE_TRIAL  line-search point rho' = rho + a r   (hybrid Picard)
//   MODE_SEED   rho' = rho (just transform an initial density)
// Reduces N', sum|rho'-rho_cur|, max|rho'-rho_cur|, non-finite count, and the
// ideal-gas / external sums of rho'.
template <int MODE>
__global__ void __launch_bounds__(NT) k_zproduce(GridDev g, ElDev el, ZTile t, const int* __restrict__ active,
                                                RedDev red) {
  extern __shared__ double2 sm[];
  __shared__ double2 s_tw[CMM_MAX_N];
  load_twiddles(s_tw, g.fz);
  __shared__ double sh[CMM_RED_SH];
  const int inst = blockIdx.y;
  if (active && !active[inst]) return;
  const int n = g.nz;
  const long long npen = (long long)g.nx * g.ny;
  const long long p0 = (long long)blockIdx.x * t.P;
  const double* V = el.vext ? el.vext + (el.vext_batched ? inst * el.fstride : 0) : nullptr;
  const double* rcur = el.rho + inst * el.fstride;
  double* rout = el.rho_out + inst * el.fstride;
  const double* coef = el.coef + inst * 2 * CMM_MMAX;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int tot = t.P * n;
  // EC elements per thread at a time; the history fields are walked in the
  // outer loop so that the EC loads of one field are in flight together (the
  // per-element sums keep the field order)
  constexpr int EC = 8;
  for (int i0 = threadIdx.x; i0 < tot; i0 += EC * blockDim.x) {
    long long idx[EC];
    double r0[EC], rn[EC];
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, tot - 1);
      const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
      idx[u] = min(p0 + pl, npen - 1) * n + e;
    }
#pragma unroll
    for (int u = 0; u < EC; ++u) r0[u] = rcur[idx[u]];
    if (MODE == MODE_MIX) {
#pragma unroll
      for (int u = 0; u < EC; ++u) rn[u] = 0.0;
      for (int j = 0; j < el.mk; ++j) {
        const double a = coef[j], b = coef[CMM_MMAX + j];
        const double* hr = el.rhohist + (long long)el.order_rho[j] * el.hstride2 + inst * el.fstride;
        const double* hs = el.rhist + (long long)el.order[j] * el.hstride + inst * el.fstride;
        double x[EC], y[EC];
#pragma unroll
        for (int u = 0; u < EC; ++u) {
          x[u] = hr[idx[u]];
          y[u] = hs[idx[u]];
        }
#pragma unroll
        for (int u = 0; u < EC; ++u) {
          rn[u] += a * x[u];
          rn[u] += b * y[u];
        }
      }
#pragma unroll
      for (int u = 0; u < EC; ++u) rn[u] = (rn[u] < 1e-12) ? 1e-30 : rn[u];
    } else if (MODE == MODE_TRIAL) {
      const double* hs = el.rhist + (long long)el.slot_new * el.hstride + inst * el.fstride;
      const double al = el.inst[inst].alpha;
#pragma unroll
      for (int u = 0; u < EC; ++u) {
        rn[u] = r0[u] + al * hs[idx[u]];
        if (el.clip_trial && rn[u] < 1e-12) rn[u] = 1e-30;
      }
    } else {
#pragma unroll
      for (int u = 0; u < EC; ++u) rn[u] = r0[u];
    }
    double vv[EC];
#pragma unroll
    for (int u = 0; u < EC; ++u) vv[u] = V ? V[idx[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < tot) {
        const int pl = (int)fdiv((unsigned)i, t.fn), e = i - pl * n;
        double v = 0.0;
        if (p0 + pl < npen) {
          v = rn[u];
          if (MODE != MODE_SEED) rout[idx[u]] = v;
          const double rc = (v < 1e-12) ? 1e-30 : v;
          const double d = fabs(v - r0[u]);
          acc[0] += v;
          acc[1] += d;
          acc[2] = fmax(acc[2], d);
          acc[3] += isfinite(v) ? 0.0 : 1.0;
          acc[4] += rc * (log(el.lam3 * rc) - 1.0);
          acc[5] += v * vv[u];
        }
        z_val(sm, t, 0, pl, e) = v;
      }
    }
  }
  __syncthreads();
  z_fft<-1>(sm, t, g.fz, s_tw);
  z_store_half(sm, el.rz + inst * el.rzstride, p0, t, npen, n, g.nzh, 0);
  const int ops[7] = {0, 0, 1, 0, 0, 0, 0};
  block_reduce<7>(acc, ops, sh);
  grid_reduce_store<7>(acc, red.partials + (long long)inst * red.pstride, gridDim.x, blockIdx.x);
}

// Per-instance totals of the partials a producer kernel stored, one per
// block (grid_reduce_store) or per wave (wave_reduce_store; nblocks counts
// them): out[q] = sum (or max, bit q of maxmask) over the
// nblocks partials of quantity q, in a fixed order (lane-strided, then the
// wavefront tree, then the waves in order).  One block per (instance,
// quantity); each thread issues its loads in groups of 8 before summing them
// in order, so the latency of one load is paid per group, not per partial.
// Quantity q goes to slot + q; quantities with bit q of skipmask are not
// stored (their slots keep what another pass put there).
__global__ void __launch_bounds__(NT) k_reduce_partials(RedDev red, int nblocks, unsigned maxmask, int slot,
                                                       const int* __restrict__ active, unsigned skipmask) {
  __shared__ double sh[NT / 64];
  const int inst = blockIdx.x, q = blockIdx.y;
  if ((active && !active[inst]) || ((skipmask >> q) & 1u)) return;
  const double* part = red.partials + (long long)inst * red.pstride + (long long)q * nblocks;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const bool mx = (maxmask >> q) & 1u;
  const double id = mx ? -INFINITY : 0.0;
  double a = id;
  for (int i0 = threadIdx.x; i0 < nblocks; i0 += 8 * blockDim.x) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * blockDim.x;
      x[u] = (i < nblocks) ? part[i] : id;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) a = mx ? fmax(a, x[u]) : a + x[u];
  }
  const double r = mx ? wave_max(a) : wave_sum(a);
  if (lane == 0) sh[wid] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int w = 1; w < nw; ++w) t = mx ? fmax(t, sh[w]) : t + sh[w];
    red.scal[(long long)inst * NSCAL + slot + q] = t;
  }
}

// G = beta exp(-beta C1) z, the density update of Solver.get_new_rho
// (solver.py:128-129), elementwise over n values.
__global__ void __launch_bounds__(NT) k_boltzmann(const double* __restrict__ c1, long long n, double beta, double fug,
                                                 double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = beta * exp(-beta * c1[i]) * fug;
}

// ---------------------------------------------------------------------------
// Anderson coefficients: minimise a^T G a over the simplex {a >= 0, sum a = 1}
// exactly, by solving the equality-constrained problem on every support and
// keeping the best feasible one (a convex QP attains its minimum on the
// relative interior of some face, where it equals that face's affine
// minimiser).  One wavefront per instance.  Also updates the Gram matrix ring
// from this step's row and the adaptive damping.
// ---------------------------------------------------------------------------
// Minimiser of a^T G a on the face {a_i = 0 for i not in S, sum a = 1} for the
// support S = mask: a_S = G_S^-1 1 / (1^T G_S^-1 1).  Solved as the fixed-size
// system A y = P 1 with A = P G P + (I - P) (P = diag(mask)), so every loop
// has a compile-time bound and the matrix lives in registers: Cholesky
// factorisation (G_S is a Gram matrix), then the two triangular solves.
// Returns false if G_S is numerically singular (pivot <= 1e-13 of the unit
// diagonal scale) or the minimiser leaves the simplex.
__device__ bool kkt_solve(const double* Gm, int mk, unsigned mask, double* alpha) {
  double Lm[CMM_MMAX][CMM_MMAX];
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) {
#pragma unroll
    for (int j = 0; j < CMM_MMAX; ++j) {
      const bool in = (i < mk) && (j < mk) && ((mask >> i) & 1u) && ((mask >> j) & 1u);
      Lm[i][j] = in ? Gm[i * CMM_MMAX + j] : (i == j ? 1.0 : 0.0);
    }
  }
#pragma unroll
  for (int j = 0; j < CMM_MMAX; ++j) {
    double d = Lm[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= Lm[j][k] * Lm[j][k];
    if (!(d > 1e-13)) return false;
    const double ljj = sqrt(d);
    Lm[j][j] = ljj;
    const double inv = 1.0 / ljj;
#pragma unroll
    for (int i = j + 1; i < CMM_MMAX; ++i) {
      double t = Lm[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= Lm[i][k] * Lm[j][k];
      Lm[i][j] = t * inv;
    }
  }
  double y[CMM_MMAX];
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) {          // L z = P 1
    double t = (i < mk && ((mask >> i) & 1u)) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) t -= Lm[i][k] * y[k];
    y[i] = t / Lm[i][i];
  }
#pragma unroll
  for (int i = CMM_MMAX - 1; i >= 0; --i) {     // L^T y = z
    double t = y[i];
#pragma unroll
    for (int k = i + 1; k < CMM_MMAX; ++k) t -= Lm[k][i] * y[k];
    y[i] = t / Lm[i][i];
  }
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) sum += (i < mk && ((mask >> i) & 1u)) ? y[i] : 0.0;
  if (!(sum > 0.0)) return false;
  const double isum = 1.0 / sum;
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) {
    const bool in = i < mk && ((mask >> i) & 1u);
    const double a = in ? y[i] * isum : 0.0;
    if (a < -1e-12) return false;
    alpha[i] = fmax(a, 0.0);
  }
  double s2 = 0.0;
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) s2 += alpha[i];
#pragma unroll
  for (int i = 0; i < CMM_MMAX; ++i) alpha[i] /= s2;
  return true;
}

__global__ void __launch_bounds__(64) k_anderson_coef(AndDev ad, const int* __restrict__ active, RedDev red) {
  const int inst = blockIdx.x;
  if (active && !active[inst]) return;
  double* Gring = ad.gram + inst * CMM_MMAX * CMM_MMAX;   // by ring slot
  const double* row = red.scal + inst * NSCAL + S_GRAM;
  const int lane = threadIdx.x;
  // this step's Gram row (slot_new against every ring slot)
  if (lane < ad.m) {
    Gring[ad.slot_new * CMM_MMAX + lane] = row[lane];
    Gring[lane * CMM_MMAX + ad.slot_new] = row[lane];
  }
  if (ad.store_only) return;   // a step without mixing keeps the ring current
  __syncthreads();
  const int mk = ad.mk;
  // ordered sub-matrix (oldest -> newest), normalised by its largest diagonal
  __shared__ double Gm[CMM_MMAX * CMM_MMAX];
  __shared__ double scale;
  if (lane == 0) {
    double mx = 0.0;
    for (int i = 0; i < mk; ++i) mx = fmax(mx, Gring[ad.order[i] * CMM_MMAX + ad.order[i]]);
    scale = (mx > 0.0 && isfinite(mx)) ? 1.0 / mx : 0.0;
  }
  __syncthreads();
  for (int t = lane; t < mk * mk; t += 64) {
    const int i = t / mk, j = t - i * mk;
    Gm[i * CMM_MMAX + j] = Gring[ad.order[i] * CMM_MMAX + ad.order[j]] * scale;
  }
  __syncthreads();
  double best = INFINITY;
  unsigned bestmask = 0;
  double alpha[CMM_MMAX];
  const unsigned nsub = (1u << mk) - 1u;
  if (scale > 0.0) {
    for (unsigned m = 1 + lane; m <= nsub; m += 64) {
      if (!kkt_solve(Gm, mk, m, alpha)) continue;
      double f = 0.0;
#pragma unroll
      for (int i = 0; i < CMM_MMAX; ++i) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < CMM_MMAX; ++j) t += (i < mk && j < mk) ? Gm[i * CMM_MMAX + j] * alpha[j] : 0.0;
        f += alpha[i] * t;
      }
      if (f < best || (f == best && m < bestmask)) { best = f; bestmask = m; }
    }
  }
  // wave argmin (value, then mask)
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const unsigned om = __shfl_xor(bestmask, o, 64);
    if (ob < best || (ob == best && om < bestmask && om != 0)) { best = ob; bestmask = om; }
  }
  if (lane == 0) {
    if (bestmask == 0 || !kkt_solve(Gm, mk, bestmask, alpha)) {
      for (int i = 0; i < mk; ++i) alpha[i] = 1.0 / mk;   // degenerate: uniform start point
    }
    // adaptive damping from the newest and previous residual norms
    double d = ad.damping[inst];
    if (ad.adaptive) {
      const double rn = Gring[ad.slot_new * CMM_MMAX + ad.slot_new];
      const double rp = (ad.slot_prev >= 0) ? Gring[ad.slot_prev * CMM_MMAX + ad.slot_prev] : 0.0;
      if (sqrt(rn) < sqrt(rp)) d = fmin(d * ad.fac_up, ad.dmax);
      else d = fmax(d * ad.fac_down, ad.dmin);
      ad.damping[inst] = d;
    }
    red.scal[inst * NSCAL + S_ADAMP] = d;
    const double dc = d * ad.cf[inst];
    double* coef = ad.coef + inst * 2 * CMM_MMAX;
    for (int i = 0; i < CMM_MMAX; ++i) {
      coef[i] = (i < mk) ? alpha[i] : 0.0;
      coef[CMM_MMAX + i] = (i < mk) ? dc * alpha[i] : 0.0;
      ad.alpha_out[inst * CMM_MMAX + i] = (i < mk) ? alpha[i] : 0.0;
    }
  }
}

// Per-instance convergence and safeguard test of one batched Anderson
// iteration (Solver._check_convergence with RIUE, solver.py:167-234, and the
// grand-potential test of Anderson.update_rho, solver.py:567-574).  Reads the
// scalars of the density just produced (N', sum|drho|, non-finite count,
// ideal/external sums) and of its functional chain (sum Phi, sum f_ex and the
// mean-field energy).
// init = 1 only records Omega of the seeded densities.
__global__ void k_batch_check(BatchDev b, RedDev red, int new_slot, int init) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= b.nb || !b.active[inst]) return;
  const double* s = red.scal + inst * NSCAL;
  const double N = s[S_PN];
  const double om = s[S_PN + 4] * b.dr * b.kT - b.mu[inst] * N * b.dr + s[S_PHI] * b.dr * b.kT + s[S_PHI + 1] * b.dr +
                    s[S_PN + 5] * b.dr + (b.mfa ? s[S_FMFA] * b.dr : 0.0);
  b.omega[inst] = om;
  if (init) {
    b.omega_prev[inst] = om;
    b.final_slot[inst] = new_slot;
    return;
  }
  const int it = b.iters[inst] + 1;
  b.iters[inst] = it;
  const bool nonfinite = s[S_PN + 3] > 0.0 || !isfinite(s[S_PHI]) || !isfinite(s[S_PHI + 1]) || !isfinite(N);
  if (nonfinite) {
    b.status[inst] = ST_NONFINITE;
    b.active[inst] = 0;
    return;
  }
  if (om > 0.8 * b.omega_prev[inst]) {     // the reference would take a hybrid Picard step here
    b.status[inst] = ST_FALLBACK;
    b.active[inst] = 0;
    return;
  }
  b.omega_prev[inst] = om;
  b.final_slot[inst] = new_slot;
  const double iue = s[S_PN + 1];
  const bool conv = (N > 0.0) ? (iue / N < b.tol) : (iue == 0.0);
  if (conv) {
    b.status[inst] = ST_CONVERGED;
    b.active[inst] = 0;
  } else if (it >= b.nsteps) {
    b.status[inst] = ST_MAXSTEPS;
    b.active[inst] = 0;
  }
}

}  // namespace cmm
