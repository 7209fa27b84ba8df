// Shared device-side building blocks of the cDFT hot path for CDNA4 (gfx950):
// complex FP64 helpers, the in-LDS Stockham FFT (radix 2/3/4/5/8, mixed,
// runtime stage plan), and deterministic block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CMM_MAX_STAGES 8
#define CMM_MAX_N 256   // longest transform axis (twiddles are staged in LDS)
#define CMM_MAX_INST 256

// Unsigned division by a runtime constant d as a multiply-high and a shift
// (round-up method, Granlund & Montgomery 1994): q = (umulhi(x, m) + x) >> l
// with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1.  Exact for
// x < 2^31 (all indices here are far below that).
struct FastDiv {
  unsigned d, m, l;
};

__host__ __device__ inline FastDiv make_fastdiv(unsigned d) {
  FastDiv f;
  f.d = d;
  unsigned l = 0;
  while ((1u << l) < d) ++l;
  f.l = l;
  f.m = (unsigned)(((((unsigned long long)1) << 32) * ((((unsigned long long)1) << l) - d)) / d + 1);
  return f;
}

__device__ __forceinline__ unsigned fdiv(unsigned x, const FastDiv& f) { return (__umulhi(x, f.m) + x) >> f.l; }

// One 1-D transform length: radix schedule, per-stage divisors and the
// twiddle table tw[j] = exp(-2*pi*i*j/n), j in [0, n).
struct Fft1D {
  int n;
  int nst;
  int rad[CMM_MAX_STAGES];
  FastDiv fns[CMM_MAX_STAGES];   // Ns of the stage (product of the previous radices)
  FastDiv fnq[CMM_MAX_STAGES];   // n / radix of the stage
  const double2* tw;
};

// Per-axis k-space tables of one grid: k values (2*pi*fftfreq(n, h)), the
// Lanczos sigma factor sinc(k h / 2 pi) and the Nyquist index (-1 if n odd).
struct KAxis {
  const double* k;
  const double* lanczos;
  int nyq;
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by (sgn * i)
__device__ __forceinline__ double2 cmul_i(double2 a, double sgn) { return make_double2(-sgn * a.y, sgn * a.x); }

// ---------------------------------------------------------------------------
// Radix-r DFT butterflies. SGN = -1: forward (exp(-2 pi i tu/r)); +1 inverse.
// ---------------------------------------------------------------------------
template <int SGN>
__device__ __forceinline__ void dft2(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int SGN>
__device__ __forceinline__ void dft3(double2* v) {
  const double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
  double2 t1 = cadd(v[1], v[2]);
  double2 t2 = csub(v[1], v[2]);
  double2 m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  double2 r = cmul_i(cscale(t2, s3), (double)SGN);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int SGN>
__device__ __forceinline__ void dft4(double2* v) {
  double2 a = cadd(v[0], v[2]);
  double2 b = csub(v[0], v[2]);
  double2 c = cadd(v[1], v[3]);
  double2 d = cmul_i(csub(v[1], v[3]), (double)SGN);
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

template <int SGN>
__device__ __forceinline__ void dft5(double2* v) {
  const double c1 = 0.30901699437494742410;   // cos(2 pi / 5)
  const double c2 = -0.80901699437494742410;  // cos(4 pi / 5)
  const double s1 = 0.95105651629515357212;   // sin(2 pi / 5)
  const double s2 = 0.58778525229247312917;   // sin(4 pi / 5)
  double2 t1 = cadd(v[1], v[4]);
  double2 t2 = cadd(v[2], v[3]);
  double2 t3 = csub(v[1], v[4]);
  double2 t4 = csub(v[2], v[3]);
  double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  double2 b1 = cmul_i(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), (double)SGN);
  double2 b2 = cmul_i(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), (double)SGN);
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int SGN>
__device__ __forceinline__ void dft8(double2* v) {
  const double h = 0.70710678118654752440;  // 1/sqrt(2)
  // two radix-4 on even/odd, then combine with W8 twiddles
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<SGN>(e);
  dft4<SGN>(o);
  // o[k] *= exp(SGN 2 pi i k / 8)
  double2 o1 = make_double2(h * (o[1].x - SGN * o[1].y), h * (o[1].y + SGN * o[1].x));
  double2 o2 = cmul_i(o[2], (double)SGN);
  double2 o3 = make_double2(-h * (o[3].x + SGN * o[3].y), h * (SGN * o[3].x - o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// x (c + SGN i s): multiply by exp(SGN i theta) given c = cos, s = sin
template <int SGN>
__device__ __forceinline__ double2 crot(double2 x, double c, double s) {
  return make_double2(x.x * c - SGN * x.y * s, x.y * c + SGN * x.x * s);
}

// 16-point DFT as 4 x 4: radix-4 transforms over the stride-4 subsequences,
// the twiddles exp(SGN 2 pi i n1 k2 / 16), radix-4 transforms across them;
// the outputs are in natural order.
template <int SGN>
__device__ __forceinline__ void dft16(double2* v) {
  const double c1 = 0.92387953251128675613;   // cos(pi / 8)
  const double s1 = 0.38268343236508977173;   // sin(pi / 8)
  const double h = 0.70710678118654752440;    // cos(pi / 4)
  double2 a[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    double2 x[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4<SGN>(x);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) a[n1][k2] = x[k2];
  }
  a[1][1] = crot<SGN>(a[1][1], c1, s1);    // exponent 1
  a[1][2] = crot<SGN>(a[1][2], h, h);      // 2
  a[1][3] = crot<SGN>(a[1][3], s1, c1);    // 3
  a[2][1] = crot<SGN>(a[2][1], h, h);      // 2
  a[2][2] = cmul_i(a[2][2], (double)SGN);  // 4
  a[2][3] = crot<SGN>(a[2][3], -h, h);     // 6
  a[3][1] = crot<SGN>(a[3][1], s1, c1);    // 3
  a[3][2] = crot<SGN>(a[3][2], -h, h);     // 6
  a[3][3] = crot<SGN>(a[3][3], -c1, -s1);  // 9
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    double2 y[4] = {a[0][k2], a[1][k2], a[2][k2], a[3][k2]};
    dft4<SGN>(y);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = y[k1];
  }
}

template <int R, int SGN>
__device__ __forceinline__ void dftr(double2* v) {
  if constexpr (R == 2) dft2<SGN>(v);
  else if constexpr (R == 3) dft3<SGN>(v);
  else if constexpr (R == 4) dft4<SGN>(v);
  else if constexpr (R == 5) dft5<SGN>(v);
  else if constexpr (R == 16) dft16<SGN>(v);
  else dft8<SGN>(v);
}

// ---------------------------------------------------------------------------
// In-place Stockham stage over `batch` sequences stored in LDS.
// Element e of sequence s lives at buf[s*bs + e*es].  EM (element-major) picks
// the butterfly -> thread mapping that keeps consecutive lanes on consecutive
// LDS addresses: sequences fastest when es==batch-stride layout [n][batch],
// elements fastest for [batch][n].  Each thread stages at most NB butterflies
// in registers: all lanes load, barrier, all lanes store, barrier.
// ---------------------------------------------------------------------------
template <int R, int SGN, bool EM, int NB>
__device__ __forceinline__ void stockham_stage(double2* buf, int batch, const FastDiv& fb, int bs, int es, const Fft1D& p,
                                               const double2* tws, int st, int Ns) {
  const int n = p.n;
  const int nq = n / R;                 // butterflies per sequence
  const int total = batch * nq;
  const int twstride = n / (Ns * R);
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  double2 v[NB][R];
  int dst[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * nt;
    dst[i] = -1;
    if (b < total) {
      int s, j;
      if (EM) { j = (int)fdiv((unsigned)b, fb); s = b - j * batch; }
      else { s = (int)fdiv((unsigned)b, p.fnq[st]); j = b - s * nq; }
      const int jn = (int)fdiv((unsigned)j, p.fns[st]);
      const int k = j - jn * Ns;
      double2* base = buf + s * bs;
#pragma unroll
      for (int t = 0; t < R; ++t) v[i][t] = base[(j + t * nq) * es];
      if (Ns > 1) {
        // t k twstride < R Ns (n / (Ns R)) = n: no wrap-around
#pragma unroll
        for (int t = 1; t < R; ++t) {
          double2 w = tws[t * k * twstride];
          if (SGN > 0) w.y = -w.y;
          v[i][t] = cmul(v[i][t], w);
        }
      }
      dftr<R, SGN>(v[i]);
      dst[i] = s * bs + (jn * Ns * R + k) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (dst[i] >= 0) {
#pragma unroll
      for (int t = 0; t < R; ++t) buf[dst[i] + t * Ns * es] = v[i][t];
    }
  }
  __syncthreads();
}

// The same stage with the length and the stage's Ns fixed at compile time:
// the index divisions become shifts or multiplications by constants and the
// twiddle strides immediates.  Same operations in the same order as
// stockham_stage, so the same results bit for bit.
template <int R, int SGN, bool EM, int NB, int N, int NS>
__device__ __forceinline__ void stockham_stage_c(double2* buf, int batch, const FastDiv& fb, int bs, int es,
                                                 const double2* tws) {
  constexpr int nq = N / R;
  constexpr int twstride = N / (NS * R);
  const int total = batch * nq;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  double2 v[NB][R];
  int dst[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * nt;
    dst[i] = -1;
    if (b < total) {
      int s, j;
      if (EM) { j = (int)fdiv((unsigned)b, fb); s = b - j * batch; }
      else { s = (int)((unsigned)b / (unsigned)nq); j = b - s * nq; }
      const int jn = (int)((unsigned)j / (unsigned)NS);
      const int k = j - jn * NS;
      double2* base = buf + s * bs;
#pragma unroll
      for (int t = 0; t < R; ++t) v[i][t] = base[(j + t * nq) * es];
      if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) {
          double2 w = tws[t * k * twstride];
          if (SGN > 0) w.y = -w.y;
          v[i][t] = cmul(v[i][t], w);
        }
      }
      dftr<R, SGN>(v[i]);
      dst[i] = s * bs + (jn * NS * R + k) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (dst[i] >= 0) {
#pragma unroll
      for (int t = 0; t < R; ++t) buf[dst[i] + t * NS * es] = v[i][t];
    }
  }
  __syncthreads();
}

// stockham_stage_c with the sequence count BATCH (the fast index: bs = 1)
// and the element stride ES fixed too, for blocks of NTB threads: every LDS
// offset of a butterfly is an immediate, the butterfly index splits by
// constant divisions, and each thread takes ceil(BATCH N / R / NTB)
// butterflies.  Same operations in the same order as stockham_stage, so the
// same results bit for bit.
template <int R, int SGN, int N, int NS, int BATCH, int ES, int NTB>
__device__ __forceinline__ void stockham_stage_t(double2* buf, const double2* tws) {
  constexpr int nq = N / R;
  constexpr int twstride = N / (NS * R);
  constexpr int total = BATCH * nq;
  constexpr int NB = (total + NTB - 1) / NTB;
  const int tid = threadIdx.x;
  double2 v[NB][R];
  int dst[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * NTB;
    dst[i] = -1;
    if (b < total) {
      const int j = (int)((unsigned)b / (unsigned)BATCH), s = b - j * BATCH;
      const int jn = (int)((unsigned)j / (unsigned)NS), k = j - jn * NS;
      const double2* base = buf + s + j * ES;
#pragma unroll
      for (int t = 0; t < R; ++t) v[i][t] = base[t * nq * ES];
      if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) {
          double2 w = tws[t * k * twstride];
          if (SGN > 0) w.y = -w.y;
          v[i][t] = cmul(v[i][t], w);
        }
      }
      dftr<R, SGN>(v[i]);
      dst[i] = s + (jn * NS * R + k) * ES;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (dst[i] >= 0) {
#pragma unroll
      for (int t = 0; t < R; ++t) buf[dst[i] + t * NS * ES] = v[i][t];
    }
  }
  __syncthreads();
}

// A whole transform of BATCH sequences (radices RS in order) with the tile
// fixed at compile time; ends with a barrier.
template <int SGN, int N, int BATCH, int ES, int NTB, int NS, int R, int... RS>
__device__ __forceinline__ void lds_fft_t(double2* buf, const double2* tws) {
  stockham_stage_t<R, SGN, N, NS, BATCH, ES, NTB>(buf, tws);
  if constexpr (sizeof...(RS) > 0) lds_fft_t<SGN, N, BATCH, ES, NTB, NS * R, RS...>(buf, tws);
}

template <int... RS>
__device__ __forceinline__ bool same_radices(const Fft1D& p) {
  constexpr int rs[] = {RS...};
  if (p.nst != (int)sizeof...(RS)) return false;
#pragma unroll
  for (int i = 0; i < (int)sizeof...(RS); ++i)
    if (p.rad[i] != rs[i]) return false;
  return true;
}

template <int SGN, bool EM, int VPT, int N, int NS, int R, int... RS>
__device__ __forceinline__ void lds_fft_c(double2* buf, int batch, const FastDiv& fb, int bs, int es,
                                          const double2* tws) {
  stockham_stage_c<R, SGN, EM, (VPT + R - 1) / R, N, NS>(buf, batch, fb, bs, es, tws);
  if constexpr (sizeof...(RS) > 0) lds_fft_c<SGN, EM, VPT, N, NS * R, RS...>(buf, batch, fb, bs, es, tws);
}

// Full transform of `batch` sequences of length p.n, in place in LDS;
// element e of sequence s at buf[s*bs + e*es]; tws: the twiddle table in LDS.  EM (element-major): the
// sequences are the fast index (bs = 1, es >= batch), so consecutive lanes
// touch consecutive addresses in every stage.  VPT bounds the complex values
// one thread holds per stage: the caller must guarantee
// batch * n <= VPT * blockDim.x, and that a barrier separates the data's
// producers from this call.  Ends with a barrier.
template <int SGN, bool EM, int VPT, bool R16 = false>
__device__ __forceinline__ void lds_fft(double2* buf, int batch, const FastDiv& fb, int bs, int es, const Fft1D& p,
                        const double2* tws) {
  // R16 (LDS-bound callers): length 128 as two stages, radix 16 then 8 --
  // two LDS round trips and one twiddled stage instead of three and two.
  // Needs batch * n / 16 <= blockDim.x, which batch * n <= VPT * blockDim.x
  // with VPT <= 16 guarantees.
  if constexpr (R16) {
    if (p.n == 128) return lds_fft_c<SGN, EM, VPT, 128, 1, 16, 8>(buf, batch, fb, bs, es, tws);
  }
  // compile-time stage sequences: the tuned plans of 128, 96, 64 and 48 and
  // the default plans of the 5-smooth lengths 48..160
#define CMM_CODELET(NN, ...)                                                          \
  if (p.n == NN && same_radices<__VA_ARGS__>(p))                                     \
    return lds_fft_c<SGN, EM, VPT, NN, 1, __VA_ARGS__>(buf, batch, fb, bs, es, tws);
  CMM_CODELET(128, 8, 4, 4)
  CMM_CODELET(128, 8, 8, 2)
  CMM_CODELET(96, 8, 3, 4)
  CMM_CODELET(96, 4, 8, 3)
  CMM_CODELET(96, 3, 8, 4)
  CMM_CODELET(64, 8, 8)
  CMM_CODELET(48, 4, 4, 3)
  CMM_CODELET(48, 3, 8, 2)
  CMM_CODELET(160, 5, 8, 4)
  CMM_CODELET(144, 3, 3, 8, 2)
  CMM_CODELET(120, 5, 3, 8)
  CMM_CODELET(100, 5, 5, 4)
  CMM_CODELET(80, 5, 8, 2)
  CMM_CODELET(72, 3, 3, 8)
#undef CMM_CODELET
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int r = p.rad[s];
    switch (r) {
      case 2: stockham_stage<2, SGN, EM, (VPT + 1) / 2>(buf, batch, fb, bs, es, p, tws, s, Ns); break;
      case 3: stockham_stage<3, SGN, EM, (VPT + 2) / 3>(buf, batch, fb, bs, es, p, tws, s, Ns); break;
      case 4: stockham_stage<4, SGN, EM, (VPT + 3) / 4>(buf, batch, fb, bs, es, p, tws, s, Ns); break;
      case 5: stockham_stage<5, SGN, EM, (VPT + 4) / 5>(buf, batch, fb, bs, es, p, tws, s, Ns); break;
      default: stockham_stage<8, SGN, EM, (VPT + 7) / 8>(buf, batch, fb, bs, es, p, tws, s, Ns); break;
    }
    Ns *= r;
  }
}

// stage an axis' twiddle table into LDS (caller adds the barrier)
__device__ __forceinline__ void load_twiddles(double2* tws, const Fft1D& p) {
  for (int i = threadIdx.x; i < p.n; i += blockDim.x) tws[i] = p.tw[i];
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  A producer kernel stores partials -- one per
// block (block_reduce + grid_reduce_store: partials[q nblocks + bid]) or one
// per wave (wave_reduce_store) -- and k_reduce_partials, launched after it,
// sums (or maxes) them in a fixed order.  Every tree is fixed, so results
// are bitwise reproducible.
// ---------------------------------------------------------------------------
// Cross-lane move of a double by a DPP control (both 32-bit halves): lanes
// whose row is outside ROWMASK get idv.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v, double idv) {
  const long long b = __double_as_longlong(v), ib = __double_as_longlong(idv);
  const int lo = __builtin_amdgcn_update_dpp((int)ib, (int)b, CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ib >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum (max) over the 64 lanes of a wave, every lane active, in registers
// (DPP moves, no LDS round trips): pairs and quads by quad permutes, half
// rows and rows by mirrors, then rows 1 and 3 take the ends of rows 0 and 2
// (row_bcast:15) and rows 2 and 3 the end of row 1 (row_bcast:31), which
// leaves the total in lane 63; every lane returns it.  A fixed tree, so the
// results are reproducible.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1, 0xF>(v, 0.0);    // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E, 0xF>(v, 0.0);    // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141, 0xF>(v, 0.0);   // row_half_mirror
  v += dpp_f64<0x140, 0xF>(v, 0.0);   // row_mirror
  v += dpp_f64<0x142, 0xA>(v, 0.0);   // row_bcast:15 into rows 1, 3
  v += dpp_f64<0x143, 0xC>(v, 0.0);   // row_bcast:31 into rows 2, 3
  return lane63(v);
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<0xB1, 0xF>(v, -INFINITY));
  v = fmax(v, dpp_f64<0x4E, 0xF>(v, -INFINITY));
  v = fmax(v, dpp_f64<0x141, 0xF>(v, -INFINITY));
  v = fmax(v, dpp_f64<0x140, 0xF>(v, -INFINITY));
  v = fmax(v, dpp_f64<0x142, 0xA>(v, -INFINITY));
  v = fmax(v, dpp_f64<0x143, 0xC>(v, -INFINITY));
  return lane63(v);
}

// LDS scratch of a block reduction: 16 wave slots (up to 1024 threads) for
// each of at most 16 quantities.
#define CMM_RED_SH 256

template <int NQ>
__device__ void block_reduce(double (&v)[NQ], const int (&ops)[NQ], double* sh /* CMM_RED_SH doubles */) {
  static_assert(NQ <= CMM_RED_SH / 16, "too many quantities for one block reduction");
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double x = ops[q] ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) sh[q * 16 + wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double a = sh[q * 16];
      for (int w = 1; w < nw; ++w) a = ops[q] ? fmax(a, sh[q * 16 + w]) : a + sh[q * 16 + w];
      v[q] = a;
    }
  }
  __syncthreads();
}

// Each wave's reduced values (op 0 = sum, 1 = max) stored as partials
// [q nparts + bid nw + wave], nparts = nblocks nw (nw waves per block): no
// block-level combine, so no barriers in the producer; k_reduce_partials
// sums or maxes the nparts partials of each quantity in a fixed order.
template <int NQ>
__device__ __forceinline__ void wave_reduce_store(const double (&v)[NQ], const int (&ops)[NQ], double* partials,
                                                  int nblocks, int bid) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const long long nparts = (long long)nblocks * nw;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double x = ops[q] ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) partials[q * nparts + (long long)bid * nw + wid] = x;
  }
}

// Called by every block with its block-reduced values valid in thread 0:
// stores them as partials[q*nblocks + bid].  The totals are formed by a
// separate small kernel launched after the producer (k_reduce_partials): the
// kernel boundary makes every block's partials visible on every XCD, where a
// last-block-done scheme inside the producer would need a device-scope
// release fence -- an L2 write-back and invalidate -- in every block.
template <int NQ>
__device__ __forceinline__ void grid_reduce_store(const double (&v)[NQ], double* partials, int nblocks, int bid) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) partials[q * nblocks + bid] = v[q];
  }
}
