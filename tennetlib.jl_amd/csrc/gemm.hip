// MFMA GEMM kernels for gfx950 (CDNA4), row-major, batched, optional
// transposes.  Three arithmetic forms:
//   F64    : v_mfma_f64_16x16x4_f64, FP64 accumulate
//   F32    : v_mfma_f32_32x32x2_f32, FP32 accumulate
//   BF16X3 : each FP32 operand split as a = hi + lo (hi = bf16(a), lo = bf16(a - hi));
//            C += hi*hi + hi*lo + lo*hi with v_mfma_f32_32x32x16_bf16, FP32 accumulate
//            (split-precision scheme of Henry, Tang & Heinecke 2019).
//   Z64    : complex FP64 (interleaved), four real f64 MFMAs per complex product.
// Tiles are staged global -> registers -> LDS with one tile of register
// prefetch; a workgroup is 4 wavefronts (2 x 2) of 64 lanes.
#include "tn_internal.h"

#include <type_traits>

namespace tn {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

struct GArgs {
  long M, N, K;
  const void* A; long lda; long sA; int tA; int cA;
  const void* B; long ldb; long sB; int tB; int cB;
  void* C; long ldc; long sC;
  double alpha, beta;
};

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs, so remap the linear id so that each XCD gets a contiguous run
// of output tiles (shared A/B panels stay in that XCD's L2).
__device__ __forceinline__ void tile_coords(int ntm, int ntn, int& tm, int& tn) {
  const int nt = ntm * ntn;
  int id = blockIdx.x;
  if (nt >= 64 && (nt % 8) == 0) {
    const int per = nt / 8;
    id = (id % 8) * per + (id / 8);
  }
  // group rows of tiles (GROUP_M = 8) for L2 reuse of B panels
  const int GM = 8;
  const int group = id / (GM * ntn);
  const int first_m = group * GM;
  const int gsz = min(ntm - first_m, GM);
  const int in_group = id % (GM * ntn);
  tm = first_m + (in_group % gsz);
  tn = in_group / gsz;
}

// ------------------------------------------------------------------------
// FP64 / FP32 kernel: LDS tiles As[BK][BM+PAD], Bs[BK][BN+PAD]
// ------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK>
struct RealTile {
  static constexpr int NT = 256;
  static constexpr int PAD = (sizeof(T) == 8) ? 2 : 4;
  static constexpr int AE = BM * BK / NT;  // A elements per thread
  static constexpr int BE = BK * BN / NT;
};

template <typename T>
__device__ __forceinline__ T ldg(const T* p) { return *p; }

template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void load_tiles(const GArgs& a, const T* __restrict__ A,
                                           const T* __restrict__ B, long m0, long n0, long k0,
                                           T (&ra)[RealTile<T, BM, BN, BK>::AE],
                                           T (&rb)[RealTile<T, BM, BN, BK>::BE]) {
  using TT = RealTile<T, BM, BN, BK>;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < TT::AE; ++i) {
    const int idx = tid + i * TT::NT;
    int mm, kk;
    if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
    const long m = m0 + mm, k = k0 + kk;
    T v = T(0);
    if (m < a.M && k < a.K) v = a.tA ? A[k * a.lda + m] : A[m * a.lda + k];
    ra[i] = v;
  }
#pragma unroll
  for (int i = 0; i < TT::BE; ++i) {
    const int idx = tid + i * TT::NT;
    int nn, kk;
    if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
    const long n = n0 + nn, k = k0 + kk;
    T v = T(0);
    if (n < a.N && k < a.K) v = a.tB ? B[n * a.ldb + k] : B[k * a.ldb + n];
    rb[i] = v;
  }
}

template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void store_tiles(const GArgs& a, T* As, T* Bs,
                                            const T (&ra)[RealTile<T, BM, BN, BK>::AE],
                                            const T (&rb)[RealTile<T, BM, BN, BK>::BE]) {
  using TT = RealTile<T, BM, BN, BK>;
  constexpr int LA = BM + TT::PAD, LB = BN + TT::PAD;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < TT::AE; ++i) {
    const int idx = tid + i * TT::NT;
    int mm, kk;
    if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
    As[kk * LA + mm] = ra[i];
  }
#pragma unroll
  for (int i = 0; i < TT::BE; ++i) {
    const int idx = tid + i * TT::NT;
    int nn, kk;
    if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
    Bs[kk * LB + nn] = rb[i];
  }
}

// FP64: wave tile (BM/2) x (BN/2) made of 16x16 MFMA blocks
template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GArgs a) {
  using T = double;
  using TT = RealTile<T, BM, BN, BK>;
  constexpr int LA = BM + TT::PAD, LB = BN + TT::PAD;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  __shared__ T As[BK * LA];
  __shared__ T Bs[BK * LB];

  const int ntm = (int)((a.M + BM - 1) / BM), ntn = (int)((a.N + BN - 1) / BN);
  int tm, tn;
  tile_coords(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const long bz = blockIdx.z;
  const T* A = reinterpret_cast<const T*>(a.A) + bz * a.sA;
  const T* B = reinterpret_cast<const T*>(a.B) + bz * a.sB;
  T* C = reinterpret_cast<T*>(a.C) + bz * a.sC;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  d4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0, 0, 0, 0};

  T ra[TT::AE], rb[TT::BE];
  load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, 0, ra, rb);
  for (long k0 = 0; k0 < a.K; k0 += BK) {
    __syncthreads();
    store_tiles<T, BM, BN, BK>(a, As, Bs, ra, rb);
    __syncthreads();
    if (k0 + BK < a.K) load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, k0 + BK, ra, rb);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double fa[FM], fb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) fa[i] = As[kr * LA + wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[j] = Bs[kr * LB + wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }
  // epilogue: register r of lane holds row lane/16 + 4r, column lane%16 of
  // each 16x16 block (v_mfma_f64_16x16x4f64 output layout)
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * WTM + i * 16 + (lane >> 4) + 4 * r;
        const long n = n0 + wn * WTN + j * 16 + (lane & 15);
        if (m < a.M && n < a.N) {
          double v = a.alpha * acc[i][j][r];
          if (a.beta != 0.0) v += a.beta * C[m * a.ldc + n];
          C[m * a.ldc + n] = v;
        }
      }
}

// FP32: wave tile (BM/2) x (BN/2) made of 32x32 MFMA blocks
template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_f32_kernel(GArgs a) {
  using T = float;
  using TT = RealTile<T, BM, BN, BK>;
  constexpr int LA = BM + TT::PAD, LB = BN + TT::PAD;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 32, FN = WTN / 32;
  __shared__ T As[BK * LA];
  __shared__ T Bs[BK * LB];

  const int ntm = (int)((a.M + BM - 1) / BM), ntn = (int)((a.N + BN - 1) / BN);
  int tm, tn;
  tile_coords(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const long bz = blockIdx.z;
  const T* A = reinterpret_cast<const T*>(a.A) + bz * a.sA;
  const T* B = reinterpret_cast<const T*>(a.B) + bz * a.sB;
  T* C = reinterpret_cast<T*>(a.C) + bz * a.sC;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  f16v acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f16v{};

  T ra[TT::AE], rb[TT::BE];
  load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, 0, ra, rb);
  for (long k0 = 0; k0 < a.K; k0 += BK) {
    __syncthreads();
    store_tiles<T, BM, BN, BK>(a, As, Bs, ra, rb);
    __syncthreads();
    if (k0 + BK < a.K) load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, k0 + BK, ra, rb);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int kr = kk + (lane >> 5);
      float fa[FM], fb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) fa[i] = As[kr * LA + wm * WTM + i * 32 + (lane & 31)];
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[j] = Bs[kr * LB + wn * WTN + j * 32 + (lane & 31)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }
  // epilogue: register r of lane -> row 8*(r/4) + 4*(lane/32) + r%4, column lane%32
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long m = m0 + wm * WTM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long n = n0 + wn * WTN + j * 32 + (lane & 31);
        if (m < a.M && n < a.N) {
          float v = (float)a.alpha * acc[i][j][r];
          if (a.beta != 0.0) v += (float)a.beta * C[m * a.ldc + n];
          C[m * a.ldc + n] = v;
        }
      }
}

// ------------------------------------------------------------------------
// BF16x3: LDS tiles hold bf16 hi/lo parts with k contiguous per row:
// Ah[BM][BK+8], Al[BM][BK+8], Bh[BN][BK+8], Bl[BN][BK+8].
// ------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}

template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GArgs a) {
  constexpr int NT = 256;
  constexpr int LK = BK + 8;  // row pitch in bf16 elements (16-byte aligned)
  constexpr int AE = BM * BK / NT, BE = BK * BN / NT;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 32, FN = WTN / 32;
  __shared__ __attribute__((aligned(16))) unsigned short Ah[BM * LK];
  __shared__ __attribute__((aligned(16))) unsigned short Al[BM * LK];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[BN * LK];
  __shared__ __attribute__((aligned(16))) unsigned short Bl[BN * LK];

  const int ntm = (int)((a.M + BM - 1) / BM), ntn = (int)((a.N + BN - 1) / BN);
  int tm, tn;
  tile_coords(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const long bz = blockIdx.z;
  const float* A = reinterpret_cast<const float*>(a.A) + bz * a.sA;
  const float* B = reinterpret_cast<const float*>(a.B) + bz * a.sB;
  float* C = reinterpret_cast<float*>(a.C) + bz * a.sC;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  f16v acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f16v{};

  float ra[AE], rb[BE];
  auto gload = [&](long k0) {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int idx = tid + i * NT;
      int mm, kk;
      if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      const long m = m0 + mm, k = k0 + kk;
      float v = 0.f;
      if (m < a.M && k < a.K) v = a.tA ? A[k * a.lda + m] : A[m * a.lda + k];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int idx = tid + i * NT;
      int nn, kk;
      if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
      const long n = n0 + nn, k = k0 + kk;
      float v = 0.f;
      if (n < a.N && k < a.K) v = a.tB ? B[n * a.ldb + k] : B[k * a.ldb + n];
      rb[i] = v;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int idx = tid + i * NT;
      int mm, kk;
      if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      const unsigned short h = f2bf_rne(ra[i]);
      Ah[mm * LK + kk] = h;
      Al[mm * LK + kk] = f2bf_rne(ra[i] - bf2f(h));
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int idx = tid + i * NT;
      int nn, kk;
      if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
      const unsigned short h = f2bf_rne(rb[i]);
      Bh[nn * LK + kk] = h;
      Bl[nn * LK + kk] = f2bf_rne(rb[i] - bf2f(h));
    }
  };

  gload(0);
  for (long k0 = 0; k0 < a.K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < a.K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kc = kk + 8 * (lane >> 5);
      bf8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int row = wm * WTM + i * 32 + (lane & 31);
        us8 h = *reinterpret_cast<const us8*>(&Ah[row * LK + kc]);
        us8 l = *reinterpret_cast<const us8*>(&Al[row * LK + kc]);
        ah[i] = __builtin_bit_cast(bf8, h);
        al[i] = __builtin_bit_cast(bf8, l);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = wn * WTN + j * 32 + (lane & 31);
        us8 h = *reinterpret_cast<const us8*>(&Bh[col * LK + kc]);
        us8 l = *reinterpret_cast<const us8*>(&Bl[col * LK + kc]);
        bh[j] = __builtin_bit_cast(bf8, h);
        bl[j] = __builtin_bit_cast(bf8, l);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long m = m0 + wm * WTM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long n = n0 + wn * WTN + j * 32 + (lane & 31);
        if (m < a.M && n < a.N) {
          float v = (float)a.alpha * acc[i][j][r];
          if (a.beta != 0.0) v += (float)a.beta * C[m * a.ldc + n];
          C[m * a.ldc + n] = v;
        }
      }
}

// ------------------------------------------------------------------------
// Complex FP64 (interleaved re/im).  LDS holds re and im planes; each complex
// multiply-accumulate is four f64 MFMAs into re/im accumulators.
// ------------------------------------------------------------------------
template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_z64_kernel(GArgs a) {
  constexpr int NT = 256;
  constexpr int LA = BM + 2, LB = BN + 2;
  constexpr int AE = BM * BK / NT, BE = BK * BN / NT;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  __shared__ double Ar[BK * LA], Ai[BK * LA];
  __shared__ double Br[BK * LB], Bi[BK * LB];

  const int ntm = (int)((a.M + BM - 1) / BM), ntn = (int)((a.N + BN - 1) / BN);
  int tm, tn;
  tile_coords(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const long bz = blockIdx.z;
  const double2* A = reinterpret_cast<const double2*>(a.A) + bz * a.sA;
  const double2* B = reinterpret_cast<const double2*>(a.B) + bz * a.sB;
  double2* C = reinterpret_cast<double2*>(a.C) + bz * a.sC;
  const double sa = a.cA ? -1.0 : 1.0, sb = a.cB ? -1.0 : 1.0;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  d4 accr[FM][FN], acci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) { accr[i][j] = d4{0, 0, 0, 0}; acci[i][j] = d4{0, 0, 0, 0}; }

  double2 ra[AE], rb[BE];
  auto gload = [&](long k0) {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int idx = tid + i * NT;
      int mm, kk;
      if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      const long m = m0 + mm, k = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (m < a.M && k < a.K) v = a.tA ? A[k * a.lda + m] : A[m * a.lda + k];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int idx = tid + i * NT;
      int nn, kk;
      if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
      const long n = n0 + nn, k = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (n < a.N && k < a.K) v = a.tB ? B[n * a.ldb + k] : B[k * a.ldb + n];
      rb[i] = v;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int idx = tid + i * NT;
      int mm, kk;
      if (!a.tA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      Ar[kk * LA + mm] = ra[i].x;
      Ai[kk * LA + mm] = sa * ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int idx = tid + i * NT;
      int nn, kk;
      if (!a.tB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
      Br[kk * LB + nn] = rb[i].x;
      Bi[kk * LB + nn] = sb * rb[i].y;
    }
  };

  gload(0);
  for (long k0 = 0; k0 < a.K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < a.K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double far_[FM], fai[FM], fbr[FN], fbi[FN], fbin[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        far_[i] = Ar[kr * LA + wm * WTM + i * 16 + (lane & 15)];
        fai[i] = Ai[kr * LA + wm * WTM + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        fbr[j] = Br[kr * LB + wn * WTN + j * 16 + (lane & 15)];
        fbi[j] = Bi[kr * LB + wn * WTN + j * 16 + (lane & 15)];
        fbin[j] = -fbi[j];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(far_[i], fbr[j], accr[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fai[i], fbin[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(far_[i], fbi[j], acci[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fai[i], fbr[j], acci[i][j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * WTM + i * 16 + (lane >> 4) + 4 * r;
        const long n = n0 + wn * WTN + j * 16 + (lane & 15);
        if (m < a.M && n < a.N) {
          double vr = a.alpha * accr[i][j][r];
          double vi = a.alpha * acci[i][j][r];
          if (a.beta != 0.0) {
            vr += a.beta * C[m * a.ldc + n].x;
            vi += a.beta * C[m * a.ldc + n].y;
          }
          C[m * a.ldc + n] = make_double2(vr, vi);
        }
      }
}

template <int BM, int BN>
dim3 grid_for(const GemmDesc& g) {
  const long ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
  return dim3((unsigned)(ntm * ntn), 1, (unsigned)g.batch);
}

GArgs to_args(const GemmDesc& g) {
  GArgs a;
  a.M = g.M; a.N = g.N; a.K = g.K;
  a.A = g.A; a.lda = g.lda; a.sA = g.sA; a.tA = g.tA; a.cA = g.cA;
  a.B = g.B; a.ldb = g.ldb; a.sB = g.sB; a.tB = g.tB; a.cB = g.cB;
  a.C = g.C; a.ldc = g.ldc; a.sC = g.sC;
  a.alpha = g.alpha; a.beta = g.beta;
  return a;
}


// ------------------------------------------------------------------------
// Grouped FP64 GEMM: output group g (grid y) and batch z (grid z) accumulate
// sum over the terms t of group g of w_t A_{ia_t} B_{ib_t}, every term an
// M x N x K product; the output tile is written once (no read of C).
// ------------------------------------------------------------------------
struct GroupedArgs {
  long M, N, K;
  const double* A; long lda, sA, tsA;  // A_t = A + ia_t tsA + z sA
  const double* B; long ldb, sB, tsB;  // B_t = B + ib_t tsB + z sB
  double* C; long ldc, sC, gsC;        // C_g = C + g gsC + z sC
  const int* off; const int* ia; const int* ib; const double* w;
};

template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_f64_grouped_kernel(GroupedArgs ga) {
  using T = double;
  using TT = RealTile<T, BM, BN, BK>;
  constexpr int LA = BM + TT::PAD, LB = BN + TT::PAD;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  __shared__ T As[BK * LA];
  __shared__ T Bs[BK * LB];
  const int ntm = (int)((ga.M + BM - 1) / BM), ntn = (int)((ga.N + BN - 1) / BN);
  int tm, tn;
  tile_coords(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const int g = blockIdx.y;
  const long z = blockIdx.z;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  GArgs a;
  a.M = ga.M; a.N = ga.N; a.K = ga.K;
  a.lda = ga.lda; a.tA = 0; a.cA = 0;
  a.ldb = ga.ldb; a.tB = 0; a.cB = 0;
  d4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0, 0, 0, 0};
  const int t1 = ga.off[g + 1];
  for (int t = ga.off[g]; t < t1; ++t) {
    const T* A = ga.A + (long)ga.ia[t] * ga.tsA + z * ga.sA;
    const T* B = ga.B + (long)ga.ib[t] * ga.tsB + z * ga.sB;
    const double wt = ga.w[t];
    T ra[TT::AE], rb[TT::BE];
    load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, 0, ra, rb);
    for (long k0 = 0; k0 < a.K; k0 += BK) {
#pragma unroll
      for (int i = 0; i < TT::AE; ++i) ra[i] *= wt;
      __syncthreads();
      store_tiles<T, BM, BN, BK>(a, As, Bs, ra, rb);
      __syncthreads();
      if (k0 + BK < a.K) load_tiles<T, BM, BN, BK>(a, A, B, m0, n0, k0 + BK, ra, rb);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = As[kr * LA + wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = Bs[kr * LB + wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  T* C = ga.C + (long)g * ga.gsC + z * ga.sC;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * WTM + i * 16 + (lane >> 4) + 4 * r;
        const long n = n0 + wn * WTN + j * 16 + (lane & 15);
        if (m < a.M && n < a.N) C[m * ga.ldc + n] = acc[i][j][r];
      }
}

// ------------------------------------------------------------------------
// Block-sparse GEMM over uneven blocks (QN sectors): one launch covers a list
// of output tiles; tile t belongs to task tiles[t].x, whose output block
// C + c_off (M x N, ld ldc) is the sum over the task's K segments of
// A_s (M x K_s, ld lda) * B_s (K_s x N, ld ldb), A_s = A + a_off, B_s = B +
// b_off.  A task without segments writes zeros.  The tile list is laid out
// XCD-major: workgroup id i runs list entry (i % 8) * per + i / 8, so the
// contiguous run of tiles one XCD runs shares its tasks' A and B panels in
// that XCD's L2; padding entries (task < 0) exit at once.
// ------------------------------------------------------------------------
struct BsArgs {
  const void* A; long lda;
  const void* B; long ldb; int tB;
  void* C; long ldc;
  const BsTask* tasks;
  const BsSeg* segs;
  const int4* tiles;
  long per;
  double alpha;
};

// SPLIT (FP32 operands only): each FP32 fragment is split into bf16 hi/lo
// parts in registers and the product accumulates lo*hi + hi*lo + hi*hi with
// the bf16 MFMA 32x32x16 (BK = 16: one step per staged K chunk).
template <typename T, int BM, int BN, int BK, bool SPLIT = false>
__global__ __launch_bounds__(256) void gemm_bs_kernel(BsArgs a) {
  using TT = RealTile<T, BM, BN, BK>;
  constexpr bool F64 = sizeof(T) == 8;
  static_assert(!SPLIT || (!F64 && BK == 16), "the split product takes FP32 operands in K chunks of 16");
  constexpr int LA = BM + TT::PAD, LB = BN + TT::PAD;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int MB = F64 ? 16 : 32;  // MFMA block edge
  constexpr int FM = WTM / MB, FN = WTN / MB;
  typedef typename std::conditional<F64, d4, f16v>::type Acc;
  __shared__ T As[BK * LA];
  __shared__ T Bs[BK * LB];
  const long id = (long)(blockIdx.x % 8) * a.per + blockIdx.x / 8;
  const int4 tl = a.tiles[id];
  if (tl.x < 0) return;
  const BsTask tk = a.tasks[tl.x];
  const long m0 = tl.y, n0 = tl.z;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  GArgs g;
  g.M = tk.M; g.N = tk.N; g.K = 0;
  g.lda = a.lda; g.tA = 0; g.cA = 0;
  g.ldb = a.ldb; g.tB = a.tB; g.cB = 0;
  Acc acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = Acc{};
  for (int sgi = tk.seg0; sgi < tk.seg0 + tk.nseg; ++sgi) {
    const BsSeg sg = a.segs[sgi];
    g.K = sg.K;
    const T* A = reinterpret_cast<const T*>(a.A) + sg.a_off;
    const T* B = reinterpret_cast<const T*>(a.B) + sg.b_off;
    T ra[TT::AE], rb[TT::BE];
    load_tiles<T, BM, BN, BK>(g, A, B, m0, n0, 0, ra, rb);
    for (long k0 = 0; k0 < g.K; k0 += BK) {
      __syncthreads();
      store_tiles<T, BM, BN, BK>(g, As, Bs, ra, rb);
      __syncthreads();
      if (k0 + BK < g.K) load_tiles<T, BM, BN, BK>(g, A, B, m0, n0, k0 + BK, ra, rb);
      if constexpr (F64) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
          const int kr = kk + (lane >> 4);
          double fa[FM], fb[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) fa[i] = As[kr * LA + wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < FN; ++j) fb[j] = Bs[kr * LB + wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
      } else if constexpr (SPLIT) {
        // lane: row / column lane % 32, the 8 k values 8 (lane / 32) .. + 7
        const int kb = 8 * (lane >> 5);
        bf8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          us8 h, l;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float x = As[(kb + q) * LA + wm * WTM + i * 32 + (lane & 31)];
            h[q] = f2bf_rne(x);
            l[q] = f2bf_rne(x - bf2f(h[q]));
          }
          ah[i] = __builtin_bit_cast(bf8, h);
          al[i] = __builtin_bit_cast(bf8, l);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          us8 h, l;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float x = Bs[(kb + q) * LB + wn * WTN + j * 32 + (lane & 31)];
            h[q] = f2bf_rne(x);
            l[q] = f2bf_rne(x - bf2f(h[q]));
          }
          bh[j] = __builtin_bit_cast(bf8, h);
          bl[j] = __builtin_bit_cast(bf8, l);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
      } else {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
          const int kr = kk + (lane >> 5);
          float fa[FM], fb[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) fa[i] = As[kr * LA + wm * WTM + i * 32 + (lane & 31)];
#pragma unroll
          for (int j = 0; j < FN; ++j) fb[j] = Bs[kr * LB + wn * WTN + j * 32 + (lane & 31)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  T* C = reinterpret_cast<T*>(a.C) + tk.c_off;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      if constexpr (F64) {
        // register r of lane: row lane/16 + 4r, column lane%16 of the 16x16 block
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = m0 + wm * WTM + i * 16 + (lane >> 4) + 4 * r;
          const long n = n0 + wn * WTN + j * 16 + (lane & 15);
          if (m < g.M && n < g.N) C[m * a.ldc + n] = a.alpha * acc[i][j][r];
        }
      } else {
        // register r of lane: row 8(r/4) + 4(lane/32) + r%4, column lane%32
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const long m = m0 + wm * WTM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const long n = n0 + wn * WTN + j * 32 + (lane & 31);
          if (m < g.M && n < g.N) C[m * a.ldc + n] = (float)a.alpha * acc[i][j][r];
        }
      }
    }
}

}  // namespace

namespace {
// C = alpha * sum_p part[p] + beta * C   (part: P slices of M x N, ld N)
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const T* __restrict__ part, int P, long M, long N,
                                                            double alpha, double beta, T* __restrict__ C, long ldc) {
  const long tot = M * N;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    double acc = 0.0;
    for (int q = 0; q < P; ++q) acc += (double)part[(long)q * tot + i];
    const long r = i / N, col = i - r * N;
    T* dst = C + r * ldc + col;
    const double prev = (beta != 0.0) ? beta * (double)*dst : 0.0;
    *dst = (T)(alpha * acc + prev);
  }
}
}  // namespace

static void gemm_impl(tn_ctx* c, Prec p, const GemmDesc& g);

// Few output tiles with a long K (K-reductions over whole tensors, Gram
// matrices of tall matrices) leave most of the chip idle: split K into
// chunks computed as one batched GEMM into FP64/FP32 partials, then reduce.
// Chunk length along K of the split (0: no split).
static long split_chunk(Prec p, const GemmDesc& g) {
  if (!(p == Prec::F64 || p == Prec::F32) || g.batch != 1) return 0;
  const long tiles = ((g.M + 127) / 128) * ((g.N + 127) / 128);
  long P0;
  if (tiles < 64 && g.K >= 4096) {
    P0 = std::min<long>(g.K / 512, 256 / tiles);
  } else {
    // mid-size products (the eigensolver's merges and compact-WY updates at
    // p of a few hundred): fewer than 64 tiles of 64 x 64 and K >= 256 run as
    // chunks of at least 64 along K
    const long t64 = ((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (t64 >= 64 || g.K < 256) return 0;
    P0 = std::min<long>(g.K / 64, 128 / t64);
  }
  if (P0 < 2) return 0;
  return ((g.K + P0 - 1) / P0 + 15) / 16 * 16;
}

size_t gemm_split_bytes(Prec p, const GemmDesc& g) {
  if (g.M <= 0 || g.N <= 0) return 0;
  const long Kc = split_chunk(p, g);
  if (Kc == 0) return 0;
  const long P = (g.K + Kc - 1) / Kc;
  return (size_t)P * g.M * g.N * ((p == Prec::F64) ? 8 : 4);
}

static bool gemm_split_k(tn_ctx* c, Prec p, const GemmDesc& g) {
  const long Kc = split_chunk(p, g);
  if (Kc == 0) return false;
  const long P = (g.K + Kc - 1) / Kc;
  const long full = g.K / Kc;  // chunks of exactly Kc
  const size_t es = (p == Prec::F64) ? 8 : 4;
  const size_t need = (size_t)P * g.M * g.N * es;
  char* part;
  if (g.split_ws) {
    TN_REQUIRE(need <= g.split_cap, TN_E_INVALID, "gemm: split-K buffer too small");
    part = (char*)g.split_ws;
  } else {
    part = (char*)c->workspace(WS_GEMM_SPLIT, need);
  }
  GemmDesc q = g;
  q.alpha = 1.0;
  q.beta = 0.0;
  q.K = Kc;
  q.C = part;
  q.ldc = g.N;
  q.sC = g.M * g.N;
  q.sA = g.tA ? Kc * g.lda : Kc;
  q.sB = g.tB ? Kc : Kc * g.ldb;
  q.batch = (int)full;
  gemm_impl(c, p, q);
  if (full < P) {  // the shorter last chunk
    GemmDesc r = q;
    r.batch = 1;
    r.K = g.K - full * Kc;
    r.A = (const char*)g.A + (size_t)(full * q.sA) * es;
    r.B = (const char*)g.B + (size_t)(full * q.sB) * es;
    r.C = part + (size_t)(full * q.sC) * es;
    gemm_impl(c, p, r);
  }
  const long tot = g.M * g.N;
  const unsigned grid = (unsigned)std::max<long>(1, std::min<long>(2048, (tot + 255) / 256));
  if (p == Prec::F64)
    hipLaunchKernelGGL(splitk_reduce_kernel<double>, dim3(grid), dim3(256), 0, c->stream, (const double*)part, (int)P,
                       g.M, g.N, g.alpha, g.beta, (double*)g.C, g.ldc);
  else
    hipLaunchKernelGGL(splitk_reduce_kernel<float>, dim3(grid), dim3(256), 0, c->stream, (const float*)part, (int)P,
                       g.M, g.N, g.alpha, g.beta, (float*)g.C, g.ldc);
  TN_HIP(hipGetLastError());
  return true;
}

void gemm(tn_ctx* c, Prec p, const GemmDesc& g) {
  TN_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0 && g.batch >= 1, TN_E_INVALID, "gemm: bad sizes");
  if (g.M == 0 || g.N == 0) return;
  TN_REQUIRE(g.A && g.B && g.C, TN_E_INVALID, "gemm: null operand");
  TN_REQUIRE(g.batch <= 65535, TN_E_INVALID, "gemm: batch too large");
  if (gemm_split_k(c, p, g)) return;
  gemm_impl(c, p, g);
}

static void gemm_impl(tn_ctx* c, Prec p, const GemmDesc& g) {
  // fewer 128 x 128 output tiles than CUs: 64 x 64 tiles spread the work
  // over four times as many workgroups (mid-size bonds are latency-bound)
  const long t128 = ((g.M + 127) / 128) * ((g.N + 127) / 128) * (long)g.batch;
  const bool few_tiles = t128 < 240;
  if (!few_tiles && gemm_fast_ok(p, g)) {
    gemm_fast(c, p, g);
    return;
  }
  GArgs a = to_args(g);
  const bool small = (g.M * g.N <= 256L * 256L) || few_tiles;
  switch (p) {
    case Prec::F64:
      if (small && ((g.M + 63) / 64) * ((g.N + 63) / 64) * (long)g.batch < 64) {
        // few 64 x 64 tiles: 32 x 32 tiles, four times the workgroups
        dim3 gr = grid_for<32, 32>(g);
        hipLaunchKernelGGL((gemm_f64_kernel<32, 32, 16>), gr, dim3(256), 0, c->stream, a);
      } else if (small) {
        dim3 gr = grid_for<64, 64>(g);
        hipLaunchKernelGGL((gemm_f64_kernel<64, 64, 16>), gr, dim3(256), 0, c->stream, a);
      } else {
        dim3 gr = grid_for<128, 128>(g);
        hipLaunchKernelGGL((gemm_f64_kernel<128, 128, 16>), gr, dim3(256), 0, c->stream, a);
      }
      break;
    case Prec::F32:
      if (small) {
        dim3 gr = grid_for<64, 64>(g);
        hipLaunchKernelGGL((gemm_f32_kernel<64, 64, 16>), gr, dim3(256), 0, c->stream, a);
      } else {
        dim3 gr = grid_for<128, 128>(g);
        hipLaunchKernelGGL((gemm_f32_kernel<128, 128, 16>), gr, dim3(256), 0, c->stream, a);
      }
      break;
    case Prec::BF16X3:
      if (small) {
        dim3 gr = grid_for<64, 64>(g);
        hipLaunchKernelGGL((gemm_bf16x3_kernel<64, 64, 32>), gr, dim3(256), 0, c->stream, a);
      } else {
        dim3 gr = grid_for<128, 128>(g);
        hipLaunchKernelGGL((gemm_bf16x3_kernel<128, 128, 32>), gr, dim3(256), 0, c->stream, a);
      }
      break;
    case Prec::Z64: {
      // small complex products (TDVP at small bond dimension): 32 x 32 tiles
      // put four times as many workgroups on the MFMA work of each tile
      const long t64 = ((g.M + 63) / 64) * ((g.N + 63) / 64) * (long)g.batch;
      if (t64 < 64) {
        dim3 gr = grid_for<32, 32>(g);
        hipLaunchKernelGGL((gemm_z64_kernel<32, 32, 8>), gr, dim3(256), 0, c->stream, a);
      } else {
        dim3 gr = grid_for<64, 64>(g);
        hipLaunchKernelGGL((gemm_z64_kernel<64, 64, 8>), gr, dim3(256), 0, c->stream, a);
      }
      break;
    }
  }
  TN_HIP(hipGetLastError());
}


void gemm_grouped_f64(tn_ctx* c, const GroupedDesc& d) {
  if (d.M <= 0 || d.N <= 0 || d.ngroups <= 0 || d.batch <= 0) return;
  TN_REQUIRE(d.ngroups <= 65535 && d.batch <= 65535, TN_E_INVALID, "grouped gemm: grid too large");
  GroupedArgs ga;
  ga.M = d.M; ga.N = d.N; ga.K = d.K;
  ga.A = d.A; ga.lda = d.lda; ga.sA = d.sA; ga.tsA = d.tsA;
  ga.B = d.B; ga.ldb = d.ldb; ga.sB = d.sB; ga.tsB = d.tsB;
  ga.C = d.C; ga.ldc = d.ldc; ga.sC = d.sC; ga.gsC = d.gsC;
  ga.off = d.off; ga.ia = d.ia; ga.ib = d.ib; ga.w = d.w;
  const long tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64);
  TN_REQUIRE(tiles < (1L << 31), TN_E_INVALID, "grouped gemm: too many tiles");
  dim3 gr((unsigned)tiles, (unsigned)d.ngroups, (unsigned)d.batch);
  hipLaunchKernelGGL((gemm_f64_grouped_kernel<64, 64, 16>), gr, dim3(256), 0, c->stream, ga);
  TN_HIP(hipGetLastError());
}

void gemm_blocksparse(tn_ctx* c, Prec p, const BsDesc& d) {
  if (d.ntiles <= 0) return;
  TN_REQUIRE(p == Prec::F64 || p == Prec::F32 || p == Prec::BF16X3, TN_E_INVALID,
             "block-sparse gemm: FP64, FP32 or BF16x3 (FP32 operands) only");
  TN_REQUIRE(d.ntiles % 8 == 0 && d.ntiles < (1L << 31), TN_E_INVALID, "block-sparse gemm: bad tile list");
  BsArgs a;
  a.A = d.A; a.lda = d.lda;
  a.B = d.B; a.ldb = d.ldb; a.tB = d.tB;
  a.C = d.C; a.ldc = d.ldc;
  a.tasks = d.tasks; a.segs = d.segs; a.tiles = d.tiles;
  a.per = d.ntiles / 8;
  a.alpha = d.alpha;
  if (p == Prec::F64)
    hipLaunchKernelGGL((gemm_bs_kernel<double, BS_TILE, BS_TILE, 16>), dim3((unsigned)d.ntiles), dim3(256), 0,
                       c->stream, a);
  else if (p == Prec::F32)
    hipLaunchKernelGGL((gemm_bs_kernel<float, BS_TILE, BS_TILE, 16>), dim3((unsigned)d.ntiles), dim3(256), 0,
                       c->stream, a);
  else
    hipLaunchKernelGGL((gemm_bs_kernel<float, BS_TILE, BS_TILE, 16, true>), dim3((unsigned)d.ntiles), dim3(256), 0,
                       c->stream, a);
  TN_HIP(hipGetLastError());
}

}  // namespace tn
