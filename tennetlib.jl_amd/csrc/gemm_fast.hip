// Double-buffered MFMA GEMMs for the large contractions of the hot path.
//
//  * gemm_fast<T, TA, TB>: FP64 (v_mfma_f64_16x16x4_f64) or FP32
//    (v_mfma_f32_32x32x2_f32), 128 x 128 tiles, 4 wavefronts of 64 x 64.
//    Each operand is staged with 16-byte loads along its contiguous
//    dimension into an LDS tile that keeps the global orientation
//    (A: [m][k] if k is contiguous, [k][m] if m is; B: [k][n] or [n][k]),
//    so staging never transposes; the MFMA fragment reads adapt instead.
//    Two LDS stages: the next tile is fetched into registers while the
//    current one feeds the MFMAs, then written to the other stage; one
//    barrier per K tile.
//  * gemm_bf16x3_nt: C = A * B^T with A (M x K) and B (N x K) given as bf16
//    hi/lo pairs (k contiguous, pre-split by split_bf16_*).  Each 32x32x16
//    step issues lo*hi + hi*lo + hi*hi with FP32 accumulation.
// Row strides must be multiples of the 16-byte vector width and bases
// 16-byte aligned; the dispatcher in gemm.hip checks this and otherwise
// uses the generic kernels.
#include <map>
#include <mutex>

#include "tn_internal.h"

#include <algorithm>
#include <cstdlib>

namespace tn {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

struct FArgs {
  long M, N, K;
  const void* A; long lda; long sA; long wA;  // wA, wB: readable row widths
  const void* B; long ldb; long sB; long wB;
  void* C; long ldc; long sC;
  double alpha, beta;
};

__device__ __forceinline__ void tile_coords2(int ntm, int ntn, int& tm, int& tn) {
  const int nt = ntm * ntn;
  int id = blockIdx.x;
  if (nt >= 64 && (nt % 8) == 0) {  // contiguous tile runs per XCD
    const int per = nt / 8;
    id = (id % 8) * per + (id / 8);
  }
  const int GM = 8;
  const int group = id / (GM * ntn);
  const int first_m = group * GM;
  const int gsz = min(ntm - first_m, GM);
  const int in_group = id % (GM * ntn);
  tm = first_m + (in_group % gsz);
  tn = in_group / gsz;
}

template <typename T> struct Vec;
template <> struct Vec<double> { typedef double2 type; static constexpr int n = 2; };
template <> struct Vec<float> { typedef float4 type; static constexpr int n = 4; };

// ---------------------------------------------------------------------------
// Tile loads through buffer resources.  A resource covers the rows of one
// tile of a row-major operand, from its first row to the last valid element
// ((rows_valid - 1) * ld + cols elements); the hardware returns zeros for any
// offset at or beyond that extent, and vectors past the last column get an
// offset far beyond it.  Tile edges therefore need no branches, the loads of
// the next tile stay in flight while the MFMAs run, and no load touches
// memory outside the operand.  Every row's valid width (cols) is a multiple
// of the 16-byte vector, so a vector is either fully valid or fully outside.
// ---------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr unsigned OOB = 0x80000000u;  // offset past every extent (< 2^31 bytes)

__device__ __forceinline__ Rsrc tile_rsrc(const void* base, long r0, long rows, long tile_rows, long ld, long cols,
                                          int esz) {
  const long rv = min(tile_rows, rows - r0);
  const long bytes = rv > 0 ? ((rv - 1) * ld + cols) * esz : 0;
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void*>(reinterpret_cast<const void*>(reinterpret_cast<const char*>(base) + r0 * ld * esz)),
      (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 ld16(Rsrc r, unsigned off) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// byte offset of the vector at (tile row rr, column c) or OOB past the columns
__device__ __forceinline__ unsigned voff(int rr, long c, long ld, long cols, int esz) {
  return c < cols ? (unsigned)((rr * ld + c) * esz) : OOB;
}
// the VN elements at (tile row rr, columns c .. c + VN - 1) as one 16-byte
// value: AL = 1, one vector load (the row start and c are vector-aligned and
// the row width is whole vectors); AL = 0, element loads, each bounded by the
// row width on its own, for operands whose rows are not 16-byte aligned
template <typename T, int AL>
__device__ __forceinline__ uint4 ldvec(Rsrc r, int rr, long c, long ld, long cols) {
  constexpr int ES = (int)sizeof(T);
  if constexpr (AL) {
    return ld16(r, voff(rr, c, ld, cols, ES));
  } else if constexpr (ES == 8) {
    uint4 v;
    const unsigned long long x0 = __builtin_bit_cast(
        unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, voff(rr, c, ld, cols, ES), 0, 0));
    const unsigned long long x1 = __builtin_bit_cast(
        unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, voff(rr, c + 1, ld, cols, ES), 0, 0));
    v.x = (unsigned)x0; v.y = (unsigned)(x0 >> 32);
    v.z = (unsigned)x1; v.w = (unsigned)(x1 >> 32);
    return v;
  } else {
    uint4 v;
    v.x = __builtin_amdgcn_raw_buffer_load_b32(r, voff(rr, c, ld, cols, ES), 0, 0);
    v.y = __builtin_amdgcn_raw_buffer_load_b32(r, voff(rr, c + 1, ld, cols, ES), 0, 0);
    v.z = __builtin_amdgcn_raw_buffer_load_b32(r, voff(rr, c + 2, ld, cols, ES), 0, 0);
    v.w = __builtin_amdgcn_raw_buffer_load_b32(r, voff(rr, c + 3, ld, cols, ES), 0, 0);
    return v;
  }
}

template <typename T, int TA, int TB>
struct FastCfg {
  static constexpr int BM = 128, BN = 128;
  static constexpr int BK = (sizeof(T) == 8) ? 16 : 32;
  static constexpr int VN = Vec<T>::n;
  static constexpr int PAD = VN;  // keeps 16-byte alignment of every LDS row
  // A stage: [BM][BK+PAD] if TA == 0 else [BK][BM+PAD]
  static constexpr int A_ROWS = TA ? BK : BM, A_COLS = TA ? BM : BK, A_LD = A_COLS + PAD;
  static constexpr int B_ROWS = TB ? BN : BK, B_COLS = TB ? BK : BN, B_LD = B_COLS + PAD;
  static constexpr int A_VECS = A_ROWS * A_COLS / VN / 256;  // per thread
  static constexpr int B_VECS = B_ROWS * B_COLS / VN / 256;
  static constexpr int A_ELEMS = A_ROWS * A_LD, B_ELEMS = B_ROWS * B_LD;
};

template <typename T, int TA, int TB, int AL>
__global__ __launch_bounds__(256) void gemm_fast_kernel(FArgs a) {
  using CF = FastCfg<T, TA, TB>;
  constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, VN = CF::VN;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T* As = reinterpret_cast<T*>(smem);                 // 2 stages
  T* Bs = As + 2 * CF::A_ELEMS;

  const int ntm = (int)((a.M + BM - 1) / BM), ntn = (int)((a.N + BN - 1) / BN);
  int tm, tn;
  tile_coords2(ntm, ntn, tm, tn);
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const long bz = blockIdx.z;
  const T* A = reinterpret_cast<const T*>(a.A) + bz * a.sA;
  const T* B = reinterpret_cast<const T*>(a.B) + bz * a.sB;
  T* C = reinterpret_cast<T*>(a.C) + bz * a.sC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  constexpr int ES = (int)sizeof(T);
  uint4 ra[CF::A_VECS], rb[CF::B_VECS];
  // tile k0 of A: rows m (TA = 0, k contiguous) or rows k (TA = 1, m contiguous);
  // a k0 past the end loads zeros (the unconditional prefetch of the last step)
  auto gload = [&](long k0) {
    const Rsrc ra_r = TA ? tile_rsrc(A, k0, a.K, BK, a.lda, a.wA, ES) : tile_rsrc(A, m0, a.M, BM, a.lda, a.wA, ES);
    const Rsrc rb_r = TB ? tile_rsrc(B, n0, a.N, BN, a.ldb, a.wB, ES) : tile_rsrc(B, k0, a.K, BK, a.ldb, a.wB, ES);
#pragma unroll
    for (int i = 0; i < CF::A_VECS; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / (CF::A_COLS / VN), cc = (idx % (CF::A_COLS / VN)) * VN;
      ra[i] = TA ? ldvec<T, AL>(ra_r, rr, m0 + cc, a.lda, a.wA) : ldvec<T, AL>(ra_r, rr, k0 + cc, a.lda, a.wA);
    }
#pragma unroll
    for (int i = 0; i < CF::B_VECS; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / (CF::B_COLS / VN), cc = (idx % (CF::B_COLS / VN)) * VN;
      rb[i] = TB ? ldvec<T, AL>(rb_r, rr, k0 + cc, a.ldb, a.wB) : ldvec<T, AL>(rb_r, rr, n0 + cc, a.ldb, a.wB);
    }
  };
  auto sstore = [&](int st) {
    T* as = As + st * CF::A_ELEMS;
    T* bs = Bs + st * CF::B_ELEMS;
#pragma unroll
    for (int i = 0; i < CF::A_VECS; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / (CF::A_COLS / VN), cc = (idx % (CF::A_COLS / VN)) * VN;
      *reinterpret_cast<uint4*>(&as[rr * CF::A_LD + cc]) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < CF::B_VECS; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / (CF::B_COLS / VN), cc = (idx % (CF::B_COLS / VN)) * VN;
      *reinterpret_cast<uint4*>(&bs[rr * CF::B_LD + cc]) = rb[i];
    }
  };
  // element (m, k) of the A stage and (k, n) of the B stage
  auto Ael = [&](const T* as, int m, int k) -> T { return TA ? as[k * CF::A_LD + m] : as[m * CF::A_LD + k]; };
  auto Bel = [&](const T* bs, int k, int n) -> T { return TB ? bs[n * CF::B_LD + k] : bs[k * CF::B_LD + n]; };

  const long nk = (a.K + BK - 1) / BK;
  if constexpr (sizeof(T) == 8) {
    constexpr int FM = 4, FN = 4;  // 64x64 wave tile of 16x16 blocks
    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = d4{0, 0, 0, 0};
    gload(0);
    sstore(0);
    __syncthreads();
    for (long t = 0; t < nk; ++t) {
      gload((t + 1) * BK);
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
      const T* as = As + (t & 1) * CF::A_ELEMS;
      const T* bs = Bs + (t & 1) * CF::B_ELEMS;
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = Ael(as, wm * 64 + i * 16 + (lane & 15), kr);
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = Bel(bs, kr, wn * 64 + j * 16 + (lane & 15));
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
      sstore((t + 1) & 1);
      __syncthreads();
    }
    // v_mfma_f64_16x16x4f64 output: register r of lane -> row lane/16 + 4r, column lane%16
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
          const long n = n0 + wn * 64 + j * 16 + (lane & 15);
          if (m < a.M && n < a.N) {
            double v = a.alpha * acc[i][j][r];
            if (a.beta != 0.0) v += a.beta * C[m * a.ldc + n];
            C[m * a.ldc + n] = (T)v;
          }
        }
  } else {
    constexpr int FM = 2, FN = 2;  // 64x64 wave tile of 32x32 blocks
    f16v acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = f16v{};
    gload(0);
    sstore(0);
    __syncthreads();
    for (long t = 0; t < nk; ++t) {
      gload((t + 1) * BK);
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
      const T* as = As + (t & 1) * CF::A_ELEMS;
      const T* bs = Bs + (t & 1) * CF::B_ELEMS;
#pragma unroll
      for (int kk = 0; kk < BK; kk += 2) {
        const int kr = kk + (lane >> 5);
        float fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = Ael(as, wm * 64 + i * 32 + (lane & 31), kr);
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = Bel(bs, kr, wn * 64 + j * 32 + (lane & 31));
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
      sstore((t + 1) & 1);
      __syncthreads();
    }
    // register r of lane -> row 8*(r/4) + 4*(lane/32) + r%4, column lane%32
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const long m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const long n = n0 + wn * 64 + j * 32 + (lane & 31);
          if (m < a.M && n < a.N) {
            float v = (float)a.alpha * acc[i][j][r];
            if (a.beta != 0.0) v += (float)a.beta * C[m * a.ldc + n];
            C[m * a.ldc + n] = v;
          }
        }
  }
}

// ---------------------------------------------------------------------------
// BF16x3, C (float) = A * B^T, A = (Ah + Al) M x K, B = (Bh + Bl) N x K
// ---------------------------------------------------------------------------
struct SArgs {
  long M, N, K;
  const unsigned short* Ah; const unsigned short* Al; long lda;
  const unsigned short* Bh; const unsigned short* Bl; long ldb;
  float* C; long ldc;
};

// 128 x 128 output tile, 4 wavefronts of 64 x 64 (2 x 2 blocks of 32 x 32);
// K tile SBK = 32 (two 80 KiB stages: two workgroups per CU) or 64 (one).
constexpr int SBM = 128, SBN = 128;
template <int SBK> struct SCfg {
  static constexpr int SLD = SBK + 8;                      // bf16 per LDS row (16 B pad)
  static constexpr int STAGE = (2 * SBM + 2 * SBN) * SLD;  // hi+lo of A and B, bf16 elements
  static constexpr int VPR = SBK / 8;                      // 16-byte vectors per row
  static constexpr int NV = SBM * VPR / 256;               // per thread, per operand part
};

template <int SBK>
__global__ __launch_bounds__(256) void gemm_bf16x3_nt_kernel(SArgs a) {
  using CF = SCfg<SBK>;
  constexpr int SLD = CF::SLD, VPR = CF::VPR, NV = CF::NV;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  unsigned short* S = reinterpret_cast<unsigned short*>(smem);
  const int ntm = (int)((a.M + SBM - 1) / SBM), ntn = (int)((a.N + SBN - 1) / SBN);
  int tm, tn;
  tile_coords2(ntm, ntn, tm, tn);
  const long m0 = (long)tm * SBM, n0 = (long)tn * SBN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // rows are zero-padded to Kp (a multiple of 8): vectors below Kp are valid
  const long Kp = (a.K + 7) & ~7L;
  const Rsrc rah_r = tile_rsrc(a.Ah, m0, a.M, SBM, a.lda, Kp, 2);
  const Rsrc ral_r = tile_rsrc(a.Al, m0, a.M, SBM, a.lda, Kp, 2);
  const Rsrc rbh_r = tile_rsrc(a.Bh, n0, a.N, SBN, a.ldb, Kp, 2);
  const Rsrc rbl_r = tile_rsrc(a.Bl, n0, a.N, SBN, a.ldb, Kp, 2);
  uint4 rah[NV], ral[NV], rbh[NV], rbl[NV];
  auto gload = [&](long k0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / VPR, cc = (idx % VPR) * 8;
      const unsigned oa = voff(rr, k0 + cc, a.lda, Kp, 2), ob = voff(rr, k0 + cc, a.ldb, Kp, 2);
      rah[i] = ld16(rah_r, oa);
      ral[i] = ld16(ral_r, oa);
      rbh[i] = ld16(rbh_r, ob);
      rbl[i] = ld16(rbl_r, ob);
    }
  };
  auto sstore = [&](int st) {
    unsigned short* Ah = S + st * CF::STAGE;
    unsigned short* Al = Ah + SBM * SLD;
    unsigned short* Bh = Al + SBM * SLD;
    unsigned short* Bl = Bh + SBN * SLD;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = tid + i * 256;
      const int rr = idx / VPR, cc = (idx % VPR) * 8;
      *reinterpret_cast<uint4*>(&Ah[rr * SLD + cc]) = rah[i];
      *reinterpret_cast<uint4*>(&Al[rr * SLD + cc]) = ral[i];
      *reinterpret_cast<uint4*>(&Bh[rr * SLD + cc]) = rbh[i];
      *reinterpret_cast<uint4*>(&Bl[rr * SLD + cc]) = rbl[i];
    }
  };
  f16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f16v{};
  const long nk = (a.K + SBK - 1) / SBK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (long t = 0; t < nk; ++t) {
    gload((t + 1) * SBK);  // past the last tile this loads zeros and touches no memory
    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
    const unsigned short* Ah = S + (t & 1) * CF::STAGE;
    const unsigned short* Al = Ah + SBM * SLD;
    const unsigned short* Bh = Al + SBM * SLD;
    const unsigned short* Bl = Bh + SBN * SLD;
#pragma unroll
    for (int kk = 0; kk < SBK; kk += 16) {
      const int kc = kk + 8 * (lane >> 5);
      bf8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wm * 64 + i * 32 + (lane & 31);
        ah[i] = __builtin_bit_cast(bf8, *reinterpret_cast<const us8*>(&Ah[row * SLD + kc]));
        al[i] = __builtin_bit_cast(bf8, *reinterpret_cast<const us8*>(&Al[row * SLD + kc]));
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = wn * 64 + j * 32 + (lane & 31);
        bh[j] = __builtin_bit_cast(bf8, *reinterpret_cast<const us8*>(&Bh[col * SLD + kc]));
        bl[j] = __builtin_bit_cast(bf8, *reinterpret_cast<const us8*>(&Bl[col * SLD + kc]));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    sstore((t + 1) & 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long n = n0 + wn * 64 + j * 32 + (lane & 31);
        if (m < a.M && n < a.N) a.C[m * a.ldc + n] = acc[i][j][r];
      }
}

// ---------------------------------------------------------------------------
// bf16 hi/lo split: hi = rne(x), lo = rne(x - hi); from FP64 or FP32 sources
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f_(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <typename T>
__device__ __forceinline__ void split1(T x, unsigned short& h, unsigned short& l) {
  h = f2bf((float)x);
  l = f2bf((float)(x - (T)bf2f_(h)));
}

// dst[r][c] (ld_dst) = split(src[r][c]); columns cols..ld_dst-1 zeroed
template <typename T>
__global__ __launch_bounds__(256) void split_rows_kernel(const T* __restrict__ src, long rows, long cols,
                                                         long lds, unsigned short* __restrict__ hi,
                                                         unsigned short* __restrict__ lo, long ldd) {
  const long total = rows * ldd;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldd, c = i % ldd;
    unsigned short h = 0, l = 0;
    if (c < cols) split1<T>(src[r * lds + c], h, l);
    hi[i] = h;
    lo[i] = l;
  }
}

// dst[c][r] (ld_dst) = split(src[r][c]) for r < rows, c < cols; 32x32 tiles
template <typename T>
__global__ __launch_bounds__(256) void split_transpose_kernel(const T* __restrict__ src, long rows, long cols,
                                                              long lds, unsigned short* __restrict__ hi,
                                                              unsigned short* __restrict__ lo, long ldd) {
  __shared__ T t[32][33];
  const long r0 = (long)blockIdx.y * 32, c0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const long r = r0 + k, c = c0 + tx;
    t[k][tx] = (r < rows && c < cols) ? src[r * lds + c] : T(0);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const long c = c0 + k, r = r0 + tx;  // output row c, column r
    if (c < cols && r < ldd) {
      unsigned short h = 0, l = 0;
      if (r < rows) split1<T>(t[tx][k], h, l);
      hi[c * ldd + r] = h;
      lo[c * ldd + r] = l;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
std::recursive_mutex& api_mutex() {
  static std::recursive_mutex m;
  return m;
}

// Every launch of a kernel with dynamic LDS goes through this first.  The
// first call for a kernel sets its dynamic-LDS limit once, to everything the
// 160 KB LDS leaves beside its static LDS, and it is never changed again, so
// no attribute change can meet a launch of the same kernel on another thread.
void raise_lds_limit(const void* fn, size_t bytes) {
  constexpr size_t LDS_BYTES = 160 * 1024;
  static std::mutex mu;
  static std::map<const void*, size_t> done;  // kernel -> dynamic-LDS limit
  std::lock_guard<std::mutex> lock(mu);
  auto it = done.find(fn);
  if (it == done.end()) {
    hipFuncAttributes attr;
    TN_HIP(hipFuncGetAttributes(&attr, fn));
    const size_t lim = LDS_BYTES - std::min(LDS_BYTES, (size_t)attr.sharedSizeBytes);
    TN_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
    it = done.emplace(fn, lim).first;
  }
  TN_REQUIRE(bytes <= it->second, TN_E_UNSUPPORTED,
             "kernel needs " + std::to_string(bytes) + " bytes of dynamic LDS, more than the " +
                 std::to_string(it->second) + " available");
}
static void allow_lds(const void* fn, size_t bytes) { raise_lds_limit(fn, bytes); }

// operands whose rows are 16-byte aligned with a valid width of whole
// vectors take the vector loads; any other operand takes element loads
// readable row widths of A and B: the valid width, rounded up to whole
// vectors for an operand the caller marks zero-padded
static long width_a(Prec p, const GemmDesc& g) {
  const int vn = (p == Prec::F64) ? 2 : 4;
  if (g.tA) return g.M;
  return g.zpA ? (g.K + vn - 1) / vn * vn : g.K;
}
static long width_b(Prec p, const GemmDesc& g) {
  const int vn = (p == Prec::F64) ? 2 : 4;
  if (g.tB) return g.K;
  return g.zpB ? (g.N + vn - 1) / vn * vn : g.N;
}
static bool fast_aligned(Prec p, const GemmDesc& g) {
  const int vn = (p == Prec::F64) ? 2 : 4;
  auto al = [&](const void* ptr, long ld, long s, long cols) {
    return ((uintptr_t)ptr % 16 == 0) && (ld % vn == 0) && (s % vn == 0) && (cols % vn == 0) && cols <= ld;
  };
  return al(g.A, g.lda, g.sA, width_a(p, g)) && al(g.B, g.ldb, g.sB, width_b(p, g));
}

bool gemm_fast_ok(Prec p, const GemmDesc& g) {
  if (p != Prec::F64 && p != Prec::F32) return false;
  if (g.cA || g.cB) return false;
  if (g.M < 128 || g.N < 128 || g.K < 16) return false;
  const size_t es = (p == Prec::F64) ? 8 : 4;
  // a 128-row tile must span less than 2^31 bytes (32-bit buffer offsets)
  return (size_t)g.lda * es * 130 < ((size_t)1 << 31) && (size_t)g.ldb * es * 130 < ((size_t)1 << 31);
}

void gemm_fast(tn_ctx* c, Prec p, const GemmDesc& g) {
  FArgs a;
  a.M = g.M; a.N = g.N; a.K = g.K;
  const bool aligned = fast_aligned(p, g);
  a.A = g.A; a.lda = g.lda; a.sA = g.sA;
  a.B = g.B; a.ldb = g.ldb; a.sB = g.sB;
  // the vector kernel reads the padded widths; the element kernel the valid ones
  a.wA = aligned ? width_a(p, g) : (g.tA ? g.M : g.K);
  a.wB = aligned ? width_b(p, g) : (g.tB ? g.K : g.N);
  a.C = g.C; a.ldc = g.ldc; a.sC = g.sC;
  a.alpha = g.alpha; a.beta = g.beta;
  const long ntm = (g.M + 127) / 128, ntn = (g.N + 127) / 128;
  dim3 grid((unsigned)(ntm * ntn), 1, (unsigned)g.batch);
#define TN_LAUNCH_FAST1(T, TA, TB, AL)                                                               \
  do {                                                                                               \
    using CF = FastCfg<T, TA, TB>;                                                                   \
    const size_t shm = 2 * (size_t)(CF::A_ELEMS + CF::B_ELEMS) * sizeof(T);                          \
    allow_lds((const void*)gemm_fast_kernel<T, TA, TB, AL>, shm);                                    \
    hipLaunchKernelGGL((gemm_fast_kernel<T, TA, TB, AL>), grid, dim3(256), shm, c->stream, a);       \
  } while (0)
#define TN_LAUNCH_FAST(T, TA, TB)                 \
  do {                                            \
    if (aligned) TN_LAUNCH_FAST1(T, TA, TB, 1);   \
    else TN_LAUNCH_FAST1(T, TA, TB, 0);           \
  } while (0)
  if (p == Prec::F64) {
    if (!g.tA && !g.tB) TN_LAUNCH_FAST(double, 0, 0);
    else if (g.tA && !g.tB) TN_LAUNCH_FAST(double, 1, 0);
    else if (!g.tA && g.tB) TN_LAUNCH_FAST(double, 0, 1);
    else TN_LAUNCH_FAST(double, 1, 1);
  } else {
    if (!g.tA && !g.tB) TN_LAUNCH_FAST(float, 0, 0);
    else if (g.tA && !g.tB) TN_LAUNCH_FAST(float, 1, 0);
    else if (!g.tA && g.tB) TN_LAUNCH_FAST(float, 0, 1);
    else TN_LAUNCH_FAST(float, 1, 1);
  }
#undef TN_LAUNCH_FAST
#undef TN_LAUNCH_FAST1
  TN_HIP(hipGetLastError());
}

void gemm_bf16x3_nt(tn_ctx* c, long M, long N, long K, const unsigned short* Ah, const unsigned short* Al,
                    long lda, const unsigned short* Bh, const unsigned short* Bl, long ldb, float* C, long ldc) {
  if (M == 0 || N == 0) return;
  TN_REQUIRE(lda % 8 == 0 && ldb % 8 == 0 && lda >= ((K + 7) / 8) * 8 && ldb >= ((K + 7) / 8) * 8,
             TN_E_INVALID, "gemm_bf16x3_nt: rows must be padded to 8 bf16");
  TN_REQUIRE((size_t)std::max(lda, ldb) * 2 * 130 < ((size_t)1 << 31), TN_E_INVALID,
             "gemm_bf16x3_nt: row stride too large for 32-bit tile offsets");
  SArgs a{M, N, K, Ah, Al, lda, Bh, Bl, ldb, C, ldc};
  const long ntm = (M + SBM - 1) / SBM, ntn = (N + SBN - 1) / SBN;
  const size_t shm = 2 * (size_t)SCfg<32>::STAGE * sizeof(unsigned short);
  allow_lds((const void*)gemm_bf16x3_nt_kernel<32>, shm);
  hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<32>, dim3((unsigned)(ntm * ntn)), dim3(256), shm, c->stream, a);
  TN_HIP(hipGetLastError());
}

void split_bf16(tn_ctx* c, const void* src, bool src_f64, long rows, long cols, long lds, bool transpose,
                unsigned short* hi, unsigned short* lo, long ldd) {
  if (rows == 0 || cols == 0) return;
  if (!transpose) {
    const long total = rows * ldd;
    const int g = (int)std::min<long>(8192, (total + 255) / 256);
    if (src_f64)
      hipLaunchKernelGGL(split_rows_kernel<double>, dim3(g), dim3(256), 0, c->stream, (const double*)src, rows,
                         cols, lds, hi, lo, ldd);
    else
      hipLaunchKernelGGL(split_rows_kernel<float>, dim3(g), dim3(256), 0, c->stream, (const float*)src, rows,
                         cols, lds, hi, lo, ldd);
  } else {
    // output is cols x ldd with ldd >= rows
    const long orows = std::max(rows, ldd);
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((orows + 31) / 32));
    TN_REQUIRE((orows + 31) / 32 <= 65535, TN_E_INVALID, "split_bf16: too many rows");
    if (src_f64)
      hipLaunchKernelGGL(split_transpose_kernel<double>, grid, dim3(256), 0, c->stream, (const double*)src, rows,
                         cols, lds, hi, lo, ldd);
    else
      hipLaunchKernelGGL(split_transpose_kernel<float>, grid, dim3(256), 0, c->stream, (const float*)src, rows,
                         cols, lds, hi, lo, ldd);
  }
  TN_HIP(hipGetLastError());
}

}  // namespace tn
