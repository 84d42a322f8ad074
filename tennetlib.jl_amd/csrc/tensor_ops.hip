// Memory-bound kernels of the hot path: the small MPO contraction ("W step")
// between the two big GEMMs of a matvec / environment update, precision
// conversions, the FP64 Krylov vector operations (fixed-order two-stage
// reductions, no atomics, so results are reproducible run to run) and the
// gather / transpose kernels used by the truncated SVD.
#include "tn_internal.h"

#include <cmath>

namespace tn {
namespace {

constexpr int RED_BLOCKS = 1024;  // first-stage reduction blocks (4 per CU)
constexpr int RED_THREADS = 256;
constexpr int MAX_DOT = 32;       // basis vectors per multi-dot pass

// ---------------------------------------------------------------------------
// W step: out[g', A, S', B] = sum_nz val * in[g, A, S, B]
// block = 64 B-columns x 4 row groups; the (g S) x 64 input slab of one A is
// staged in LDS, each wavefront walks a quarter of the CSR output rows (rows
// are wave-uniform, so the CSR entries come through the scalar cache).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// SPLIT: T is float and the result is stored as bf16 hi/lo pairs
template <typename T, bool SPLIT>
__global__ __launch_bounds__(256) void wstep_kernel(int G, int Gp, int D, int Dp, long nA, long nB,
                                                    const T* __restrict__ in, long in_sg, long in_sA,
                                                    long in_sS, T* __restrict__ out, long out_sg,
                                                    long out_sA, long out_sS,
                                                    const int* __restrict__ rowptr,
                                                    const int* __restrict__ colidx,
                                                    const double* __restrict__ vals,
                                                    unsigned short* __restrict__ ohi,
                                                    unsigned short* __restrict__ olo) {
  extern __shared__ unsigned char smem_raw[];
  T* slab = reinterpret_cast<T*>(smem_raw);  // [K][64]
  const int K = G * D;
  const int tb = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const long A = blockIdx.y;
  const long B = (long)blockIdx.x * 64 + tb;
  const bool valid = B < nB;
  for (int i = ty; i < K; i += 4) {
    const int g = i / D, S = i % D;
    T v = T(0);
    if (valid) v = in[g * in_sg + A * in_sA + S * in_sS + B];
    slab[i * 64 + tb] = v;
  }
  __syncthreads();
  if (!valid) return;
  const int rows = Gp * Dp;
  for (int o = ty; o < rows; o += 4) {
    const int b = rowptr[o], e = rowptr[o + 1];
    T acc = T(0);
    for (int z = b; z < e; ++z) acc += (T)vals[z] * slab[colidx[z] * 64 + tb];
    const int gp = o / Dp, Sp = o % Dp;
    const long off = gp * out_sg + A * out_sA + Sp * out_sS + B;
    if constexpr (SPLIT) {
      const unsigned short h = f2bf_rne((float)acc);
      ohi[off] = h;
      olo[off] = f2bf_rne((float)acc - __uint_as_float(((unsigned int)h) << 16));
    } else {
      out[off] = acc;
    }
  }
}

// The same W step with VEC consecutive B per thread (16-byte loads and
// stores: 64 VEC columns per workgroup, so every row segment it touches is
// 1 KiB contiguous instead of 256 B); shapes and strides multiples of VEC.
// Dynamic LDS: K x 64 VEC elements.
template <typename T, int VEC>
__global__ __launch_bounds__(256) void wstep_vec_kernel(int G, int Gp, int D, int Dp, long nB,
                                                        const T* __restrict__ in, long in_sg, long in_sA,
                                                        long in_sS, T* __restrict__ out, long out_sg,
                                                        long out_sA, long out_sS,
                                                        const int* __restrict__ rowptr,
                                                        const int* __restrict__ colidx,
                                                        const double* __restrict__ vals) {
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  extern __shared__ unsigned char smem_raw[];
  vec_t* slab = reinterpret_cast<vec_t*>(smem_raw);  // [K][64] vectors
  const int K = G * D;
  const int tb = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const long A = blockIdx.y;
  const long B = ((long)blockIdx.x * 64 + tb) * VEC;
  const bool valid = B < nB;
  for (int i = ty; i < K; i += 4) {
    const int g = i / D, S = i % D;
    vec_t v = (vec_t)(T(0));
    if (valid) v = *reinterpret_cast<const vec_t*>(in + g * in_sg + A * in_sA + S * in_sS + B);
    slab[i * 64 + tb] = v;
  }
  __syncthreads();
  if (!valid) return;
  const int rows = Gp * Dp;
  for (int o = ty; o < rows; o += 4) {
    const int b = rowptr[o], e = rowptr[o + 1];
    vec_t acc = (vec_t)(T(0));
    for (int z = b; z < e; ++z) acc += (T)vals[z] * slab[colidx[z] * 64 + tb];
    const int gp = o / Dp, Sp = o % Dp;
    *reinterpret_cast<vec_t*>(out + gp * out_sg + A * out_sA + Sp * out_sS + B) = acc;
  }
}

// The W step of a dense operator (a compressed MPO pair: every output row
// mixes most of the (w, S) inputs) as an MFMA product per workgroup:
// out[(g', S'), cols] = Op[(g', S'), (g, S)] * slab[(g, S), cols] for one A
// index and WM_NB columns.  Op (Mp x Kp, zero-padded) and the input slab
// (Kp x WM_NB) sit in LDS; each wave owns 32 output columns and all row
// blocks (FP32: v_mfma_f32_32x32x2f32, FP64: v_mfma_f64_16x16x4f64).
constexpr int WM_NB = 128;
typedef float f16v __attribute__((ext_vector_type(16)));
typedef double d4 __attribute__((ext_vector_type(4)));

// SPLIT (T = float): the result goes out as bf16 hi/lo pairs (ohi, olo).
template <typename T, bool SPLIT>
__global__ __launch_bounds__(256) void wstep_mfma_kernel(int G, int Gp, int D, int Dp, long nB,
                                                         const T* __restrict__ in, long in_sg, long in_sA,
                                                         long in_sS, T* __restrict__ out, long out_sg,
                                                         long out_sA, long out_sS, const T* __restrict__ opd,
                                                         int Mp, int Kp, unsigned short* __restrict__ ohi,
                                                         unsigned short* __restrict__ olo, int vecin) {
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int VN = 16 / (int)sizeof(T);  // elements per 16-byte vector
  typedef T vec_t __attribute__((ext_vector_type(VN)));
  extern __shared__ unsigned char smem_raw[];
  const int M = Gp * Dp, K = G * D;
  const int lda = Kp + 1;
  T* Aop = reinterpret_cast<T*>(smem_raw);  // Mp x lda
  T* slab = Aop + Mp * lda;                 // Kp x WM_NB
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long Ai = blockIdx.y;
  const long B0 = (long)blockIdx.x * WM_NB;
  for (int t = tid; t < Mp * Kp; t += 256) {
    const int m = t / Kp, k = t % Kp;
    Aop[m * lda + k] = (m < M && k < K) ? opd[m * K + k] : T(0);
  }
  if (vecin) {  // 16-byte loads: strides, nB and the input base are whole vectors
    for (int t = tid; t < Kp * (WM_NB / VN); t += 256) {
      const int i = t / (WM_NB / VN), col = (t % (WM_NB / VN)) * VN;
      vec_t v = (vec_t)(T(0));
      if (i < K && B0 + col < nB) {
        const int g = i / D, S = i % D;
        v = *reinterpret_cast<const vec_t*>(in + g * in_sg + Ai * in_sA + S * in_sS + B0 + col);
      }
      *reinterpret_cast<vec_t*>(slab + i * WM_NB + col) = v;
    }
  } else {
    for (int t = tid; t < Kp * WM_NB; t += 256) {
      const int i = t / WM_NB, col = t % WM_NB;
      T v = T(0);
      if (i < K && B0 + col < nB) {
        const int g = i / D, S = i % D;
        v = in[g * in_sg + Ai * in_sA + S * in_sS + B0 + col];
      }
      slab[i * WM_NB + col] = v;
    }
  }
  __syncthreads();
  if constexpr (!F64) {
    const int nmb = Mp / 32;  // <= 4
    f16v acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = f16v{};
    for (int k0 = 0; k0 < Kp; k0 += 2) {
      const int kr = k0 + (lane >> 5);
      const float bv = slab[kr * WM_NB + wave * 32 + (lane & 31)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (i < nmb) acc[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(Aop[(i * 32 + (lane & 31)) * lda + kr], bv,
                                                                   acc[i], 0, 0, 0);
    }
    // register r of lane: row 8(r/4) + 4(lane/32) + r%4, column lane%32
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= nmb) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long col = B0 + wave * 32 + (lane & 31);
        if (row < M && col < nB) {
          const int gp = row / Dp, Sp = row % Dp;
          const long off = gp * out_sg + Ai * out_sA + Sp * out_sS + col;
          if constexpr (SPLIT) {
            const float x = acc[i][r];
            const unsigned short h = f2bf_rne(x);
            ohi[off] = h;
            olo[off] = f2bf_rne(x - __uint_as_float(((unsigned int)h) << 16));
          } else {
            out[off] = acc[i][r];
          }
        }
      }
    }
  } else {
    const int nmb = Mp / 16;  // <= 8
    d4 acc[8][2];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i][0] = acc[i][1] = d4{0, 0, 0, 0};
    for (int k0 = 0; k0 < Kp; k0 += 4) {
      const int kr = k0 + (lane >> 4);
      const double b0 = slab[kr * WM_NB + wave * 32 + (lane & 15)];
      const double b1 = slab[kr * WM_NB + wave * 32 + 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (i >= nmb) continue;
        const double av = Aop[(i * 16 + (lane & 15)) * lda + kr];
        acc[i][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b0, acc[i][0], 0, 0, 0);
        acc[i][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1, acc[i][1], 0, 0, 0);
      }
    }
    // register r of lane: row lane/16 + 4r, column lane%16
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i >= nmb) continue;
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = i * 16 + (lane >> 4) + 4 * r;
          const long col = B0 + wave * 32 + j * 16 + (lane & 15);
          if (row < M && col < nB) {
            const int gp = row / Dp, Sp = row % Dp;
            out[gp * out_sg + Ai * out_sA + Sp * out_sS + col] = acc[i][j][r];
          }
        }
    }
  }
}

// ---------------------------------------------------------------------------
// complex <-> real FP32 layouts of the complex mixed-precision product: a
// complex product C = A B runs as one real GEMM [Cr | Ci] = [Ar | Ai] *
// [[Br, Bi], [-Bi, Br]] (four real products' flops, FP32 MFMA)
// ---------------------------------------------------------------------------
// dst (rows x 2 cols): [re | im] of the interleaved complex src (rows x cols)
__global__ void cplx_split_rows_kernel(const double* __restrict__ src, long rows, long cols,
                                       float* __restrict__ dst) {
  const long n = rows * cols;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / cols, c = i % cols;
    dst[r * 2 * cols + c] = (float)src[2 * i];
    dst[r * 2 * cols + cols + c] = (float)src[2 * i + 1];
  }
}
// dst (2 rows x 2 cols) = [[re, im], [-im, re]] of the interleaved complex src
__global__ void cplx_block_kernel(const double* __restrict__ src, long rows, long cols, float* __restrict__ dst) {
  const long n = rows * cols;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / cols, c = i % cols;
    const float re = (float)src[2 * i], im = (float)src[2 * i + 1];
    dst[r * 2 * cols + c] = re;
    dst[r * 2 * cols + cols + c] = im;
    dst[(rows + r) * 2 * cols + c] = -im;
    dst[(rows + r) * 2 * cols + cols + c] = re;
  }
}
// interleaved complex dst (rows x cols) from src (rows x 2 cols) = [re | im]
__global__ void cplx_join_rows_kernel(const float* __restrict__ src, long rows, long cols,
                                      double* __restrict__ dst) {
  const long n = rows * cols;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / cols, c = i % cols;
    dst[2 * i] = (double)src[r * 2 * cols + c];
    dst[2 * i + 1] = (double)src[r * 2 * cols + cols + c];
  }
}

// ---------------------------------------------------------------------------
// conversions / fills
// ---------------------------------------------------------------------------
__global__ void cvt_d2f_kernel(const double* __restrict__ x, float* __restrict__ y, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
// rows x cols block with row strides lds (source) and ldd (destination);
// destination columns cols .. ldd - 1 are set to zero
__global__ void cvt_d2f_2d_kernel(const double* __restrict__ x, long rows, long cols, long lds,
                                  float* __restrict__ y, long ldd) {
  const long n = rows * ldd;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldd, c = i % ldd;
    y[i] = c < cols ? (float)x[r * lds + c] : 0.0f;
  }
}
// rows x cols FP64 block copied to row stride ldd, columns cols .. ldd - 1 zero
__global__ void pad_rows_f64_kernel(const double* __restrict__ x, long rows, long cols, long lds,
                                    double* __restrict__ y, long ldd) {
  const long n = rows * ldd;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldd, c = i % ldd;
    y[i] = c < cols ? x[r * lds + c] : 0.0;
  }
}
__global__ void zero_cols_f64_kernel(double* __restrict__ y, long rows, long c0, long c1, long ld) {
  const long w = c1 - c0, n = rows * w;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[(i / w) * ld + c0 + i % w] = 0.0;
}
// columns c0 .. c1 - 1 of a rows x ld FP32 matrix set to zero
__global__ void zero_cols_f32_kernel(float* __restrict__ y, long rows, long c0, long c1, long ld) {
  const long w = c1 - c0, n = rows * w;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[(i / w) * ld + c0 + i % w] = 0.0f;
}
__global__ void zero_cols_u16_kernel(unsigned short* __restrict__ y, long rows, long c0, long c1, long ld) {
  const long w = c1 - c0, n = rows * w;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[(i / w) * ld + c0 + i % w] = 0;
}
__global__ void cvt_f2d_kernel(const float* __restrict__ x, double* __restrict__ y, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}
__global__ void fill_kernel(double* __restrict__ x, double v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void scaled_copy_kernel(double* __restrict__ y, const double* __restrict__ x, double a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
// y = x / sqrt(max(0, *r2)) with r2 read from device memory (the squared
// norm a Krylov step left there), so the next basis vector needs no host trip
__global__ void scale_rnorm_kernel(double* __restrict__ y, const double* __restrict__ x,
                                   const double* __restrict__ r2, long n) {
  const double a = 1.0 / sqrt(fmax(0.0, *r2));
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}

// ---------------------------------------------------------------------------
// Krylov: multi-dot (first stage + fixed-order second stage)
// part[v * nblk + blk] = partial sum of <basis_v, x> over the block's slice.
// Complex: data are interleaved (re, im); <u, x> = sum conj(u) x -> two parts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) { return wave_sum64(v); }

template <bool CPLX>
__global__ __launch_bounds__(RED_THREADS) void multidot_stage1(const double* __restrict__ basis, long ld,
                                                               int count, const double* __restrict__ x,
                                                               long n, double* __restrict__ part,
                                                               double* __restrict__ hout) {
  __shared__ double red[RED_THREADS / 64][2 * MAX_DOT];
  const int nblk = gridDim.x;
  double acc[2 * MAX_DOT];
#pragma unroll
  for (int v = 0; v < 2 * MAX_DOT; ++v) acc[v] = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!CPLX) {
      const double xi = x[i];
#pragma unroll
      for (int v = 0; v < MAX_DOT; ++v)
        if (v < count) acc[v] = fma(basis[v * ld + i], xi, acc[v]);
    } else {
      const double xr = x[2 * i], xi = x[2 * i + 1];
#pragma unroll
      for (int v = 0; v < MAX_DOT; ++v)
        if (v < count) {
          const double ur = basis[v * ld + 2 * i], ui = basis[v * ld + 2 * i + 1];
          acc[2 * v] = fma(ur, xr, fma(ui, xi, acc[2 * v]));
          acc[2 * v + 1] = fma(ur, xi, fma(-ui, xr, acc[2 * v + 1]));
        }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nacc = CPLX ? 2 * count : count;
#pragma unroll
  for (int v = 0; v < 2 * MAX_DOT; ++v) {
    if (v < nacc) {
      double s = wave_sum(acc[v]);
      if (lane == 0) red[wave][v] = s;
    }
  }
  __syncthreads();
  for (int v = threadIdx.x; v < nacc; v += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < RED_THREADS / 64; ++w) s += red[w][v];
    part[(long)v * nblk + blockIdx.x] = s;
    if (hout) hout[v] = s;  // one-workgroup launch: s is the final value
  }
}

__global__ __launch_bounds__(256) void multidot_stage2(const double* __restrict__ part, int nblk,
                                                       int nacc, double* __restrict__ out,
                                                       double* __restrict__ hout) {
  __shared__ double red[4];
  const int v = blockIdx.x;
  if (v >= nacc) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[(long)v * nblk + i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = red[0] + red[1] + red[2] + red[3];
    out[v] = r;
    if (hout) hout[v] = r;
  }
}

// x -= sum_v coef[v] * basis_v (real or complex coefficients)
template <bool CPLX>
__global__ __launch_bounds__(256) void multi_axpy_kernel(const double* __restrict__ basis, long ld, int count,
                                                         const double* __restrict__ coef, double* __restrict__ x,
                                                         long n, int sub) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!CPLX) {
      double acc = 0.0;
      for (int v = 0; v < count; ++v) acc = fma(coef[v], basis[v * ld + i], acc);
      x[i] = sub ? x[i] - acc : acc;
    } else {
      double ar = 0.0, ai = 0.0;
      for (int v = 0; v < count; ++v) {
        const double cr = coef[2 * v], ci = coef[2 * v + 1];
        const double br = basis[v * ld + 2 * i], bi = basis[v * ld + 2 * i + 1];
        ar = fma(cr, br, fma(-ci, bi, ar));
        ai = fma(cr, bi, fma(ci, br, ai));
      }
      if (sub) { x[2 * i] -= ar; x[2 * i + 1] -= ai; }
      else { x[2 * i] = ar; x[2 * i + 1] = ai; }
    }
  }
}

// A whole Krylov orthogonalisation step for a short vector in one
// workgroup: two passes of w <- w - V (V^H w) over the K basis vectors, then
// |w|^2.  The pass coefficients go to coef[0..) and coef[128..) (complex
// pairs when cplx) and |w|^2 to coef[256], each also to hout (mapped host
// memory) at the same offsets.  Dots: one wave per basis vector (DPP sums).
constexpr int KS_THREADS = 1024;
template <bool CPLX>
__global__ __launch_bounds__(KS_THREADS) void krylov_ortho_small_kernel(const double* __restrict__ basis, long ld,
                                                                        int K, double* __restrict__ w, long n,
                                                                        double* __restrict__ coef,
                                                                        double* __restrict__ hout) {
  __shared__ double cf[128];
  __shared__ double red[KS_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = KS_THREADS / 64;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = wave; i < K; i += NW) {
      const double* v = basis + (long)i * ld;
      double re = 0.0, im = 0.0;
      if (!CPLX) {
        for (long j = lane; j < n; j += 64) re = fma(v[j], w[j], re);
      } else {
        for (long j = lane; j < n; j += 64) {
          const double vr = v[2 * j], vi = v[2 * j + 1], wr = w[2 * j], wi = w[2 * j + 1];
          re = fma(vr, wr, fma(vi, wi, re));
          im = fma(vr, wi, fma(-vi, wr, im));
        }
      }
      re = wave_sum64(re);
      if (CPLX) im = wave_sum64(im);
      if (lane == 0) {
        if (CPLX) { cf[2 * i] = re; cf[2 * i + 1] = im; }
        else cf[i] = re;
      }
    }
    __syncthreads();
    for (long j = tid; j < n; j += KS_THREADS) {
      if (!CPLX) {
        double acc = 0.0;
        for (int v = 0; v < K; ++v) acc = fma(cf[v], basis[v * ld + j], acc);
        w[j] -= acc;
      } else {
        double ar = 0.0, ai = 0.0;
        for (int v = 0; v < K; ++v) {
          const double cr = cf[2 * v], ci = cf[2 * v + 1];
          const double br = basis[v * ld + 2 * j], bi = basis[v * ld + 2 * j + 1];
          ar = fma(cr, br, fma(-ci, bi, ar));
          ai = fma(cr, bi, fma(ci, br, ai));
        }
        w[2 * j] -= ar;
        w[2 * j + 1] -= ai;
      }
    }
    const int nc = CPLX ? 2 * K : K;
    for (int i = tid; i < nc; i += KS_THREADS) {
      coef[128 * pass + i] = cf[i];
      hout[128 * pass + i] = cf[i];
    }
    __syncthreads();  // w complete and cf read before the next pass rewrites cf
  }
  const long nd = CPLX ? 2 * n : n;
  double s = 0.0;
  for (long j = tid; j < nd; j += KS_THREADS) s = fma(w[j], w[j], s);
  s = wave_sum64(s);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int i = 0; i < NW; ++i) t += red[i];
    coef[256] = t;
    hout[256] = t;
  }
}

// row sums of squares: one block per row
__global__ __launch_bounds__(256) void row_norms2_kernel(const double* __restrict__ X, long L, long ld,
                                                         double* __restrict__ out) {
  __shared__ double red[4];
  const long r = blockIdx.x;
  double s = 0.0;
  for (long j = threadIdx.x; j < L; j += blockDim.x) {
    const double v = X[r * ld + j];
    s = fma(v, v, s);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[r] = red[0] + red[1] + red[2] + red[3];
}

// out (C x R) = in^T, in is R x C; 32x32 LDS tiles
__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ in, long R, long C,
                                                        double* __restrict__ out) {
  __shared__ double t[32][33];
  const long r0 = (long)blockIdx.y * 32, c0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const long r = r0 + k, cc = c0 + tx;
    t[k][tx] = (r < R && cc < C) ? in[r * C + cc] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const long cc = c0 + k, r = r0 + tx;
    if (cc < C && r < R) out[cc * R + r] = t[tx][k];
  }
}

__global__ __launch_bounds__(256) void gather_rows_kernel(const double* __restrict__ src, long ld,
                                                          const int* __restrict__ perm,
                                                          const double* __restrict__ scale, long L,
                                                          double* __restrict__ out) {
  const long k = blockIdx.y;
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < L) out[k * L + j] = src[(long)perm[k] * ld + j] * scale[k];
}

// out[j][k] = src[perm[k]][j] * scale[k]; 32x32 tiles through LDS
__global__ __launch_bounds__(256) void gather_rows_T_kernel(const double* __restrict__ src, long ld,
                                                            const int* __restrict__ perm,
                                                            const double* __restrict__ scale, long kcount,
                                                            long L, double* __restrict__ out) {
  __shared__ double t[32][33];
  const long k0 = (long)blockIdx.y * 32, j0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int a = ty; a < 32; a += 8) {
    const long k = k0 + a, j = j0 + tx;
    t[a][tx] = (k < kcount && j < L) ? src[(long)perm[k] * ld + j] * scale[k] : 0.0;
  }
  __syncthreads();
  for (int a = ty; a < 32; a += 8) {
    const long j = j0 + a, k = k0 + tx;
    if (j < L && k < kcount) out[j * kcount + k] = t[tx][a];
  }
}

// complex: out[j][k] = conj?(src[perm[k]][j]) * scale[k]; 16x16 complex tiles
__global__ __launch_bounds__(256) void gather_rows_T_z_kernel(const double2* __restrict__ src, long ld,
                                                              const int* __restrict__ perm,
                                                              const double* __restrict__ scale, long kcount,
                                                              long L, double2* __restrict__ out, int conj) {
  __shared__ double2 t[16][17];
  const long k0 = (long)blockIdx.y * 16, j0 = (long)blockIdx.x * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16
  {
    const long k = k0 + ty, j = j0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (k < kcount && j < L) {
      v = src[(long)perm[k] * ld + j];
      if (conj) v.y = -v.y;
      v.x *= scale[k];
      v.y *= scale[k];
    }
    t[ty][tx] = v;
  }
  __syncthreads();
  const long j = j0 + ty, k = k0 + tx;
  if (j < L && k < kcount) out[j * kcount + k] = t[tx][ty];
}

inline int grid1d(long n, int threads = 256) {
  long g = (n + threads - 1) / threads;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// ---------------------------------------------------------------------------
void build_sparse_op(SparseOp& op, int rows, int cols, const std::vector<double>& dense) {
  std::vector<int> rp(rows + 1, 0), ci;
  std::vector<double> v;
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      const double x = dense[(size_t)r * cols + c];
      if (x != 0.0) { ci.push_back(c); v.push_back(x); }
    }
    rp[r + 1] = (int)ci.size();
  }
  op.rows = rows;
  op.cols = cols;
  op.nnz = (int)ci.size();
  op.rowptr.ensure(rp.size() * sizeof(int));
  op.colidx.ensure(std::max<size_t>(1, ci.size()) * sizeof(int));
  op.vals.ensure(std::max<size_t>(1, v.size()) * sizeof(double));
  TN_HIP(hipMemcpy(op.rowptr.p, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!ci.empty()) {
    TN_HIP(hipMemcpy(op.colidx.p, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(op.vals.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  op.dense = rows <= 128 && cols <= 128 && op.nnz >= 8 * rows;
  if (op.dense) {
    std::vector<float> f(dense.begin(), dense.begin() + (size_t)rows * cols);
    op.d64.ensure((size_t)rows * cols * sizeof(double));
    op.d32.ensure((size_t)rows * cols * sizeof(float));
    TN_HIP(hipMemcpy(op.d64.p, dense.data(), (size_t)rows * cols * sizeof(double), hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(op.d32.p, f.data(), (size_t)rows * cols * sizeof(float), hipMemcpyHostToDevice));
  }
}

static void wstep_launch(tn_ctx* c, bool fp32, bool split, const WStepDesc& w, unsigned short* hi,
                         unsigned short* lo) {
  TN_REQUIRE(w.op && w.op->rows == w.Gp * w.Dp && w.op->cols == w.G * w.D, TN_E_SHAPE,
             "wstep: operator shape mismatch");
  if (w.nA == 0 || w.nB == 0) return;
  TN_REQUIRE(w.nA <= 65535L, TN_E_INVALID, "wstep: nA exceeds grid.y");
  const int K = w.G * w.D;
  const int* rp = w.op->rowptr.as<int>();
  const int* ci = w.op->colidx.as<int>();
  const double* vv = w.op->vals.as<double>();
  if (w.op->dense) {
    // dense operator: MFMA product per workgroup when Op and the slab fit in LDS
    const int mb = fp32 ? 32 : 16, ks = fp32 ? 2 : 4;
    const int M = w.Gp * w.Dp;
    const int Mp = (M + mb - 1) / mb * mb, Kp = (K + ks - 1) / ks * ks;
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    const size_t mshm = ((size_t)Mp * (Kp + 1) + (size_t)Kp * WM_NB) * es;
    if (mshm <= 160 * 1024 && Mp <= 128) {
      const long nbt = (w.nB + WM_NB - 1) / WM_NB;
      const int vn = fp32 ? 4 : 2;
      auto whole = [&](long x) { return x % vn == 0; };
      const int vecin = whole(w.nB) && whole(w.in_sg) && whole(w.in_sA) && whole(w.in_sS) &&
                        (uintptr_t)w.in % 16 == 0;
      dim3 grid((unsigned)nbt, (unsigned)w.nA);
      if (split) {
        raise_lds_limit((const void*)wstep_mfma_kernel<float, true>, mshm);
        hipLaunchKernelGGL((wstep_mfma_kernel<float, true>), grid, dim3(256), mshm, c->stream, w.G, w.Gp, w.D,
                           w.Dp, w.nB, (const float*)w.in, w.in_sg, w.in_sA, w.in_sS, (float*)nullptr, w.out_sg,
                           w.out_sA, w.out_sS, w.op->d32.as<float>(), Mp, Kp, hi, lo, vecin);
      } else if (fp32) {
        raise_lds_limit((const void*)wstep_mfma_kernel<float, false>, mshm);
        hipLaunchKernelGGL((wstep_mfma_kernel<float, false>), grid, dim3(256), mshm, c->stream, w.G, w.Gp, w.D,
                           w.Dp, w.nB, (const float*)w.in, w.in_sg, w.in_sA, w.in_sS, (float*)w.out, w.out_sg,
                           w.out_sA, w.out_sS, w.op->d32.as<float>(), Mp, Kp, (unsigned short*)nullptr,
                           (unsigned short*)nullptr, vecin);
      } else {
        raise_lds_limit((const void*)wstep_mfma_kernel<double, false>, mshm);
        hipLaunchKernelGGL((wstep_mfma_kernel<double, false>), grid, dim3(256), mshm, c->stream, w.G, w.Gp, w.D,
                           w.Dp, w.nB, (const double*)w.in, w.in_sg, w.in_sA, w.in_sS, (double*)w.out, w.out_sg,
                           w.out_sA, w.out_sS, w.op->d64.as<double>(), Mp, Kp, (unsigned short*)nullptr,
                           (unsigned short*)nullptr, vecin);
      }
      TN_HIP(hipGetLastError());
      return;
    }
  }
  if (!split) {
    // 16-byte vectors along B when every offset allows them and the slab fits
    const int vec = fp32 ? 4 : 2;
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    const size_t vshm = (size_t)K * 64 * vec * es;
    auto ok = [&](long x) { return x % vec == 0; };
    if (w.nB >= 64 * vec && ok(w.nB) && ok(w.in_sg) && ok(w.in_sA) && ok(w.in_sS) && ok(w.out_sg) &&
        ok(w.out_sA) && ok(w.out_sS) && (uintptr_t)w.in % 16 == 0 && (uintptr_t)w.out % 16 == 0 &&
        vshm <= 160 * 1024) {
      dim3 grid((unsigned)((w.nB / vec + 63) / 64), (unsigned)w.nA);
      if (fp32) {
        raise_lds_limit((const void*)wstep_vec_kernel<float, 4>, vshm);
        hipLaunchKernelGGL((wstep_vec_kernel<float, 4>), grid, dim3(256), vshm, c->stream, w.G, w.Gp, w.D, w.Dp,
                           w.nB, (const float*)w.in, w.in_sg, w.in_sA, w.in_sS, (float*)w.out, w.out_sg, w.out_sA,
                           w.out_sS, rp, ci, vv);
      } else {
        raise_lds_limit((const void*)wstep_vec_kernel<double, 2>, vshm);
        hipLaunchKernelGGL((wstep_vec_kernel<double, 2>), grid, dim3(256), vshm, c->stream, w.G, w.Gp, w.D, w.Dp,
                           w.nB, (const double*)w.in, w.in_sg, w.in_sA, w.in_sS, (double*)w.out, w.out_sg,
                           w.out_sA, w.out_sS, rp, ci, vv);
      }
      TN_HIP(hipGetLastError());
      return;
    }
  }
  const size_t shm = (size_t)K * 64 * (fp32 ? sizeof(float) : sizeof(double));
  TN_REQUIRE(shm <= 64 * 1024, TN_E_UNSUPPORTED, "wstep: MPO slab exceeds 64 KiB of LDS");
  dim3 grid((unsigned)((w.nB + 63) / 64), (unsigned)w.nA);
  if (split)
    hipLaunchKernelGGL((wstep_kernel<float, true>), grid, dim3(256), shm, c->stream, w.G, w.Gp, w.D, w.Dp, w.nA,
                       w.nB, (const float*)w.in, w.in_sg, w.in_sA, w.in_sS, (float*)nullptr, w.out_sg, w.out_sA,
                       w.out_sS, rp, ci, vv, hi, lo);
  else if (fp32)
    hipLaunchKernelGGL((wstep_kernel<float, false>), grid, dim3(256), shm, c->stream, w.G, w.Gp, w.D, w.Dp, w.nA,
                       w.nB, (const float*)w.in, w.in_sg, w.in_sA, w.in_sS, (float*)w.out, w.out_sg, w.out_sA,
                       w.out_sS, rp, ci, vv, (unsigned short*)nullptr, (unsigned short*)nullptr);
  else
    hipLaunchKernelGGL((wstep_kernel<double, false>), grid, dim3(256), shm, c->stream, w.G, w.Gp, w.D, w.Dp, w.nA,
                       w.nB, (const double*)w.in, w.in_sg, w.in_sA, w.in_sS, (double*)w.out, w.out_sg, w.out_sA,
                       w.out_sS, rp, ci, vv, (unsigned short*)nullptr, (unsigned short*)nullptr);
  TN_HIP(hipGetLastError());
}

void wstep(tn_ctx* c, bool fp32, const WStepDesc& w) { wstep_launch(c, fp32, false, w, nullptr, nullptr); }

void wstep_split(tn_ctx* c, const WStepDesc& w, unsigned short* hi, unsigned short* lo) {
  wstep_launch(c, true, true, w, hi, lo);
}

void cvt_f64_to_f32(tn_ctx* c, const double* x, float* y, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cvt_d2f_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, x, y, n);
  TN_HIP(hipGetLastError());
}
void cvt_f64_to_f32_2d(tn_ctx* c, const double* x, long rows, long cols, long lds, float* y, long ldd) {
  if (rows * cols <= 0) return;
  if (lds == cols && ldd == cols) {
    cvt_f64_to_f32(c, x, y, rows * cols);
    return;
  }
  hipLaunchKernelGGL(cvt_d2f_2d_kernel, dim3(grid1d(rows * ldd)), dim3(256), 0, c->stream, x, rows, cols, lds, y,
                     ldd);
  TN_HIP(hipGetLastError());
}
void pad_rows_f64(tn_ctx* c, const double* x, long rows, long cols, long lds, double* y, long ldd) {
  if (rows * ldd <= 0) return;
  hipLaunchKernelGGL(pad_rows_f64_kernel, dim3(grid1d(rows * ldd)), dim3(256), 0, c->stream, x, rows, cols, lds, y,
                     ldd);
  TN_HIP(hipGetLastError());
}
void zero_cols_f64(tn_ctx* c, double* y, long rows, long c0, long c1, long ld) {
  if (rows <= 0 || c1 <= c0) return;
  hipLaunchKernelGGL(zero_cols_f64_kernel, dim3(grid1d(rows * (c1 - c0))), dim3(256), 0, c->stream, y, rows, c0, c1,
                     ld);
  TN_HIP(hipGetLastError());
}
void zero_cols_f32(tn_ctx* c, float* y, long rows, long c0, long c1, long ld) {
  if (rows <= 0 || c1 <= c0) return;
  hipLaunchKernelGGL(zero_cols_f32_kernel, dim3(grid1d(rows * (c1 - c0))), dim3(256), 0, c->stream, y, rows, c0, c1,
                     ld);
  TN_HIP(hipGetLastError());
}
void zero_cols_u16(tn_ctx* c, unsigned short* y, long rows, long c0, long c1, long ld) {
  if (rows <= 0 || c1 <= c0) return;
  hipLaunchKernelGGL(zero_cols_u16_kernel, dim3(grid1d(rows * (c1 - c0))), dim3(256), 0, c->stream, y, rows, c0, c1,
                     ld);
  TN_HIP(hipGetLastError());
}
void cvt_f32_to_f64(tn_ctx* c, const float* x, double* y, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cvt_f2d_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, x, y, n);
  TN_HIP(hipGetLastError());
}
void cplx_split_rows_f32(tn_ctx* c, const double* src, long rows, long cols, float* dst) {
  if (rows * cols <= 0) return;
  hipLaunchKernelGGL(cplx_split_rows_kernel, dim3(grid1d(rows * cols)), dim3(256), 0, c->stream, src, rows, cols,
                     dst);
  TN_HIP(hipGetLastError());
}
void cplx_block_f32(tn_ctx* c, const double* src, long rows, long cols, float* dst) {
  if (rows * cols <= 0) return;
  hipLaunchKernelGGL(cplx_block_kernel, dim3(grid1d(rows * cols)), dim3(256), 0, c->stream, src, rows, cols, dst);
  TN_HIP(hipGetLastError());
}
void cplx_join_rows_f64(tn_ctx* c, const float* src, long rows, long cols, double* dst) {
  if (rows * cols <= 0) return;
  hipLaunchKernelGGL(cplx_join_rows_kernel, dim3(grid1d(rows * cols)), dim3(256), 0, c->stream, src, rows, cols,
                     dst);
  TN_HIP(hipGetLastError());
}
void fill_f64(tn_ctx* c, double* x, double v, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, x, v, n);
  TN_HIP(hipGetLastError());
}
void copy_f64(tn_ctx* c, double* dst, const double* src, long n) {
  if (n <= 0) return;
  TN_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}
void scale_f64(tn_ctx* c, double* x, double s, long n) { scaled_copy(c, x, x, s, n); }
void scaled_copy(tn_ctx* c, double* y, const double* x, double a, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scaled_copy_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, y, x, a, n);
  TN_HIP(hipGetLastError());
}
void scale_rnorm(tn_ctx* c, double* y, const double* x, const double* r2_dev, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_rnorm_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, y, x, r2_dev, n);
  TN_HIP(hipGetLastError());
}

void multi_dot(tn_ctx* c, const double* basis, long ld, int count, const double* x, long n, bool cplx,
               double* out_dev, double* out_host) {
  TN_REQUIRE(count >= 1, TN_E_INVALID, "multi_dot: count");
  // short vectors (small bonds): one workgroup sums everything and writes
  // the results itself (part[v * 1 + 0] is out[v]), so no second launch
  const bool one = n <= 16L * RED_THREADS;
  const int nblk = one ? 1 : (int)std::min<long>(RED_BLOCKS, std::max<long>(1, (n + RED_THREADS - 1) / RED_THREADS));
  double* part = one ? nullptr : (double*)c->workspace(WS_KRY_PART, (size_t)2 * MAX_DOT * nblk * sizeof(double));
  for (int v0 = 0; v0 < count; v0 += MAX_DOT) {
    const int cnt = std::min(MAX_DOT, count - v0);
    const int nacc = cplx ? 2 * cnt : cnt;
    double* o = out_dev + (cplx ? 2 * v0 : v0);
    double* ho = out_host ? out_host + (cplx ? 2 * v0 : v0) : nullptr;
    double* dst = one ? o : part;
    double* hdst = one ? ho : nullptr;
    if (cplx)
      hipLaunchKernelGGL(multidot_stage1<true>, dim3(nblk), dim3(RED_THREADS), 0, c->stream, basis + v0 * ld,
                         ld, cnt, x, n, dst, hdst);
    else
      hipLaunchKernelGGL(multidot_stage1<false>, dim3(nblk), dim3(RED_THREADS), 0, c->stream, basis + v0 * ld,
                         ld, cnt, x, n, dst, hdst);
    TN_HIP(hipGetLastError());
    if (one) continue;
    hipLaunchKernelGGL(multidot_stage2, dim3(nacc), dim3(256), 0, c->stream, part, nblk, nacc, o, ho);
    TN_HIP(hipGetLastError());
  }
}

namespace {
// y += scale * coef[0] * x (coef on the device; complex pair when cplx)
__global__ __launch_bounds__(256) void add_scaled_dev_kernel(double* __restrict__ y, const double* __restrict__ x,
                                                             const double* __restrict__ coef, double scale, long n,
                                                             int cplx) {
  const double cr = scale * coef[0], ci = cplx ? scale * coef[1] : 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!cplx) {
      y[i] = fma(cr, x[i], y[i]);
    } else {
      const double xr = x[2 * i], xi = x[2 * i + 1];
      y[2 * i] += cr * xr - ci * xi;
      y[2 * i + 1] += cr * xi + ci * xr;
    }
  }
}

}  // namespace

void add_scaled_dev(tn_ctx* c, double* y, const double* x, const double* coef_dev, double scale, long n, bool cplx) {
  if (n <= 0) return;
  hipLaunchKernelGGL(add_scaled_dev_kernel, dim3(grid1d(n)), dim3(256), 0, c->stream, y, x, coef_dev, scale, n,
                     cplx ? 1 : 0);
  TN_HIP(hipGetLastError());
}

bool krylov_ortho_small(tn_ctx* c, const double* basis, long ld, int K, double* w, long n, bool cplx,
                        double* coef_dev, double* hout) {
  if (n > 4L * KS_THREADS || K < 1 || K > 61) return false;
  if (cplx)
    hipLaunchKernelGGL(krylov_ortho_small_kernel<true>, dim3(1), dim3(KS_THREADS), 0, c->stream, basis, ld, K, w,
                       n, coef_dev, hout);
  else
    hipLaunchKernelGGL(krylov_ortho_small_kernel<false>, dim3(1), dim3(KS_THREADS), 0, c->stream, basis, ld, K, w,
                       n, coef_dev, hout);
  TN_HIP(hipGetLastError());
  return true;
}

void multi_axpy_sub(tn_ctx* c, const double* basis, long ld, int count, const double* coef_dev, double* x,
                    long n, bool cplx) {
  if (n <= 0 || count <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(multi_axpy_kernel<true>, dim3(grid1d(n)), dim3(256), 0, c->stream, basis, ld, count,
                       coef_dev, x, n, 1);
  else
    hipLaunchKernelGGL(multi_axpy_kernel<false>, dim3(grid1d(n)), dim3(256), 0, c->stream, basis, ld, count,
                       coef_dev, x, n, 1);
  TN_HIP(hipGetLastError());
}

void lincomb(tn_ctx* c, const double* basis, long ld, int count, const double* coef_dev, double* out, long n,
             bool cplx) {
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL(multi_axpy_kernel<true>, dim3(grid1d(n)), dim3(256), 0, c->stream, basis, ld, count,
                       coef_dev, out, n, 0);
  else
    hipLaunchKernelGGL(multi_axpy_kernel<false>, dim3(grid1d(n)), dim3(256), 0, c->stream, basis, ld, count,
                       coef_dev, out, n, 0);
  TN_HIP(hipGetLastError());
}

void row_norms2(tn_ctx* c, const double* X, long rows, long L, long ld, double* out) {
  if (rows <= 0) return;
  TN_REQUIRE(rows <= 2147483647L, TN_E_INVALID, "row_norms2: rows");
  hipLaunchKernelGGL(row_norms2_kernel, dim3((unsigned)rows), dim3(256), 0, c->stream, X, L, ld, out);
  TN_HIP(hipGetLastError());
}

void transpose_f64(tn_ctx* c, const double* in, long R, long C, double* out) {
  if (R <= 0 || C <= 0) return;
  dim3 grid((unsigned)((C + 31) / 32), (unsigned)((R + 31) / 32));
  TN_REQUIRE((R + 31) / 32 <= 65535, TN_E_INVALID, "transpose: too many rows");
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, c->stream, in, R, C, out);
  TN_HIP(hipGetLastError());
}

void gather_rows(tn_ctx* c, const double* src, long ld_src, const int* perm, const double* scale, long kcount,
                 long L, double* out) {
  if (kcount <= 0 || L <= 0) return;
  TN_REQUIRE(kcount <= 65535, TN_E_INVALID, "gather_rows: kcount");
  dim3 grid((unsigned)((L + 255) / 256), (unsigned)kcount);
  hipLaunchKernelGGL(gather_rows_kernel, grid, dim3(256), 0, c->stream, src, ld_src, perm, scale, L, out);
  TN_HIP(hipGetLastError());
}

void gather_rows_T_z(tn_ctx* c, const double* src, long ld_src, const int* perm, const double* scale,
                     long kcount, long L, double* out, bool conj) {
  if (kcount <= 0 || L <= 0) return;
  dim3 grid((unsigned)((L + 15) / 16), (unsigned)((kcount + 15) / 16));
  TN_REQUIRE((kcount + 15) / 16 <= 65535, TN_E_INVALID, "gather_rows_T_z: kcount");
  hipLaunchKernelGGL(gather_rows_T_z_kernel, grid, dim3(256), 0, c->stream, (const double2*)src, ld_src, perm,
                     scale, kcount, L, (double2*)out, (int)conj);
  TN_HIP(hipGetLastError());
}

void gather_rows_T(tn_ctx* c, const double* src, long ld_src, const int* perm, const double* scale,
                   long kcount, long L, double* out) {
  if (kcount <= 0 || L <= 0) return;
  dim3 grid((unsigned)((L + 31) / 32), (unsigned)((kcount + 31) / 32));
  TN_REQUIRE((kcount + 31) / 32 <= 65535, TN_E_INVALID, "gather_rows_T: kcount");
  hipLaunchKernelGGL(gather_rows_T_kernel, grid, dim3(256), 0, c->stream, src, ld_src, perm, scale, kcount, L,
                     out);
  TN_HIP(hipGetLastError());
}

}  // namespace tn
