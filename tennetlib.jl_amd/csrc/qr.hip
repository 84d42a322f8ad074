// Householder QR of tall or wide FP64 matrices and small tensor permutes,
// for the isometry moves of tree tensor networks (ITensors qr(A, uinds),
// called from moveisometry_to_next!, src/ttn/ttn.jl:292-306).
//
// The matrix M (m x n) is held column by column: X (n x m), row c = column c
// of M, so that every column is contiguous.  Step j (one launch, n - j
// workgroups): every workgroup rebuilds reflector j from column j (LAPACK
// dlarfg convention: beta = -sign(alpha)|x|, tau = (beta - alpha)/beta,
// v_0 = 1, v_i = x_i / (alpha - beta)), workgroup 0 stores it, and workgroup
// c > 0 applies it to column j + c.  Column j is only read during step j, so
// no workgroup waits for another.  Q^T is then formed from [I | 0] by the
// compact-WY blocks of 64 reflectors, in reverse order, with FP64 MFMA GEMMs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tn_internal.h"

namespace tn {
namespace {

__device__ double qr_block_sum(double v, double* red) {
  v = wave_sum64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void qr_step_kernel(double* __restrict__ X, long m, int j, double* __restrict__ Vt,
                                                      double* __restrict__ tau, double* __restrict__ beta) {
  __shared__ double red[4];
  const double* xj = X + (long)j * m;
  const long c = (long)j + blockIdx.x;
  double* y = X + c * m;
  const double alpha = xj[j];
  const double yj = y[j];  // read before this workgroup writes column c
  double s = 0.0;
  for (long i = j + 1 + threadIdx.x; i < m; i += 256) {
    const double v = xj[i];
    s = fma(v, v, s);
  }
  const double sigma = qr_block_sum(s, red);
  double t = 0.0, b = alpha, sc = 0.0;
  if (sigma > 0.0) {
    const double nrm = sqrt(fma(alpha, alpha, sigma));
    b = alpha >= 0.0 ? -nrm : nrm;
    t = (b - alpha) / b;
    sc = 1.0 / (alpha - b);
  }
  if (blockIdx.x == 0) {
    double* v = Vt + (long)j * m;
    for (long i = threadIdx.x; i < m; i += 256) v[i] = i < j ? 0.0 : (i == j ? 1.0 : xj[i] * sc);
    if (threadIdx.x == 0) {
      tau[j] = t;
      beta[j] = b;
    }
    return;
  }
  if (t == 0.0) return;
  double d = 0.0;
  for (long i = j + 1 + threadIdx.x; i < m; i += 256) d = fma(xj[i] * sc, y[i], d);
  d = qr_block_sum(d, red);
  const double f = t * (yj + d);
  for (long i = j + threadIdx.x; i < m; i += 256) y[i] -= f * (i == j ? 1.0 : xj[i] * sc);
}

// Register-resident step for m - j - 1 <= QR_THREADS * ITEMS: each thread
// loads its ITEMS entries of columns j and c once (all loads in flight),
// both sums of the step come from one pass (|x|^2 for the reflector and
// x.y for the update), and column c is written once.
constexpr int QR_THREADS = 512;

__device__ void qr_block_sum2(double& a, double& b, double* red) {
  a = wave_sum64(a);
  b = wave_sum64(b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wave] = a;
    red[16 + wave] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int w = 0; w < QR_THREADS / 64; ++w) {
    sa += red[w];
    sb += red[16 + w];
  }
  a = sa;
  b = sb;
}

template <int ITEMS>
__global__ __launch_bounds__(QR_THREADS) void qr_step_reg_kernel(double* __restrict__ X, long m, int j,
                                                                 double* __restrict__ Vt, double* __restrict__ tau,
                                                                 double* __restrict__ beta) {
  __shared__ double red[32];
  const double* xj = X + (long)j * m;
  const long c = (long)j + blockIdx.x;
  double* y = X + c * m;
  const bool upd = blockIdx.x != 0;
  const int tid = threadIdx.x;
  const long base = (long)j + 1 + tid;
  const double alpha = xj[j];
  const double yj = y[j];
  double xv[ITEMS], yv[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const long i = base + (long)QR_THREADS * k;
    xv[k] = i < m ? xj[i] : 0.0;
    yv[k] = (upd && i < m) ? y[i] : 0.0;
  }
  double s = 0.0, d = 0.0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    s = fma(xv[k], xv[k], s);
    d = fma(xv[k], yv[k], d);
  }
  qr_block_sum2(s, d, red);
  double t = 0.0, b = alpha, sc = 0.0;
  if (s > 0.0) {
    const double nrm = sqrt(fma(alpha, alpha, s));
    b = alpha >= 0.0 ? -nrm : nrm;
    t = (b - alpha) / b;
    sc = 1.0 / (alpha - b);
  }
  if (!upd) {
    double* v = Vt + (long)j * m;
    for (long i = tid; i <= j; i += QR_THREADS) v[i] = i < j ? 0.0 : 1.0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const long i = base + (long)QR_THREADS * k;
      if (i < m) v[i] = xv[k] * sc;
    }
    if (tid == 0) {
      tau[j] = t;
      beta[j] = b;
    }
    return;
  }
  if (t == 0.0) return;
  const double f = t * fma(sc, d, yj);
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const long i = base + (long)QR_THREADS * k;
    if (i < m) y[i] = fma(-f * sc, xv[k], yv[k]);
  }
  if (tid == 0) y[j] = yj - f;
}

// Qt (k x m) = [I_k | 0]
__global__ void qr_eye_kernel(double* __restrict__ Qt, long k, long m) {
  const long tot = k * m;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long r = i / m, col = i - r * m;
    Qt[i] = (r == col) ? 1.0 : 0.0;
  }
}

// R[r][c] (k x n): beta on the diagonal, X[c][r] above it, zero below
__global__ void qr_r_kernel(const double* __restrict__ X, long m, const double* __restrict__ beta, long k, long n,
                            double* __restrict__ R) {
  const long tot = k * n;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long r = i / n, c = i - r * n;
    R[i] = c < r ? 0.0 : (c == r ? beta[r] : X[c * m + r]);
  }
}

__global__ void swap01_kernel(const double* __restrict__ A, long d0, long d1, long d2, double* __restrict__ out) {
  const long tot = d0 * d1 * d2;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long i2 = i % d2, r = i / d2, i1 = r % d1, i0 = r / d1;
    out[(i1 * d0 + i0) * d2 + i2] = A[i];
  }
}

__global__ void concat3_kernel(const double* __restrict__ A, long a0, long a1, long a2, int ax,
                               const double* __restrict__ P, long p, double* __restrict__ out) {
  const long o0 = a0 + (ax == 0 ? p : 0), o1 = a1 + (ax == 1 ? p : 0), o2 = a2 + (ax == 2 ? p : 0);
  const long tot = o0 * o1 * o2;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long i2 = i % o2, r = i / o2, i1 = r % o1, i0 = r / o1;
    const long idx[3] = {i0, i1, i2};
    const long lim[3] = {a0, a1, a2};
    if (idx[ax] < lim[ax]) {
      out[i] = A[(i0 * a1 + i1) * a2 + i2];
    } else {
      const long q0 = ax == 0 ? i0 - a0 : i0, q1 = ax == 1 ? i1 - a1 : i1, q2 = ax == 2 ? i2 - a2 : i2;
      const long p1 = ax == 1 ? p : a1, p2 = ax == 2 ? p : a2;
      out[i] = P[(q0 * p1 + q1) * p2 + q2];
    }
  }
}

unsigned grid1(long tot) { return (unsigned)std::max<long>(1, std::min<long>(4096, (tot + 255) / 256)); }

}  // namespace

void householder_qr(tn_ctx* c, double* X, long m, long n, double* Qt, double* R) {
  TN_REQUIRE(m >= 1 && n >= 1, TN_E_INVALID, "qr: empty matrix");
  const long k = std::min(m, n);
  double* Vt = (double*)c->workspace(WS_QR_V, (size_t)k * m * sizeof(double));
  double* tb = (double*)c->workspace(WS_QR_S, (size_t)(2 * k + WY_NB * WY_NB) * sizeof(double));
  double* tau = tb;
  double* beta = tb + k;
  double* S = tb + 2 * k;
  for (long j = 0; j < k; ++j) {
    const long below = m - j - 1;
    const dim3 grid((unsigned)(n - j));
    if (below <= QR_THREADS * 1L)
      hipLaunchKernelGGL(qr_step_reg_kernel<1>, grid, dim3(QR_THREADS), 0, c->stream, X, m, (int)j, Vt, tau, beta);
    else if (below <= QR_THREADS * 4L)
      hipLaunchKernelGGL(qr_step_reg_kernel<4>, grid, dim3(QR_THREADS), 0, c->stream, X, m, (int)j, Vt, tau, beta);
    else if (below <= QR_THREADS * 16L)
      hipLaunchKernelGGL(qr_step_reg_kernel<16>, grid, dim3(QR_THREADS), 0, c->stream, X, m, (int)j, Vt, tau, beta);
    else if (below <= QR_THREADS * 32L)
      hipLaunchKernelGGL(qr_step_reg_kernel<32>, grid, dim3(QR_THREADS), 0, c->stream, X, m, (int)j, Vt, tau, beta);
    else
      hipLaunchKernelGGL(qr_step_kernel, grid, dim3(256), 0, c->stream, X, m, (int)j, Vt, tau, beta);
    TN_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(qr_r_kernel, dim3(grid1(k * n)), dim3(256), 0, c->stream, X, m, beta, k, n, R);
  hipLaunchKernelGGL(qr_eye_kernel, dim3(grid1(k * m)), dim3(256), 0, c->stream, Qt, k, m);
  TN_HIP(hipGetLastError());
  double* W = (double*)c->workspace(WS_QR_W, (size_t)WY_NB * k * sizeof(double));
  for (long j0 = ((k - 1) / WY_NB) * WY_NB; j0 >= 0; j0 -= WY_NB) {
    const int nb = (int)std::min<long>(WY_NB, k - j0);
    const long cols = m - j0;   // columns touched by the block
    const long rows = k - j0;   // rows of Q^T that can be nonzero there
    const double* Vs = Vt + j0 * m + j0;
    fill_f64(c, S, 0.0, WY_NB * WY_NB);
    {  // S = Vs Vs^T
      GemmDesc g;
      g.M = nb; g.N = nb; g.K = cols;
      g.A = Vs; g.lda = m;
      g.B = Vs; g.ldb = m; g.tB = 1;
      g.C = S; g.ldc = WY_NB;
      gemm(c, Prec::F64, g);
    }
    {  // W = Vs Qt[j0:, j0:]^T   (nb x rows)
      GemmDesc g;
      g.M = nb; g.N = rows; g.K = cols;
      g.A = Vs; g.lda = m;
      g.B = Qt + j0 * m + j0; g.ldb = m; g.tB = 1;
      g.C = W; g.ldc = rows;
      gemm(c, Prec::F64, g);
    }
    wy_tsolve(c, S, tau, (int)j0, nb, W, rows);
    {  // Qt[j0:, j0:] -= W^T Vs
      GemmDesc g;
      g.M = rows; g.N = cols; g.K = nb;
      g.A = W; g.lda = rows; g.tA = 1;
      g.B = Vs; g.ldb = m;
      g.C = Qt + j0 * m + j0; g.ldc = m;
      g.alpha = -1.0; g.beta = 1.0;
      gemm(c, Prec::F64, g);
    }
  }
}

void swap01_f64(tn_ctx* c, const double* A, long d0, long d1, long d2, double* out) {
  const long tot = d0 * d1 * d2;
  if (tot <= 0) return;
  hipLaunchKernelGGL(swap01_kernel, dim3(grid1(tot)), dim3(256), 0, c->stream, A, d0, d1, d2, out);
  TN_HIP(hipGetLastError());
}

void concat3_f64(tn_ctx* c, const double* A, const long* dims, int ax, const double* P, long p, double* out) {
  TN_REQUIRE(ax >= 0 && ax < 3 && p >= 0, TN_E_INVALID, "concat: axis");
  long tot = 1;
  for (int i = 0; i < 3; ++i) tot *= dims[i] + (i == ax ? p : 0);
  if (tot <= 0) return;
  hipLaunchKernelGGL(concat3_kernel, dim3(grid1(tot)), dim3(256), 0, c->stream, A, dims[0], dims[1], dims[2], ax, P, p,
                     out);
  TN_HIP(hipGetLastError());
}

}  // namespace tn
