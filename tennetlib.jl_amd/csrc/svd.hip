// Truncated SVD for replacebond!: one-sided (Hestenes) block Jacobi in FP64.
//
// The p = min(m, n) vectors of length q = max(m, n) (columns of M when
// n <= m, rows of M otherwise) are stored as rows of X (p x q).  They are
// split into blocks of 32 rows; a sweep visits every pair of blocks once in a
// round-robin tournament (nb - 1 rounds of nb/2 independent pairs).  For each
// pair of 64 rows:
//   1. the Gram matrix G = X_pair X_pair^T is formed afresh (FP64 MFMA,
//      split over q and summed in a fixed order),
//   2. G is diagonalised by parallel-order cyclic Jacobi in LDS with the
//      relative threshold |g_ij| > tol * sqrt(g_ii g_jj) (Demmel & Veselic:
//      on a nearly-diagonal scaled Gram this keeps high relative accuracy),
//      accumulating the rotation J (64 x 64),
//   3. X_pair <- J^T X_pair and Qt_pair <- J^T Qt_pair in place (FP64 MFMA).
// Sweeps stop when a whole sweep performs no rotation.  The singular values
// are the row norms of X; the rows of Qt are the other singular vectors.
// Small singular values keep high relative accuracy, which the truncation
// rule (a cumulative sum of discarded weights against cutoff) relies on.
#include "tn_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace tn {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BR = 32;      // rows per block
constexpr int PR = 2 * BR;  // rows per pair
constexpr int LDP = PR + 1;
constexpr int GBK = 32;     // k-chunk of the Gram kernel

// partial Gram of one pair over one k-range; grid (npairs, ksplit)
__global__ __launch_bounds__(256) void jac_gram_kernel(const double* __restrict__ X, long q, long ldx,
                                                       const int* __restrict__ pairs, long kc,
                                                       double* __restrict__ part) {
  __shared__ double Xs[GBK][PR + 2];
  const int pr = blockIdx.x, ks = blockIdx.y, nks = gridDim.y;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  const long kbeg = (long)ks * kc, kend = min(q, kbeg + kc);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0, 0, 0, 0};
  for (long k0 = kbeg; k0 < kend; k0 += GBK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < (PR * GBK) / 256; ++i) {
      const int idx = tid + i * 256;
      const int kk = idx % GBK, rr = idx / GBK;
      const long row = (rr < BR) ? ((long)bi * BR + rr) : ((long)bj * BR + rr - BR);
      const long k = k0 + kk;
      Xs[kk][rr] = (k < kend) ? X[row * ldx + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i] = Xs[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) fb[j] = Xs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }
  double* out = part + ((long)pr * nks + ks) * PR * PR;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * 32 + i * 16 + (lane >> 4) + 4 * r;  // f64 MFMA output layout
        const int col = wn * 32 + j * 16 + (lane & 15);
        out[row * PR + col] = acc[i][j][r];
      }
}

// Jacobi diagonalisation of one 64 x 64 pair Gram; grid (npairs)
__global__ __launch_bounds__(256) void jac_inner_kernel(const double* __restrict__ part, int nks, double tol,
                                                        int max_inner, double* __restrict__ Jout,
                                                        int* __restrict__ rotcount) {
  __shared__ double Gs[PR][LDP];
  __shared__ double Js[PR][LDP];
  __shared__ double cs[PR / 2], sn[PR / 2];
  __shared__ int pi_[PR / 2], pj_[PR / 2], fl[PR / 2];
  __shared__ int any;
  const int pr = blockIdx.x, tid = threadIdx.x;
  const double* src = part + (long)pr * nks * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) {
    double s = 0.0;
    for (int k = 0; k < nks; ++k) s += src[(long)k * PR * PR + idx];
    Gs[idx / PR][idx % PR] = s;
    Js[idx / PR][idx % PR] = (idx / PR == idx % PR) ? 1.0 : 0.0;
  }
  int total = 0;
  __syncthreads();
  for (int sweep = 0; sweep < max_inner; ++sweep) {
    if (tid == 0) any = 0;
    __syncthreads();
    for (int step = 0; step < PR - 1; ++step) {
      if (tid < PR / 2) {
        // circle method: position 0 fixed, the others rotate by `step`
        const int pa = tid, pb = PR - 1 - tid;
        const int i = (pa == 0) ? 0 : 1 + ((pa - 1 + step) % (PR - 1));
        const int j = 1 + ((pb - 1 + step) % (PR - 1));
        const double a = Gs[i][i], b = Gs[j][j], g = Gs[i][j];
        int f = 0;
        double c = 1.0, s = 0.0;
        if (a > 0.0 && b > 0.0 && fabs(g) > tol * sqrt(a) * sqrt(b)) {
          const double zeta = (b - a) / (2.0 * g);
          double t;
          if (fabs(zeta) > 1e150) t = 0.5 / zeta;
          else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          c = 1.0 / sqrt(1.0 + t * t);
          s = c * t;
          f = 1;
          any = 1;
        }
        pi_[tid] = i; pj_[tid] = j; cs[tid] = c; sn[tid] = s; fl[tid] = f;
      }
      __syncthreads();
      // rows: row_i' = c row_i - s row_j ; row_j' = s row_i + c row_j
      for (int idx = tid; idx < (PR / 2) * PR; idx += 256) {
        const int p = idx / PR, col = idx % PR;
        if (fl[p]) {
          const int i = pi_[p], j = pj_[p];
          const double c = cs[p], s = sn[p];
          const double gi = Gs[i][col], gj = Gs[j][col];
          Gs[i][col] = c * gi - s * gj;
          Gs[j][col] = s * gi + c * gj;
        }
      }
      __syncthreads();
      // columns of G and J
      for (int idx = tid; idx < (PR / 2) * PR; idx += 256) {
        const int p = idx / PR, row = idx % PR;
        if (fl[p]) {
          const int i = pi_[p], j = pj_[p];
          const double c = cs[p], s = sn[p];
          const double gi = Gs[row][i], gj = Gs[row][j];
          Gs[row][i] = c * gi - s * gj;
          Gs[row][j] = s * gi + c * gj;
          const double ji = Js[row][i], jj = Js[row][j];
          Js[row][i] = c * ji - s * jj;
          Js[row][j] = s * ji + c * jj;
        }
      }
      __syncthreads();
    }
    const int a_ = any;
    __syncthreads();
    if (!a_) break;
    ++total;
  }
  double* J = Jout + (long)pr * PR * PR;
  if (total > 0) {
    // J is a product of a few hundred plane rotations and is orthogonal only
    // to ~1e-14; one Newton-Schulz step J <- J (3I - J^T J) / 2 squares that
    // error, so repeated application keeps the vectors' norms exact.
    for (int idx = tid; idx < PR * PR; idx += 256) {
      const int i = idx / PR, j = idx % PR;
      double s = 0.0;
      for (int k = 0; k < PR; ++k) s = fma(Js[k][i], Js[k][j], s);
      Gs[i][j] = (i == j ? 1.5 : 0.0) - 0.5 * s;
    }
    __syncthreads();
    for (int idx = tid; idx < PR * PR; idx += 256) {
      const int r = idx / PR, cc = idx % PR;
      double s = 0.0;
      for (int k = 0; k < PR; ++k) s = fma(Js[r][k], Gs[k][cc], s);
      J[idx] = s;
    }
  } else {
    for (int idx = tid; idx < PR * PR; idx += 256) J[idx] = Js[idx / PR][idx % PR];
  }
  if (tid == 0) rotcount[pr] = total;
}

// rows of Q (unit vectors up to rounding) rescaled to unit norm; grid (rows)
__global__ __launch_bounds__(256) void normalize_rows_kernel(double* __restrict__ Q, long L, long ld) {
  __shared__ double red[4];
  __shared__ double inv;
  const long r = blockIdx.x;
  double s = 0.0;
  for (long j = threadIdx.x; j < L; j += blockDim.x) s = fma(Q[r * ld + j], Q[r * ld + j], s);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = red[0] + red[1] + red[2] + red[3];
    inv = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
  }
  __syncthreads();
  for (long j = threadIdx.x; j < L; j += blockDim.x) Q[r * ld + j] *= inv;
}

// X_pair <- J^T X_pair (and Qt_pair) in place; grid (npairs, tilesX + tilesQ)
__global__ __launch_bounds__(256) void jac_apply_kernel(double* __restrict__ X, long q, long ldx,
                                                        double* __restrict__ Q, long qq, long ldq,
                                                        const int* __restrict__ pairs,
                                                        const double* __restrict__ Jall,
                                                        const int* __restrict__ rotcount, int tilesX) {
  __shared__ double Jt[PR][PR + 2];   // Jt[k][m] = J[k][m]  (A^T operand, k-major)
  __shared__ double Ys[PR][PR + 2];   // Ys[k][n]
  const int pr = blockIdx.x;
  if (rotcount[pr] == 0) return;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  double* base;
  long L, ld, c0;
  if ((int)blockIdx.y < tilesX) { base = X; L = q; ld = ldx; c0 = (long)blockIdx.y * PR; }
  else { base = Q; L = qq; ld = ldq; c0 = (long)(blockIdx.y - tilesX) * PR; }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double* J = Jall + (long)pr * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) {
    const int r = idx / PR, cc = idx % PR;
    Jt[r][cc] = J[idx];
    const long row = (r < BR) ? ((long)bi * BR + r) : ((long)bj * BR + r - BR);
    const long col = c0 + cc;
    Ys[r][cc] = (col < L) ? base[row * ld + col] : 0.0;
  }
  __syncthreads();
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < PR; kk += 4) {
    const int kr = kk + (lane >> 4);
    double fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) fa[i] = Jt[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < 2; ++j) fb[j] = Ys[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = wm * 32 + i * 16 + (lane >> 4) + 4 * r;  // f64 MFMA output layout
        const int cc = wn * 32 + j * 16 + (lane & 15);
        const long row = (rr < BR) ? ((long)bi * BR + rr) : ((long)bj * BR + rr - BR);
        const long col = c0 + cc;
        if (col < L) base[row * ld + col] = acc[i][j][r];
      }
}

__global__ void set_identity_kernel(double* __restrict__ Q, long n) {
  const long total = n * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    Q[i] = (i / n == i % n) ? 1.0 : 0.0;
}

struct Schedule {
  long nb = 0;
  std::vector<int> pairs;  // (nb-1) rounds x nb/2 pairs x 2
};

Schedule make_schedule(long nb) {
  Schedule s;
  s.nb = nb;
  const long np = nb / 2;
  s.pairs.resize((size_t)(nb - 1) * np * 2);
  for (long r = 0; r < nb - 1; ++r)
    for (long t = 0; t < np; ++t) {
      const long pa = t, pb = nb - 1 - t;
      const long i = (pa == 0) ? 0 : 1 + ((pa - 1 + r) % (nb - 1));
      const long j = 1 + ((pb - 1 + r) % (nb - 1));
      s.pairs[(size_t)(r * np + t) * 2] = (int)std::min(i, j);
      s.pairs[(size_t)(r * np + t) * 2 + 1] = (int)std::max(i, j);
    }
  return s;
}

}  // namespace

void svd_jacobi(tn_ctx* c, const double* M, long m, long n, SvdOut& out) {
  TN_REQUIRE(m > 0 && n > 0, TN_E_INVALID, "svd: empty matrix");
  const long p = std::min(m, n), q = std::max(m, n);
  const long ppad = ((p + PR - 1) / PR) * PR;
  const long nb = ppad / BR;
  out.p = p;
  out.q = q;
  out.vec_cols = (n <= m);
  double* X = (double*)c->workspace(WS_SVD_X, (size_t)ppad * q * sizeof(double));
  double* Q = (double*)c->workspace(WS_SVD_Q, (size_t)ppad * ppad * sizeof(double));
  out.X = X;
  out.Qt = Q;
  // vectors as rows of X
  if (out.vec_cols) transpose_f64(c, M, m, n, X);  // X = M^T (n x m)
  else copy_f64(c, X, M, p * q);                    // X = M   (m x n)
  if (ppad > p) fill_f64(c, X + p * q, 0.0, (ppad - p) * q);
  {
    long tot = ppad * ppad;
    int g = (int)std::min<long>(4096, (tot + 255) / 256);
    hipLaunchKernelGGL(set_identity_kernel, dim3(g), dim3(256), 0, c->stream, Q, ppad);
    TN_HIP(hipGetLastError());
  }
  const Schedule sch = make_schedule(nb);
  const long np = nb / 2;
  const long nrounds = nb - 1;
  int* dpairs = (int*)c->workspace(WS_SVD_PAIRS, sch.pairs.size() * sizeof(int));
  TN_HIP(hipMemcpyAsync(dpairs, sch.pairs.data(), sch.pairs.size() * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  // k-split of the Gram: aim for ~512 workgroups per round
  long ksplit = std::max<long>(1, std::min<long>((512 + np - 1) / np, (q + 127) / 128));
  long kc = (q + ksplit - 1) / ksplit;
  kc = ((kc + GBK - 1) / GBK) * GBK;
  ksplit = (q + kc - 1) / kc;
  double* part = (double*)c->workspace(WS_SVD_GRAM, (size_t)np * ksplit * PR * PR * sizeof(double));
  double* J = (double*)c->workspace(WS_SVD_J, (size_t)np * PR * PR * sizeof(double));
  int* rot = (int*)c->workspace(WS_SVD_CNT, (size_t)nrounds * np * sizeof(int));
  const double tol = 2.220446049250313e-16 * std::max(1.0, std::sqrt((double)q));
  const int tilesX = (int)((q + PR - 1) / PR), tilesQ = (int)((ppad + PR - 1) / PR);
  std::vector<int> hrot((size_t)nrounds * np);
  const int max_sweeps = 60;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    for (long r = 0; r < nrounds; ++r) {
      const int* pr = dpairs + r * np * 2;
      hipLaunchKernelGGL(jac_gram_kernel, dim3((unsigned)np, (unsigned)ksplit), dim3(256), 0, c->stream, X, q,
                         q, pr, kc, part);
      TN_HIP(hipGetLastError());
      hipLaunchKernelGGL(jac_inner_kernel, dim3((unsigned)np), dim3(256), 0, c->stream, part, (int)ksplit, tol,
                         4, J, rot + r * np);
      TN_HIP(hipGetLastError());
      hipLaunchKernelGGL(jac_apply_kernel, dim3((unsigned)np, (unsigned)(tilesX + tilesQ)), dim3(256), 0,
                         c->stream, X, q, q, Q, ppad, ppad, pr, J, rot + r * np, tilesX);
      TN_HIP(hipGetLastError());
    }
    TN_HIP(hipMemcpyAsync(hrot.data(), rot, hrot.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    long tot = 0;
    for (int v : hrot) tot += v;
    if (tot == 0) break;
  }
  out.sweeps = sweep + 1;
  hipLaunchKernelGGL(normalize_rows_kernel, dim3((unsigned)p), dim3(256), 0, c->stream, Q, p, ppad);
  TN_HIP(hipGetLastError());
  double* s2 = (double*)c->workspace(WS_SVD_S, (size_t)p * sizeof(double));
  row_norms2(c, X, p, q, q, s2);
  out.s.resize(p);
  TN_HIP(hipMemcpyAsync(out.s.data(), s2, p * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
  for (auto& v : out.s) v = std::sqrt(std::max(0.0, v));
}

TruncRule truncate_weights(const std::vector<double>& Pin, long maxdim, long mindim, double cutoff) {
  // NDTensors truncate!! with the relative cutoff (ITensors' default)
  std::vector<double> P = Pin;
  const long origm = (long)P.size();
  TruncRule tr{0, 0.0, 0.0};
  if (origm == 0) return tr;
  for (long k = origm - 1; k >= 0; --k) {
    if (P[k] >= 0.0) break;
    P[k] = 0.0;
  }
  if (origm == 1) { tr.n = 1; tr.docut = std::fabs(P[0]) / 2; return tr; }
  long n = origm;
  double truncerr = 0.0;
  maxdim = std::min(maxdim, origm);
  mindim = std::max<long>(1, std::min(mindim, maxdim));
  while (n > maxdim) { truncerr += P[n - 1]; --n; }
  double scale = 0.0;
  for (double v : P) scale += v;
  if (scale == 0.0) scale = 1.0;
  while (n > mindim && truncerr + P[n - 1] <= cutoff * scale) { truncerr += P[n - 1]; --n; }
  truncerr /= scale;
  n = std::max<long>(n, 1);
  double docut = 0.0;
  if (n < origm) {
    docut = (P[n - 1] + P[n]) / 2;
    if (std::fabs(P[n - 1] - P[n]) < 1e-3 * P[n - 1]) docut += 1e-3 * P[n - 1];
  }
  tr.n = n;
  tr.truncerr = truncerr;
  tr.docut = docut;
  return tr;
}

}  // namespace tn
