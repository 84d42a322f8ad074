// Quantum-number (U(1), U(1)xU(1)) block kernels of the truncating
// factorisation (replacebond! of a QN-conserving two-site tensor):
//
//  * qn_small_eig_kernel: every QN sector of the combined (link, site) index
//    whose density-matrix block is at most 32 x 32 gets one wavefront, in one
//    launch: the block rho_q = M_q M_q^T is formed from the sector's gathered
//    rows of M (staged through LDS 64 columns at a time), then diagonalised by
//    cyclic Jacobi rotations in LDS (relative accuracy for these positive
//    semi-definite blocks); eigenvalues (unsorted) and eigenvectors (as rows)
//    go to the sector's slots of the block outputs.
//  * qn_scatter_kernel: the kept eigenvectors of every sector become the
//    columns of the orthonormal factor (rows of its transpose), each written
//    to the rows of M that belong to its sector; everything else stays zero.
#include "tn_internal.h"

#include <algorithm>
#include <cmath>

namespace tn {
namespace {

constexpr int QB = 32;  // largest block handled by qn_small_eig_kernel

// one wavefront per block; blk[b] = {row offset in M, size p, output offset of
// the eigenvalues, output offset of the eigenvector rows}
__global__ __launch_bounds__(64) void qn_small_eig_kernel(const double* __restrict__ M, long n,
                                                          const int* __restrict__ blk, double* __restrict__ w,
                                                          double* __restrict__ V, int max_sweeps) {
  __shared__ double A[QB][QB + 1];
  __shared__ double U[QB][QB + 1];
  __shared__ double stage[QB][65];
  const int b = blockIdx.x;
  const int r0 = blk[4 * b + 0], p = blk[4 * b + 1], wo = blk[4 * b + 2], vo = blk[4 * b + 3];
  const int lane = threadIdx.x;
  // rho = M_q M_q^T: pairs (i, j), i <= j, spread over the lanes
  const int npair = p * (p + 1) / 2;
  double acc[9];
#pragma unroll
  for (int u = 0; u < 9; ++u) acc[u] = 0.0;
  for (long c0 = 0; c0 < n; c0 += 64) {
    const int cw = (int)min((long)64, n - c0);
    for (int i = 0; i < p; ++i) stage[i][lane] = (lane < cw) ? M[(long)(r0 + i) * n + c0 + lane] : 0.0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 9; ++u) {
      const int t = lane + 64 * u;
      if (t < npair) {
        // t -> (i, j) with i <= j, row-major over the upper triangle
        int i = 0, rem = t;
        while (rem >= p - i) { rem -= p - i; ++i; }
        const int j = i + rem;
        double s = 0.0;
        for (int cc = 0; cc < 64; ++cc) s += stage[i][cc] * stage[j][cc];
        acc[u] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 9; ++u) {
    const int t = lane + 64 * u;
    if (t < npair) {
      int i = 0, rem = t;
      while (rem >= p - i) { rem -= p - i; ++i; }
      const int j = i + rem;
      A[i][j] = acc[u];
      A[j][i] = acc[u];
    }
  }
  for (int t = lane; t < QB * QB; t += 64) U[t / QB][t % QB] = (t / QB == t % QB) ? 1.0 : 0.0;
  __syncthreads();
  // cyclic Jacobi: rotations (pq) one at a time, applied by the lanes
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int t = lane; t < p * p; t += 64) {
      const int i = t / p, j = t % p;
      const double x = A[i][j] * A[i][j];
      if (i == j) dia += x; else off += x;
    }
    off = wave_sum64(off);
    dia = wave_sum64(dia);
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int pp = 0; pp < p - 1; ++pp) {
      for (int qq = pp + 1; qq < p; ++qq) {
        const double apq = A[pp][qq];
        if (apq == 0.0) continue;  // uniform: every lane reads the same entry
        const double app = A[pp][pp], aqq = A[qq][qq];
        const double zeta = (aqq - app) / (2.0 * apq);
        const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
        __syncthreads();
        if (lane < p) {  // columns pp, qq of A
          const double akp = A[lane][pp], akq = A[lane][qq];
          A[lane][pp] = cs * akp - sn * akq;
          A[lane][qq] = sn * akp + cs * akq;
        } else if (lane >= 32 && lane - 32 < p) {  // columns pp, qq of U
          const int k = lane - 32;
          const double ukp = U[k][pp], ukq = U[k][qq];
          U[k][pp] = cs * ukp - sn * ukq;
          U[k][qq] = sn * ukp + cs * ukq;
        }
        __syncthreads();
        if (lane < p) {  // rows pp, qq of A
          const double apk = A[pp][lane], aqk = A[qq][lane];
          A[pp][lane] = cs * apk - sn * aqk;
          A[qq][lane] = sn * apk + cs * aqk;
        }
        __syncthreads();
      }
    }
  }
  for (int i = lane; i < p; i += 64) w[wo + i] = A[i][i];
  for (int t = lane; t < p * p; t += 64) {
    const int i = t / p, r = t % p;  // eigenvector i (column i of U), component r
    V[vo + (long)i * p + r] = U[r][i];
  }
}

// kept column j: eigenvector col[4j+2] of the sector whose rows of M are
// rows[col[4j+1] .. +col[4j+3]), stored at V + col[4j] (rows of length p);
// out (rows x k, ld k) gets out[rows[roff + r]][j], or with trans the k x rows
// transpose out[j][rows[roff + r]]
__global__ __launch_bounds__(256) void qn_scatter_kernel(const double* __restrict__ V, const int* __restrict__ col,
                                                         const int* __restrict__ rows, int k, long ldo, int trans,
                                                         double* __restrict__ out) {
  const int j = blockIdx.x;
  const long vo = col[4 * j + 0];
  const int roff = col[4 * j + 1], idx = col[4 * j + 2], p = col[4 * j + 3];
  for (int r = threadIdx.x; r < p; r += 256) {
    const double x = V[vo + (long)idx * p + r];
    const long row = rows[roff + r];
    if (trans) out[(long)j * ldo + row] = x;
    else out[row * ldo + j] = x;
  }
}

}  // namespace

void qn_small_eig(tn_ctx* c, const double* M, long n, const int* blk_dev, int nblk, double* w, double* V) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(qn_small_eig_kernel, dim3((unsigned)nblk), dim3(64), 0, c->stream, M, n, blk_dev, w, V, 60);
  TN_HIP(hipGetLastError());
}

void qn_scatter(tn_ctx* c, const double* V, const int* col_dev, const int* rows_dev, int k, long ldo, bool trans,
                double* out) {
  if (k <= 0) return;
  hipLaunchKernelGGL(qn_scatter_kernel, dim3((unsigned)k), dim3(256), 0, c->stream, V, col_dev, rows_dev, k, ldo,
                     trans ? 1 : 0, out);
  TN_HIP(hipGetLastError());
}

}  // namespace tn
