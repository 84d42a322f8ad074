// Quantum-number (U(1), U(1)xU(1)) block kernels of the truncating
// factorisation (replacebond! of a QN-conserving two-site tensor):
//
//  * qn_jacobi_kernel: every QN sector of the combined (link, site) index
//    whose density-matrix block rho_q is at most 96 x 96 gets one workgroup,
//    in one launch per size class (32, 64, 96): rho_q, formed beforehand by
//    one block-sparse GEMM launch, is diagonalised in LDS by parallel-order
//    cyclic Jacobi; eigenvalues (unsorted) and eigenvectors (as rows) go to
//    the sector's slots of the block outputs.
//  * qn_scatter_kernel: the kept eigenvectors of every sector become the
//    columns of the orthonormal factor (rows of its transpose), each written
//    to the rows of M that belong to its sector; everything else stays zero.
#include "tn_internal.h"

#include <algorithm>
#include <cmath>

namespace tn {
namespace {


// One workgroup per block.  blk[b] = {row offset of the block in the stacked
// density matrices R (ld QN_RHO_LD), size p <= PB, output offset of the
// eigenvalues, output offset of the eigenvector rows}.  The block and its
// eigenvector matrix U sit in LDS; cyclic Jacobi in the round-robin
// (parallel) order: each round applies p/2 disjoint rotations at once,
// A <- J^T A J and U <- U J, and a sweep is p - 1 rounds (p rounded up to
// even; pairs with the padding index do nothing).  A rotation sets its pair's
// off-diagonal entries to zero; pairs with |a_ij| <= 1e-3 eps max_i a_ii are
// left alone, and the sweeps end when a whole sweep rotates nothing (every
// off-diagonal entry below that bound: eigenvalues to ~p 1e-3 eps |A|).
constexpr int JT = 1024;  // threads of qn_jacobi_kernel: 16 waves hide the LDS latency of one block per CU

template <int PB>
__global__ __launch_bounds__(JT) void qn_jacobi_kernel(const double* __restrict__ R, const int* __restrict__ blk,
                                                        double* __restrict__ w, double* __restrict__ V,
                                                        int max_sweeps) {
  constexpr int LD = PB + 1;
  extern __shared__ double lds[];
  double* A = lds;             // PB x LD
  double* U = lds + PB * LD;   // PB x LD
  __shared__ double rc[PB / 2], rs[PB / 2];
  __shared__ int ri[PB / 2], rj[PB / 2];
  __shared__ double red[JT / 64];
  __shared__ int nrot;
  const int b = blockIdx.x;
  const int r0 = blk[4 * b + 0], p = blk[4 * b + 1], wo = blk[4 * b + 2], vo = blk[4 * b + 3];
  const int tid = threadIdx.x;
  double dmax = 0.0;
  for (int t = tid; t < PB * PB; t += JT) {
    const int i = t / PB, j = t % PB;
    if (i >= p || j >= p) continue;
    const double x = R[(long)(r0 + i) * QN_RHO_LD + j];
    A[i * LD + j] = x;
    U[i * LD + j] = (i == j) ? 1.0 : 0.0;
    if (i == j) dmax = fmax(dmax, fabs(x));
  }
  for (int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = dmax;
  __syncthreads();
  for (int k = 0; k < JT / 64; ++k) dmax = fmax(dmax, red[k]);
  const double tol = 1e-3 * 2.220446049250313e-16 * dmax;
  const int n = p + (p & 1);  // players of the round-robin, padded to even
  const int np = n / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) nrot = 0;
    __syncthreads();
    for (int round = 0; round < n - 1; ++round) {
      if (tid < np) {
        // circle method: player n-1 fixed, the others rotate
        int i, j;
        if (tid == 0) { i = round; j = n - 1; }
        else { i = (round + tid) % (n - 1); j = (round - tid + n - 1) % (n - 1); }
        double c = 1.0, sn = 0.0;
        if (i < p && j < p) {
          const double apq = A[i * LD + j];
          if (fabs(apq) > tol) {
            const double zeta = (A[j * LD + j] - A[i * LD + i]) / (2.0 * apq);
            const double az = fabs(zeta);
            const double t = (az > 1e150) ? 0.5 / zeta
                                           : (zeta >= 0.0 ? 1.0 : -1.0) / (az + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
            atomicAdd(&nrot, 1);
          } else {
            i = j = PB;  // nothing to do for this pair
          }
        }
        ri[tid] = i; rj[tid] = j; rc[tid] = c; rs[tid] = sn;
      }
      __syncthreads();
      // columns i, j of A and U (every row)
      for (int t = tid; t < np * PB * 2; t += JT) {
        int k = t / PB;
        const int r = t % PB;
        double* X = A;
        if (k >= np) { k -= np; X = U; }
        const int i = ri[k], j = rj[k];
        if (i >= p || j >= p || r >= p) continue;
        const double c = rc[k], sn = rs[k];
        const double xi = X[r * LD + i], xj = X[r * LD + j];
        X[r * LD + i] = c * xi - sn * xj;
        X[r * LD + j] = sn * xi + c * xj;
      }
      __syncthreads();
      // rows i, j of A (every column); the pair's own entries become zero
      for (int t = tid; t < np * PB; t += JT) {
        const int k = t / PB, r = t % PB;
        const int i = ri[k], j = rj[k];
        if (i >= p || j >= p || r >= p) continue;
        const double c = rc[k], sn = rs[k];
        const double xi = A[i * LD + r], xj = A[j * LD + r];
        A[i * LD + r] = (r == j) ? 0.0 : c * xi - sn * xj;
        A[j * LD + r] = (r == i) ? 0.0 : sn * xi + c * xj;
      }
      __syncthreads();
    }
    const int nr = nrot;  // read after the last round's barrier
    __syncthreads();      // every thread has read it before thread 0 resets it
    if (nr == 0) break;
  }
  for (int i = tid; i < p; i += JT) w[wo + i] = A[i * LD + i];
  for (int t = tid; t < p * p; t += JT) {
    const int i = t / p, r = t % p;  // eigenvector i (column i of U), component r
    V[vo + (long)i * p + r] = U[r * LD + i];
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

void qn_block_eig(tn_ctx* c, const double* R, const int* blk_dev, int nblk, int pmax, double* w, double* V) {
  if (nblk <= 0) return;
  TN_REQUIRE(pmax >= 1 && pmax <= QN_JACOBI_MAX, TN_E_INVALID, "QN block eigensolver: block too large");
  auto launch = [&](auto kern, int pb) {
    const size_t shm = (size_t)2 * pb * (pb + 1) * sizeof(double);
    raise_lds_limit((const void*)kern, shm);
    hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(JT), shm, c->stream, R, blk_dev, w, V, 40);
  };
  if (pmax <= 32) launch(qn_jacobi_kernel<32>, 32);
  else if (pmax <= 64) launch(qn_jacobi_kernel<64>, 64);
  else launch(qn_jacobi_kernel<96>, 96);
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
