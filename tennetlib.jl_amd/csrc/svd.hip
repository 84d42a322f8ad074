// Truncated factorisation for replacebond!: symmetric eigen-decomposition of
// the two-site density matrix on the device, in FP64.
//
// For phi (m x n) the Gram matrix G = phi phi^T (m x m, "left") or
// G = phi^T phi (n x n, "right") is formed with one FP64 MFMA GEMM and
// diagonalised by two-sided block Jacobi:
//   * indices are grouped in blocks of 32; a sweep visits every pair of
//     blocks once in a round-robin tournament (nb - 1 rounds of nb/2
//     disjoint pairs);
//   * per pair, the 64 x 64 principal sub-matrix is diagonalised in LDS by
//     parallel-order cyclic Jacobi (rotation of (i, j) when |g_ij| exceeds
//     max(tol_rel * sqrt(g_ii g_jj), tol_abs)), accumulating J;
//   * G <- J^T G J (rows, then columns of the pair) and V^T <- J^T V^T with
//     FP64 MFMA tiles.
// Sweeps stop when a sweep performs no rotation.  The eigenvalues are the
// weights P = s^2 with absolute accuracy ~eps * sum(P) (the accuracy of a
// density-matrix eigensolver), and the eigenvectors are orthonormal to
// rounding, so the orthonormal MPS factor is taken from them directly and
// the other factor is the exact product V^T phi (or phi V): no division by
// small singular values anywhere.
#include "tn_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>

namespace tn {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BR = 32;      // indices per block
constexpr int PR = 2 * BR;  // indices per pair
constexpr int LDP = PR + 1;

__device__ __forceinline__ long pair_index(int bi, int bj, int r) {
  return (r < BR) ? ((long)bi * BR + r) : ((long)bj * BR + r - BR);
}

// Jacobi diagonalisation of the 64 x 64 principal sub-matrix of G for each
// pair of a round; grid (npairs), 256 threads.  Thread t owns pair t/8 of a
// step and 8 columns (rows) of it, so every update loop has 8 independent
// LDS read-modify-writes in flight.
__global__ __launch_bounds__(256) void jac_inner_kernel(const double* __restrict__ G, long ldg,
                                                        const int* __restrict__ pairs, double tol_rel,
                                                        double tol_abs, int max_inner, double* __restrict__ Jout,
                                                        int* __restrict__ rotcount) {
  __shared__ double Gs[PR][LDP];
  __shared__ double Js[PR][LDP];
  __shared__ double cs[PR / 2], sn[PR / 2];
  __shared__ int pi_[PR / 2], pj_[PR / 2], fl[PR / 2];
  const int pr = blockIdx.x, tid = threadIdx.x;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  for (int idx = tid; idx < PR * PR; idx += 256) {
    const int r = idx / PR, cc = idx % PR;
    Gs[r][cc] = G[pair_index(bi, bj, r) * ldg + pair_index(bi, bj, cc)];
    Js[r][cc] = (r == cc) ? 1.0 : 0.0;
  }
  __syncthreads();
  int total = 0;
  const int p = tid >> 3, sub = tid & 7;  // pair slot and 8-wide column group
  for (int sweep = 0; sweep < max_inner; ++sweep) {
    int nsweep = 0;
    for (int step = 0; step < PR - 1; ++step) {
      int f = 0;
      if (tid < PR / 2) {
        // circle method: position 0 fixed, the others rotate by `step`
        const int pa = tid, pb = PR - 1 - tid;
        const int i = (pa == 0) ? 0 : 1 + ((pa - 1 + step) % (PR - 1));
        const int j = 1 + ((pb - 1 + step) % (PR - 1));
        const double a = Gs[i][i], b = Gs[j][j], g = Gs[i][j];
        double c = 1.0, s = 0.0;
        if (fabs(g) > fmax(tol_rel * sqrt(fabs(a)) * sqrt(fabs(b)), tol_abs)) {
          const double zeta = (b - a) / (2.0 * g);
          double t;
          if (fabs(zeta) > 1e150) t = 0.5 / zeta;
          else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          c = 1.0 / sqrt(1.0 + t * t);
          s = c * t;
          f = 1;
        }
        pi_[tid] = i; pj_[tid] = j; cs[tid] = c; sn[tid] = s; fl[tid] = f;
      }
      const int nrot = __syncthreads_count(f);
      if (nrot == 0) continue;  // uniform: no rotation in this step
      nsweep += nrot;
      if (fl[p]) {  // rows: row_i' = c row_i - s row_j ; row_j' = s row_i + c row_j
        const int i = pi_[p], j = pj_[p];
        const double c = cs[p], s = sn[p];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int col = sub * 8 + u;
          const double gi = Gs[i][col], gj = Gs[j][col];
          Gs[i][col] = c * gi - s * gj;
          Gs[j][col] = s * gi + c * gj;
        }
      }
      __syncthreads();
      if (fl[p]) {  // columns of G and J
        const int i = pi_[p], j = pj_[p];
        const double c = cs[p], s = sn[p];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int row = sub * 8 + u;
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
    if (nsweep == 0) break;
    ++total;
  }
  double* J = Jout + (long)pr * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) J[idx] = Js[idx / PR][idx % PR];
  if (tid == 0) rotcount[pr] = total;
}

// Row update: rows of each pair of A (ld, L columns) <- J^T rows.  Used for
// G (rows phase) and V^T.  grid (npairs, tiles), 64-column tiles.
// Column update (COLS): columns of each pair of G <- columns * J, 64-row tiles.
template <bool COLS>
__global__ __launch_bounds__(256) void jac_apply_kernel(double* __restrict__ A, long L, long ld,
                                                        const int* __restrict__ pairs,
                                                        const double* __restrict__ Jall,
                                                        const int* __restrict__ rotcount) {
  __shared__ double Xs[PR][PR + 2];   // A operand, [m][k]
  __shared__ double Ys[PR][PR + 2];   // B operand, [k][n]
  const int pr = blockIdx.x;
  if (rotcount[pr] == 0) return;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  const long t0 = (long)blockIdx.y * PR;  // first free index of the tile
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double* J = Jall + (long)pr * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) {
    const int r = idx / PR, cc = idx % PR;
    if (!COLS) {
      // out[rr][col] = sum_k J[k][rr] * A[pair k][col]: Xs = J^T, Ys = A rows
      Xs[r][cc] = J[cc * PR + r];
      const long col = t0 + cc;
      Ys[r][cc] = (col < L) ? A[pair_index(bi, bj, r) * ld + col] : 0.0;
    } else {
      // out[row][cc] = sum_k A[row][pair k] * J[k][cc]: Xs = A tile, Ys = J
      const long row = t0 + r;
      Xs[r][cc] = (row < L) ? A[row * ld + pair_index(bi, bj, cc)] : 0.0;
      Ys[r][cc] = J[r * PR + cc];
    }
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
    for (int i = 0; i < 2; ++i) fa[i] = Xs[wm * 32 + i * 16 + (lane & 15)][kr];
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
        if (!COLS) {
          const long col = t0 + cc;
          if (col < L) A[pair_index(bi, bj, rr) * ld + col] = acc[i][j][r];
        } else {
          const long row = t0 + rr;
          if (row < L) A[row * ld + pair_index(bi, bj, cc)] = acc[i][j][r];
        }
      }
}

// ---------------------------------------------------------------------------
// Complex Hermitian variant (interleaved re/im).  For g_ij = |g| e^{i alpha}
// the unitary R = diag(1, e^{-i alpha}) R_real zeroes the pair:
//   R[:, i] = (c, -s e^{-i alpha}),  R[:, j] = (s, c e^{-i alpha}),
// rows:    X_i <- c X_i - s e^{i alpha} X_j,  X_j <- s X_i + c e^{i alpha} X_j   (R^H X)
// columns: X_i <- c X_i - s e^{-i alpha} X_j, X_j <- s X_i + c e^{-i alpha} X_j  (X R)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void jac_inner_z_kernel(const double2* __restrict__ G, long ldg,
                                                          const int* __restrict__ pairs, double tol_rel,
                                                          double tol_abs, int max_inner,
                                                          double2* __restrict__ Jout, int* __restrict__ rotcount) {
  __shared__ double Gr[PR][LDP], Gi[PR][LDP];
  __shared__ double Jr[PR][LDP], Ji[PR][LDP];
  __shared__ double cs[PR / 2], sn[PR / 2], er[PR / 2], ei[PR / 2];  // e^{i alpha}
  __shared__ int pi_[PR / 2], pj_[PR / 2], fl[PR / 2];
  const int pr = blockIdx.x, tid = threadIdx.x;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  for (int idx = tid; idx < PR * PR; idx += 256) {
    const int r = idx / PR, cc = idx % PR;
    const double2 v = G[pair_index(bi, bj, r) * ldg + pair_index(bi, bj, cc)];
    Gr[r][cc] = v.x;
    Gi[r][cc] = v.y;
    Jr[r][cc] = (r == cc) ? 1.0 : 0.0;
    Ji[r][cc] = 0.0;
  }
  __syncthreads();
  int total = 0;
  const int p = tid >> 3, sub = tid & 7;
  for (int sweep = 0; sweep < max_inner; ++sweep) {
    int nsweep = 0;
    for (int step = 0; step < PR - 1; ++step) {
      int f = 0;
      if (tid < PR / 2) {
        const int pa = tid, pb = PR - 1 - tid;
        const int i = (pa == 0) ? 0 : 1 + ((pa - 1 + step) % (PR - 1));
        const int j = 1 + ((pb - 1 + step) % (PR - 1));
        const double a = Gr[i][i], b = Gr[j][j];
        const double gr = Gr[i][j], gi = Gi[i][j];
        const double ag = sqrt(gr * gr + gi * gi);
        double c = 1.0, s = 0.0, ur = 1.0, ui = 0.0;
        if (ag > fmax(tol_rel * sqrt(fabs(a)) * sqrt(fabs(b)), tol_abs)) {
          const double zeta = (b - a) / (2.0 * ag);
          double t;
          if (fabs(zeta) > 1e150) t = 0.5 / zeta;
          else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          c = 1.0 / sqrt(1.0 + t * t);
          s = c * t;
          ur = gr / ag;
          ui = gi / ag;
          f = 1;
        }
        pi_[tid] = i; pj_[tid] = j; cs[tid] = c; sn[tid] = s; er[tid] = ur; ei[tid] = ui; fl[tid] = f;
      }
      const int nrot = __syncthreads_count(f);
      if (nrot == 0) continue;
      nsweep += nrot;
      if (fl[p]) {  // rows (R^H G): uses e^{i alpha}
        const int i = pi_[p], j = pj_[p];
        const double c = cs[p], s = sn[p], ur = er[p], ui = ei[p];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int col = sub * 8 + u;
          const double xr = Gr[i][col], xi = Gi[i][col], yr = Gr[j][col], yi = Gi[j][col];
          const double wr = ur * yr - ui * yi, wi = ur * yi + ui * yr;  // e^{i alpha} X_j
          Gr[i][col] = c * xr - s * wr;
          Gi[i][col] = c * xi - s * wi;
          Gr[j][col] = s * xr + c * wr;
          Gi[j][col] = s * xi + c * wi;
        }
      }
      __syncthreads();
      if (fl[p]) {  // columns (G R) and J R: uses e^{-i alpha}
        const int i = pi_[p], j = pj_[p];
        const double c = cs[p], s = sn[p], ur = er[p], ui = -ei[p];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int row = sub * 8 + u;
          {
            const double xr = Gr[row][i], xi = Gi[row][i], yr = Gr[row][j], yi = Gi[row][j];
            const double wr = ur * yr - ui * yi, wi = ur * yi + ui * yr;
            Gr[row][i] = c * xr - s * wr;
            Gi[row][i] = c * xi - s * wi;
            Gr[row][j] = s * xr + c * wr;
            Gi[row][j] = s * xi + c * wi;
          }
          {
            const double xr = Jr[row][i], xi = Ji[row][i], yr = Jr[row][j], yi = Ji[row][j];
            const double wr = ur * yr - ui * yi, wi = ur * yi + ui * yr;
            Jr[row][i] = c * xr - s * wr;
            Ji[row][i] = c * xi - s * wi;
            Jr[row][j] = s * xr + c * wr;
            Ji[row][j] = s * xi + c * wi;
          }
        }
      }
      __syncthreads();
    }
    if (nsweep == 0) break;
    ++total;
  }
  double2* J = Jout + (long)pr * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) J[idx] = make_double2(Jr[idx / PR][idx % PR], Ji[idx / PR][idx % PR]);
  if (tid == 0) rotcount[pr] = total;
}

// rows mode: pair rows of A <- J^H rows; COLS: pair columns of A <- columns J.
// Scalar complex FMAs; thread t owns row t/4 and 16 columns of the 64x64 tile.
template <bool COLS>
__global__ __launch_bounds__(256) void jac_apply_z_kernel(double2* __restrict__ A, long L, long ld,
                                                          const int* __restrict__ pairs,
                                                          const double2* __restrict__ Jall,
                                                          const int* __restrict__ rotcount) {
  __shared__ double Xr[PR][LDP], Xi[PR][LDP];  // tile of A (row-mode: [pair k][col]; col-mode: [row][pair k])
  __shared__ double Kr[PR][LDP], Ki[PR][LDP];  // J
  const int pr = blockIdx.x;
  if (rotcount[pr] == 0) return;
  const int bi = pairs[2 * pr], bj = pairs[2 * pr + 1];
  const long t0 = (long)blockIdx.y * PR;
  const int tid = threadIdx.x;
  const double2* J = Jall + (long)pr * PR * PR;
  for (int idx = tid; idx < PR * PR; idx += 256) {
    const int r = idx / PR, cc = idx % PR;
    const double2 jv = J[idx];
    Kr[r][cc] = jv.x;
    Ki[r][cc] = jv.y;
    double2 v = make_double2(0.0, 0.0);
    if (!COLS) {
      const long col = t0 + cc;
      if (col < L) v = A[pair_index(bi, bj, r) * ld + col];
    } else {
      const long row = t0 + r;
      if (row < L) v = A[row * ld + pair_index(bi, bj, cc)];
    }
    Xr[r][cc] = v.x;
    Xi[r][cc] = v.y;
  }
  __syncthreads();
  const int rr = tid >> 2, c0 = (tid & 3) * 16;
  double orr[16], oi[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) { orr[u] = 0.0; oi[u] = 0.0; }
  for (int k = 0; k < PR; ++k) {
    if (!COLS) {
      // out[rr][col] = sum_k conj(J[k][rr]) X[k][col]
      const double jr = Kr[k][rr], ji = -Ki[k][rr];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const double xr = Xr[k][c0 + u], xi = Xi[k][c0 + u];
        orr[u] = fma(jr, xr, fma(-ji, xi, orr[u]));
        oi[u] = fma(jr, xi, fma(ji, xr, oi[u]));
      }
    } else {
      // out[rr][cc] = sum_k X[rr][k] J[k][cc]
      const double xr = Xr[rr][k], xi = Xi[rr][k];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const double jr = Kr[k][c0 + u], ji = Ki[k][c0 + u];
        orr[u] = fma(xr, jr, fma(-xi, ji, orr[u]));
        oi[u] = fma(xr, ji, fma(xi, jr, oi[u]));
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int cc = c0 + u;
    if (!COLS) {
      const long col = t0 + cc;
      if (col < L) A[pair_index(bi, bj, rr) * ld + col] = make_double2(orr[u], oi[u]);
    } else {
      const long row = t0 + rr;
      if (row < L) A[row * ld + pair_index(bi, bj, cc)] = make_double2(orr[u], oi[u]);
    }
  }
}

__global__ void set_identity_z_kernel(double2* __restrict__ Q, long n) {
  const long total = n * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    Q[i] = make_double2((i / n == i % n) ? 1.0 : 0.0, 0.0);
}

__global__ void diag_z_kernel(const double2* __restrict__ G, long ld, long n, double* __restrict__ d) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    d[i] = G[i * ld + i].x;
}

__global__ void set_identity_kernel(double* __restrict__ Q, long n) {
  const long total = n * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    Q[i] = (i / n == i % n) ? 1.0 : 0.0;
}

__global__ void diag_kernel(const double* __restrict__ G, long ld, long n, double* __restrict__ d) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    d[i] = G[i * ld + i];
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

void sym_eig_jacobi(tn_ctx* c, void* G, long p, long ldg, void* Vt, long ldv, bool cplx, std::vector<double>& w,
                    int* sweeps_out) {
  // G (p x p symmetric / Hermitian, leading dimension ldg >= ppad, padding
  // rows/columns zero) is overwritten by its (nearly) diagonal form; Vt
  // (ppad x ppad) receives V^T (real) or V^H (complex): eigenvectors as rows
  // (conjugated in the complex case).
  const long ppad = ((p + PR - 1) / PR) * PR;
  const long nb = ppad / BR;
  const unsigned gd = (unsigned)std::min<long>(1024, (p + 255) / 256);
  {
    const long tot = ppad * ppad;
    const int g = (int)std::min<long>(4096, (tot + 255) / 256);
    TN_REQUIRE(ldv == ppad, TN_E_INVALID, "sym_eig: V leading dimension");
    if (cplx) hipLaunchKernelGGL(set_identity_z_kernel, dim3(g), dim3(256), 0, c->stream, (double2*)Vt, ppad);
    else hipLaunchKernelGGL(set_identity_kernel, dim3(g), dim3(256), 0, c->stream, (double*)Vt, ppad);
    TN_HIP(hipGetLastError());
  }
  double* d = (double*)c->workspace(WS_SVD_S, (size_t)ppad * sizeof(double));
  auto diag = [&]() {
    if (cplx) hipLaunchKernelGGL(diag_z_kernel, dim3(gd), dim3(256), 0, c->stream, (const double2*)G, ldg, p, d);
    else hipLaunchKernelGGL(diag_kernel, dim3(gd), dim3(256), 0, c->stream, (const double*)G, ldg, p, d);
    TN_HIP(hipGetLastError());
  };
  diag();
  std::vector<double> d0(p);
  TN_HIP(hipMemcpyAsync(d0.data(), d, p * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
  double trace = 0.0;
  for (double v : d0) trace += std::fabs(v);
  const Schedule sch = make_schedule(nb);
  const long np = nb / 2, nrounds = nb - 1;
  int* dpairs = (int*)c->workspace(WS_SVD_PAIRS, sch.pairs.size() * sizeof(int));
  TN_HIP(hipMemcpyAsync(dpairs, sch.pairs.data(), sch.pairs.size() * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  void* J = c->workspace(WS_SVD_J, (size_t)np * PR * PR * (cplx ? 16 : 8));
  int* rot = (int*)c->workspace(WS_SVD_CNT, (size_t)nrounds * np * sizeof(int));
  // rotate (i, j) while |g_ij| > max(tol_rel sqrt(g_ii g_jj), tol_abs): the
  // absolute floor eps * trace(G) is the accuracy a density-matrix
  // eigensolver reaches; below it a rotation only moves rounding noise.
  const double eps = 2.220446049250313e-16;
  const double tol_rel = 64.0 * eps, tol_abs = eps * trace;
  const int tiles = (int)(ppad / PR);
  const dim3 ga((unsigned)np, (unsigned)tiles);
  std::vector<int> hrot((size_t)nrounds * np);
  const int max_sweeps = 40;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    for (long r = 0; r < nrounds; ++r) {
      const int* prs = dpairs + r * np * 2;
      int* rc = rot + r * np;
      if (!cplx) {
        double* Gd = (double*)G;
        double* Jd = (double*)J;
        hipLaunchKernelGGL(jac_inner_kernel, dim3((unsigned)np), dim3(256), 0, c->stream, Gd, ldg, prs, tol_rel,
                           tol_abs, 3, Jd, rc);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL((jac_apply_kernel<false>), ga, dim3(256), 0, c->stream, Gd, ppad, ldg, prs, Jd, rc);
        hipLaunchKernelGGL((jac_apply_kernel<true>), ga, dim3(256), 0, c->stream, Gd, ppad, ldg, prs, Jd, rc);
        hipLaunchKernelGGL((jac_apply_kernel<false>), ga, dim3(256), 0, c->stream, (double*)Vt, ppad, ldv, prs, Jd,
                           rc);
      } else {
        double2* Gz = (double2*)G;
        double2* Jz = (double2*)J;
        hipLaunchKernelGGL(jac_inner_z_kernel, dim3((unsigned)np), dim3(256), 0, c->stream, Gz, ldg, prs, tol_rel,
                           tol_abs, 3, Jz, rc);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL((jac_apply_z_kernel<false>), ga, dim3(256), 0, c->stream, Gz, ppad, ldg, prs, Jz, rc);
        hipLaunchKernelGGL((jac_apply_z_kernel<true>), ga, dim3(256), 0, c->stream, Gz, ppad, ldg, prs, Jz, rc);
        hipLaunchKernelGGL((jac_apply_z_kernel<false>), ga, dim3(256), 0, c->stream, (double2*)Vt, ppad, ldv, prs,
                           Jz, rc);
      }
      TN_HIP(hipGetLastError());
    }
    TN_HIP(hipMemcpyAsync(hrot.data(), rot, hrot.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    long tot = 0;
    for (int v : hrot) tot += v;
    if (tot == 0) break;
  }
  if (sweeps_out) *sweeps_out = sweep + 1;
  static const bool dbg = std::getenv("TNHIP_DEBUG_SVD") != nullptr;
  if (dbg) std::fprintf(stderr, "[tnhip eig] %ld%s: %d sweeps, %ld rounds/sweep\n", p, cplx ? " (complex)" : "",
                        sweep + 1, nrounds);
  diag();
  w.resize(p);
  TN_HIP(hipMemcpyAsync(w.data(), d, p * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
}

// eigen-decomposition of the padded Hermitian matrix G (ld ppad): divide and
// conquer from p = 16, two-sided Jacobi below
static void eig_padded(tn_ctx* c, void* G, long p, long ppad, void* Vt, bool cplx, GramEig& out) {
  out.p = p;
  out.ppad = ppad;
  out.cplx = cplx;
  out.Vt = Vt;
  constexpr long dc_min = 16;  // smallest p for the divide-and-conquer path
  if (!cplx && p >= std::max(3L, dc_min)) {
    sym_eig_dc(c, (double*)G, p, ppad, (double*)Vt, ppad, out.w);
    out.sweeps = 0;
  } else if (cplx && p >= std::max(3L, dc_min)) {
    sym_eig_dc_z(c, (double*)G, p, ppad, (double*)Vt, ppad, out.w);
    out.sweeps = 0;
  } else {
    sym_eig_jacobi(c, G, p, ppad, Vt, ppad, cplx, out.w, &out.sweeps);
  }
}

void gram_eig(tn_ctx* c, const void* M, long m, long n, bool left, bool cplx, GramEig& out, const void* X, long kx,
              double xs, long ldx) {
  TN_REQUIRE(m > 0 && n > 0, TN_E_INVALID, "svd: empty matrix");
  const long p = left ? m : n;
  const long ppad = ((p + PR - 1) / PR) * PR;
  const size_t es = cplx ? 16 : 8;
  void* G = c->workspace(WS_SVD_X, (size_t)ppad * ppad * es);
  void* Vt = c->workspace(WS_SVD_Q, (size_t)ppad * ppad * es);
  fill_f64(c, (double*)G, 0.0, ppad * ppad * (cplx ? 2 : 1));
  GemmDesc g;
  g.M = p; g.N = p; g.C = G; g.ldc = ppad;
  if (left) {  // G = M M^T  (M M^H)
    g.K = n; g.A = M; g.lda = n; g.B = M; g.ldb = n; g.tB = 1; g.cB = cplx;
  } else {     // G = M^T M  (M^H M)
    g.K = m; g.A = M; g.lda = n; g.tA = 1; g.cA = cplx; g.B = M; g.ldb = n;
  }
  gemm(c, cplx ? Prec::Z64 : Prec::F64, g);
  if (X && kx > 0 && xs != 0.0) {  // density-matrix perturbation
    GemmDesc h;
    h.M = p; h.N = p; h.K = kx; h.C = G; h.ldc = ppad;
    h.alpha = xs; h.beta = 1.0;
    const long ld = ldx > 0 ? ldx : kx;
    if (left) { h.A = X; h.lda = ld; h.B = X; h.ldb = ld; h.tB = 1; h.cB = cplx; }
    else      { h.A = X; h.lda = p; h.tA = 1; h.cA = cplx; h.B = X; h.ldb = p; }
    gemm(c, cplx ? Prec::Z64 : Prec::F64, h);
  }
  out.left = left;
  eig_padded(c, G, p, ppad, Vt, cplx, out);
}

void herm_eig(tn_ctx* c, const void* H, long p, long ldh, bool cplx, GramEig& out) {
  TN_REQUIRE(p > 0, TN_E_INVALID, "eigen: empty matrix");
  const long ppad = ((p + PR - 1) / PR) * PR;
  const size_t es = cplx ? 16 : 8;
  void* G = c->workspace(WS_SVD_X, (size_t)ppad * ppad * es);
  void* Vt = c->workspace(WS_SVD_Q, (size_t)ppad * ppad * es);
  fill_f64(c, (double*)G, 0.0, ppad * ppad * (cplx ? 2 : 1));
  TN_HIP(hipMemcpy2DAsync(G, (size_t)ppad * es, H, (size_t)ldh * es, (size_t)p * es, (size_t)p,
                          hipMemcpyDeviceToDevice, c->stream));
  out.left = true;
  eig_padded(c, G, p, ppad, Vt, cplx, out);
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
