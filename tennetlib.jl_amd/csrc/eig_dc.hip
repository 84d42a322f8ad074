// Symmetric eigen-decomposition of the density matrix for p > 64, in FP64:
//
//  1. Householder tridiagonalisation G = H T H^T on the device, one column
//     per step: trd_reflect_kernel forms the reflector from row j (every
//     workgroup redundantly, so no extra launch) and y = tau G22 v;
//     trd_update_kernel applies G22 -= v w^T + w v^T.  The reflectors are
//     kept below the diagonal of G.
//  2. Divide and conquer on T (Cuppen's tearing, bottom-up): leaves of at
//     most 32 rows are diagonalised by parallel Jacobi in LDS (one
//     wavefront per leaf); each merge of two neighbouring blocks solves
//     D + rho z z^T: deflation (small |rho z_i|, or a Givens rotation for
//     nearly equal d_i) on the host, the secular equation for every
//     remaining root on the device (one wavefront per root, bracketed Newton
//     in tau = lambda - d_origin measured from the nearer pole), the
//     Gu-Eisenstat z-hat, the eigenvectors z-hat / (d - lambda), and the
//     block update Q <- Q U as an FP64 MFMA GEMM.  The host keeps the
//     eigenvalues and only reads z and the roots once per level.
//  3. Back-transformation Q <- H Q with blocks of 64 reflectors in the
//     compact WY form I - V T V^T: three GEMMs per block and a triangular
//     solve with T^{-1} = diag(1/tau) + striu(V^T V) in place of T.
// Accuracy: eigenvalues to ~eps * ||G||, orthonormal eigenvectors; the
// same absolute accuracy as the block-Jacobi solver used for p <= 64.
#include "tn_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <numeric>

namespace tn {
namespace {

constexpr double EPS = 2.220446049250313e-16;
constexpr int HR = 4;     // rows of y per workgroup in the reflector step (one per wave)
constexpr int LEAF = 32;  // leaf size of the divide and conquer
constexpr int NBW = 64;   // reflectors per compact-WY block

// sum over the 256 threads of a workgroup; every thread gets the same value
__device__ double block_sum(double v, double* red) {
  v = wave_sum64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// two block sums with one pair of barriers; each result has the bits block_sum
// would give
__device__ void block_sum2(double u, double v, double* red, double& su, double& sv) {
  u = wave_sum64(u);
  v = wave_sum64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wave] = u;
    red[4 + wave] = v;
  }
  __syncthreads();
  su = (red[0] + red[1]) + (red[2] + red[3]);
  sv = (red[4] + red[5]) + (red[6] + red[7]);
}

// x times the value of the lane that the DPP control CTRL selects
template <int CTRL>
__device__ __forceinline__ double dpp_mul(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return x * __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// product over the 64 lanes of a wavefront, the same pattern as wave_sum64
// (every lane gets the same value; every lane must be active)
__device__ __forceinline__ double wave_prod64(double x) {
  x = dpp_mul<0xB1>(x);
  x = dpp_mul<0x4E>(x);
  x = dpp_mul<0x141>(x);
  x = dpp_mul<0x140>(x);
  return (read_lane(x, 0) * read_lane(x, 16)) * (read_lane(x, 32) * read_lane(x, 48));
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation
// ---------------------------------------------------------------------------
// Step j (trailing block G22 = G[j+1:, j+1:], m = p - j - 1 rows).  Grid:
// ceil(m / HR) workgroups, dynamic LDS m doubles.
__global__ __launch_bounds__(256) void trd_reflect_kernel(double* __restrict__ A, long ld, int p, int j,
                                                          double* __restrict__ y, double* __restrict__ part,
                                                          double* __restrict__ dv, double* __restrict__ ev,
                                                          double* __restrict__ tauv, double* __restrict__ vcont) {
  extern __shared__ double vs[];
  __shared__ double red[4];
  const int m = p - j - 1, tid = threadIdx.x;
  const double* x = A + (long)j * ld + (j + 1);  // row j = column j (symmetric)
  double ss = 0.0;
  for (int c = tid; c < m; c += 256) {
    const double t = x[c];
    vs[c] = t;
    if (c > 0) ss += t * t;
  }
  ss = block_sum(ss, red);
  const double x0 = vs[0];
  double tau = 0.0, alpha = x0, scale = 0.0;
  if (ss > 0.0) {
    const double nrm = sqrt(x0 * x0 + ss);
    alpha = (x0 >= 0.0) ? -nrm : nrm;
    tau = (alpha - x0) / alpha;
    scale = 1.0 / (x0 - alpha);
  }
  __syncthreads();
  for (int c = tid; c < m; c += 256) vs[c] = (c == 0) ? 1.0 : vs[c] * scale;
  __syncthreads();
  if (blockIdx.x == 0) {
    // v below the diagonal of column j (never read again as matrix data)
    for (int c = tid; c < m; c += 256) {
      A[(long)(j + 1 + c) * ld + j] = vs[c];
      if (vcont) vcont[c] = vs[c];
    }
    if (tid == 0) {
      dv[j] = A[(long)j * ld + j];
      ev[j] = alpha;
      tauv[j] = tau;
    }
  }
  if (tau == 0.0) {  // uniform over the workgroup
    if (tid == 0) part[blockIdx.x] = 0.0;
    return;
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int r0 = blockIdx.x * HR;
  double dot = 0.0;
  for (int rr = wave; rr < HR; rr += 4) {
    const int r = r0 + rr;
    if (r >= m) break;
    const double* row = A + (long)(j + 1 + r) * ld + (j + 1);
    double s = 0.0;
    for (int c = lane; c < m; c += 64) s += row[c] * vs[c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) {
      const double yr = tau * s;
      y[r] = yr;
      dot += yr * vs[r];
    }
  }
  dot = block_sum(dot, red);
  if (tid == 0) part[blockIdx.x] = dot;
}

// G22 -= v w^T + w v^T with w = y - (tau/2)(y^T v) v; 64 x 64 tiles.
__global__ __launch_bounds__(256) void trd_update_kernel(double* __restrict__ A, long ld, int p, int j,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ part, int nparts,
                                                         const double* __restrict__ tauv) {
  const double tau = tauv[j];
  if (tau == 0.0) return;
  __shared__ double red[4];
  __shared__ double vr[64], wr[64], vc[64], wc[64];
  const int m = p - j - 1, tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < nparts; i += 256) s += part[i];
  s = block_sum(s, red);
  const double hs = 0.5 * tau * s;
  const int tr0 = blockIdx.y * 64, tc0 = blockIdx.x * 64;
  if (tid < 64) {
    const int r = tr0 + tid;
    if (r < m) {
      const double v = A[(long)(j + 1 + r) * ld + j];
      vr[tid] = v;
      wr[tid] = y[r] - hs * v;
    } else {
      vr[tid] = 0.0;
      wr[tid] = 0.0;
    }
  } else if (tid < 128) {
    const int c = tc0 + tid - 64;
    if (c < m) {
      const double v = A[(long)(j + 1 + c) * ld + j];
      vc[tid - 64] = v;
      wc[tid - 64] = y[c] - hs * v;
    } else {
      vc[tid - 64] = 0.0;
      wc[tid - 64] = 0.0;
    }
  }
  __syncthreads();
  const int cc = tid & 63;
  const int c = tc0 + cc;
  if (c >= m) return;
  for (int rr = tid >> 6; rr < 64; rr += 4) {
    const int r = tr0 + rr;
    if (r >= m) break;
    double* a = A + (long)(j + 1 + r) * ld + (j + 1 + c);
    *a -= vr[rr] * wc[cc] + wr[rr] * vc[cc];
  }
}

// ---- fused step: one launch per column.  Launch j applies the rank-2
// update of step j to the rows it owns (rows >= j + 2; row j + 1 is never
// written again) and, at the same time, forms reflector j + 1 and
// y_{j+1} = tau_{j+1} G22^{(j+1)} v_{j+1} from the old rows:
//   new row j+1 = old row j+1 - v_j[j+1] w_j - w_j[j+1] v_j   (every workgroup)
//   y_{j+1}[r] = tau (old_r . v' - v_j[r] (w_j . v') - w_j[r] (v_j . v'))
// so each row is read once and written once per column, by one workgroup.
// Vectors are indexed from row j + 1 (v_j, w_j) and from row j + 2 (v').
// Dynamic LDS: 3 (p - j - 1) doubles.
__global__ __launch_bounds__(256) void trd_step_kernel(double* __restrict__ A, long ld, int p, int j,
                                                       const double* __restrict__ vj_in,
                                                       const double* __restrict__ yj,
                                                       const double* __restrict__ partj, int npj,
                                                       double* __restrict__ vn_out, double* __restrict__ yn,
                                                       double* __restrict__ partn, double* __restrict__ dv,
                                                       double* __restrict__ ev, double* __restrict__ tauv) {
  extern __shared__ double lds[];
  __shared__ double red[4];
  __shared__ double red2[8];
  const int mj = p - j - 1, mn = mj - 1;
  double* vj = lds;
  double* wj = vj + mj;
  double* vn = wj + mj;
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // issue the loads of row j+1 and of this wave's first own row up front, so
  // their latency overlaps the vector set-up (rows up to 6144 / 1024 long)
  constexpr int OV = 24, RV = 16;
  const double* orow = A + (long)(j + 1) * ld + (j + 2);
  double ov[OV];
#pragma unroll
  for (int u = 0; u < OV; ++u) {
    const int t = tid + 256 * u;
    ov[u] = (t < mn) ? orow[t] : 0.0;
  }
  const int rr0 = g * 4 + wave;
  const bool pre = (mn <= 64 * RV) && (rr0 < mn);
  double rv[RV];
  {
    const double* row0 = A + (long)(j + 2 + (pre ? rr0 : 0)) * ld + (j + 2);
#pragma unroll
    for (int u = 0; u < RV; ++u) {
      const int t = lane + 64 * u;
      rv[u] = (pre && t < mn) ? row0[t] : 0.0;
    }
  }
  // v_j and y_j (up to 1024 entries) are loaded before the reduction of the
  // partial sums, so both global round trips overlap
  constexpr int PV = 4;
  double pv[PV], py[PV];
#pragma unroll
  for (int u = 0; u < PV; ++u) {
    const int c = tid + 256 * u;
    pv[u] = (c < mj) ? vj_in[c] : 0.0;
    py[u] = (c < mj) ? yj[c] : 0.0;
  }
  const double tj = tauv[j];
  double s = 0.0;
  if (tj != 0.0)
    for (int i = tid; i < npj; i += 256) s += partj[i];
  s = block_sum(s, red);
  const double hs = 0.5 * tj * s;
#pragma unroll
  for (int u = 0; u < PV; ++u) {
    const int c = tid + 256 * u;
    if (c < mj) {
      vj[c] = pv[u];
      wj[c] = (tj != 0.0) ? py[u] - hs * pv[u] : 0.0;
    }
  }
  for (int c = tid + 256 * PV; c < mj; c += 256) {
    const double v = vj_in[c];
    vj[c] = v;
    wj[c] = (tj != 0.0) ? yj[c] - hs * v : 0.0;
  }
  __syncthreads();
  // new row j+1, columns j+2.. (index t), and its reflector
  const double v0 = vj[0], w0 = wj[0];
  double ss = 0.0;
#pragma unroll
  for (int u = 0; u < OV; ++u) {
    const int t = tid + 256 * u;
    if (t < mn) {
      const double x = ov[u] - v0 * wj[t + 1] - w0 * vj[t + 1];
      vn[t] = x;
      if (t > 0) ss += x * x;
    }
  }
  ss = block_sum(ss, red);
  const double x0 = vn[0];
  double tau = 0.0, alpha = x0, scale = 0.0;
  if (ss > 0.0) {
    const double nrm = sqrt(x0 * x0 + ss);
    alpha = (x0 >= 0.0) ? -nrm : nrm;
    tau = (alpha - x0) / alpha;
    scale = 1.0 / (x0 - alpha);
  }
  __syncthreads();
  double pa = 0.0, pb = 0.0;
  for (int t = tid; t < mn; t += 256) {
    const double v = (t == 0) ? 1.0 : vn[t] * scale;
    vn[t] = v;
    pa += wj[t + 1] * v;
    pb += vj[t + 1] * v;
  }
  double a, b;
  block_sum2(pa, pb, red2, a, b);
  if (g == 0) {
    for (int t = tid; t < mn; t += 256) {
      A[(long)(j + 2 + t) * ld + (j + 1)] = vn[t];
      vn_out[t] = vn[t];
    }
    if (tid == 0) {
      dv[j + 1] = A[(long)(j + 1) * ld + (j + 1)] - 2.0 * v0 * w0;
      ev[j + 1] = alpha;
      tauv[j + 1] = tau;
    }
  }
  // own rows: update of step j and the rows of y_{j+1}
  double dot = 0.0;
  for (int rr = rr0; rr < mn; rr += G * 4) {
    double* row = A + (long)(j + 2 + rr) * ld + (j + 2);
    const double vr = vj[rr + 1], wr = wj[rr + 1];
    double acc = 0.0;
    if (pre && rr == rr0) {
#pragma unroll
      for (int u = 0; u < RV; ++u) {
        const int t = lane + 64 * u;
        if (t < mn) {
          const double old = rv[u];
          acc += old * vn[t];
          if (tj != 0.0) row[t] = old - vr * wj[t + 1] - wr * vj[t + 1];
        }
      }
    } else {
      for (int t = lane; t < mn; t += 64) {
        const double old = row[t];
        acc += old * vn[t];
        if (tj != 0.0) row[t] = old - vr * wj[t + 1] - wr * vj[t + 1];
      }
    }
    acc = wave_sum64(acc);
    if (lane == 0) {
      const double yv = tau * (acc - vr * a - wr * b);
      yn[rr] = yv;
      dot += yv * vn[rr];
    }
  }
  dot = block_sum(dot, red);
  if (tid == 0) partn[g] = dot;
}

// ---- blocked reduction for large p (LAPACK dlatrd panels of NBT columns).
// Within a panel starting at column j0 the trailing block is left as it was
// at j0; column j's row and its matrix-vector product are corrected with the
// panel's earlier reflectors v_i and vectors w_i (i = 0..k-1, k = j - j0):
//   row_j   = A[j, j+1:] - sum_i (v_i[j] w_i + w_i[j] v_i)
//   y_j     = tau_j (A22 v_j - sum_i (v_i (w_i^T v_j) + w_i (v_i^T v_j)))
//   w_j     = y_j - (tau_j / 2)(y_j^T v_j) v_j
// and the panel ends with one MFMA GEMM, A[j1:, j1:] -= V W^T + W V^T.  Each
// column reads the trailing block once (the product), not twice (product and
// rank-2 update).  Vectors are stored by global row in the panel buffers
// P = [V; W] and Q = [W; V] (2 NBT rows of length p; unused rows zero).
constexpr int NBT = 32;
constexpr int RCHUNK = 256;  // row entries per workgroup in trd_panel_row_kernel (one per thread)
// dynamic LDS of trd_panel_row_kernel: four waves' products, (2 NBT + 1) rows of 65
constexpr size_t TRD_ROW_LDS = (size_t)4 * (2 * NBT + 1) * 65 * sizeof(double);

// Column j: forms w_{k-1} (the previous column's, from y_{j-1} and the
// partial sums of y_{j-1}^T v_{j-1}), then the current row x (stored in xbuf,
// entries c = 0..m-1 for columns j+1+c) with partial sums of sum_{c>=1} x_c^2,
// v_i . x and w_i . x over c >= 1 (i < k); d_j from workgroup 0.
// One entry per thread: its panel entries are loaded first, and the values
// every thread shares (y_{j-1}^T v_{j-1}, v_i[j], w_i[j]) are read by each
// wave itself, so nothing waits on a barrier before the row is formed.
// Grid: ceil(m / RCHUNK) workgroups.
__global__ __launch_bounds__(256) void trd_panel_row_kernel(const double* __restrict__ A, long ld, int p, int j, int k,
                                                            double* __restrict__ P, double* __restrict__ Q,
                                                            const double* __restrict__ yprev,
                                                            const double* __restrict__ partB, int nB,
                                                            const double* __restrict__ tauv,
                                                            double* __restrict__ xbuf, double* __restrict__ partA,
                                                            double* __restrict__ dv) {
  __shared__ double accs[4][2 * NBT + 1];
  extern __shared__ double tr[];  // per wave (2 NBT + 1) x 65: the products, transposed for the sums
  const int m = p - j - 1, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = blockIdx.x * RCHUNK + tid;
  const bool act = c < m;
  const long r = j + 1 + c;  // global row / column index
  // this entry's loads
  double V[NBT], W[NBT];
#pragma unroll
  for (int i = 0; i < NBT; ++i) {
    V[i] = (act && i < k) ? P[(long)i * p + r] : 0.0;
    W[i] = (act && i < k - 1) ? P[(long)(NBT + i) * p + r] : 0.0;
  }
  const double a = act ? A[r * ld + j] : 0.0;  // column j below the diagonal (the lower triangle is current)
  const double yp = (act && k >= 1) ? yprev[c + 1] : 0.0;
  // hs = (tau_{j-1} / 2) y_{j-1}^T v_{j-1}, reduced by every wave alike
  double hs = 0.0;
  if (k >= 1) {
    double s = 0.0;
    for (int i = lane; i < nB; i += 64) s += partB[i];
    hs = 0.5 * tauv[j - 1] * wave_sum64(s);
  }
  // w_{k-1} = y_{j-1} - hs v_{k-1} over rows >= j
  double wj = 0.0;
  if (k >= 1) {
    wj = yprev[0] - hs * P[(long)(k - 1) * p + j];
    double vk1 = 0.0;  // V[k - 1], selected without a runtime index (keeps V in registers)
#pragma unroll
    for (int i = 0; i < NBT; ++i)
      if (i == k - 1) vk1 = V[i];
    const double wn = yp - hs * vk1;
#pragma unroll
    for (int i = 0; i < NBT; ++i)
      if (i == k - 1) W[i] = wn;
    if (act) {
      P[(long)(NBT + k - 1) * p + r] = wn;
      Q[(long)(k - 1) * p + r] = wn;
    }
    if (blockIdx.x == 0 && tid == 0) {
      P[(long)(NBT + k - 1) * p + j] = wj;
      Q[(long)(k - 1) * p + j] = wj;
    }
  }
  // x = A[j, r] - sum_i (v_i[j] w_i[r] + w_i[j] v_i[r])
  double x = a;
#pragma unroll
  for (int i = 0; i < NBT; ++i) {
    if (i < k) {
      const double svi = P[(long)i * p + j];
      const double swi = (i == k - 1) ? wj : P[(long)(NBT + i) * p + j];
      x -= svi * W[i] + swi * V[i];
    }
  }
  if (act) xbuf[c] = x;
  const double xs = (act && c >= 1) ? x : 0.0;
  const int nq = 1 + 2 * NBT;
  // the wave's 64 x nq products go to LDS as rows q of 65 (padded), and lane
  // q sums row q (lane 0 also row 64): one LDS pass instead of nq cross-lane
  // reductions
  double* T = tr + (long)wave * nq * 65;
#pragma unroll
  for (int q = 0; q < nq; ++q) {
    if (q == 0 || (q - 1) % NBT < k) {
      const double f = q == 0 ? xs * xs : (q <= NBT ? V[q - 1] * xs : W[q - 1 - NBT] * xs);
      T[q * 65 + lane] = f;
    }
  }
  __syncthreads();
  for (int q = lane; q < nq; q += 64) {
    if (q == 0 || (q - 1) % NBT < k) {
      const double* row = T + q * 65;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int t = 0; t < 64; t += 4) {
        s0 += row[t];
        s1 += row[t + 1];
        s2 += row[t + 2];
        s3 += row[t + 3];
      }
      accs[wave][q] = (s0 + s1) + (s2 + s3);
    }
  }
  __syncthreads();
  for (int q = tid; q < nq; q += 256)
    if (q == 0 || (q - 1) % NBT < k)
      partA[(long)blockIdx.x * nq + q] = (accs[0][q] + accs[1][q]) + (accs[2][q] + accs[3][q]);
  if (blockIdx.x == 0 && tid == 0) {
    double dj = A[(long)j * ld + j];
    for (int i = 0; i < k; ++i) {
      const double swi = (i == k - 1) ? wj : P[(long)(NBT + i) * p + j];
      dj -= 2.0 * P[(long)i * p + j] * swi;
    }
    dv[j] = dj;
  }
}

// Column j: the reflector from the partial sums and xbuf (every workgroup),
// y_j for this workgroup's rows from one pass over the stored trailing
// block, partial sums of y_j . v_j; workgroup 0 stores v_j (below the
// diagonal of column j and in P, Q), e_j and tau_j.  Dynamic LDS: m doubles.
__global__ __launch_bounds__(256) void trd_panel_symv_kernel(double* __restrict__ A, long ld, int p, int j, int k,
                                                             double* __restrict__ P, double* __restrict__ Q,
                                                             const double* __restrict__ xbuf,
                                                             const double* __restrict__ partA, int nA,
                                                             double* __restrict__ y, double* __restrict__ partB,
                                                             double* __restrict__ ev, double* __restrict__ tauv) {
  extern __shared__ double vs[];
  __shared__ double red[4];
  __shared__ double sums[2 * NBT + 1];
  __shared__ double cvw[2 * NBT];  // (v_i . v_j), (w_i . v_j)
  const int m = p - j - 1, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int G = gridDim.x, g = blockIdx.x;
  const int nq = 1 + 2 * NBT;
  for (int q = tid; q < nq; q += 256) {
    double s = 0.0;
    if (q == 0 || (q - 1) % NBT < k)
      for (int b = 0; b < nA; ++b) s += partA[(long)b * nq + q];
    sums[q] = s;
  }
  __syncthreads();
  const double x0 = xbuf[0], ss = sums[0];
  double tau = 0.0, alpha = x0, scale = 0.0;
  if (ss > 0.0) {
    const double nrm = sqrt(x0 * x0 + ss);
    alpha = (x0 >= 0.0) ? -nrm : nrm;
    tau = (alpha - x0) / alpha;
    scale = 1.0 / (x0 - alpha);
  }
  for (int c = tid; c < m; c += 256) vs[c] = (c == 0) ? 1.0 : xbuf[c] * scale;
  if (tid < k) {
    const long r1 = j + 1;
    cvw[tid] = sums[1 + tid] * scale + P[(long)tid * p + r1];                    // v_i . v_j
    cvw[NBT + tid] = sums[1 + NBT + tid] * scale + P[(long)(NBT + tid) * p + r1];  // w_i . v_j
  }
  __syncthreads();
  if (g == 0) {
    for (int c = tid; c < m; c += 256) {
      const long r = j + 1 + c;
      A[r * ld + j] = vs[c];
      P[(long)k * p + r] = vs[c];
      Q[(long)(NBT + k) * p + r] = vs[c];
    }
    if (tid == 0) {
      ev[j] = alpha;
      tauv[j] = tau;
    }
  }
  // two rows per wave, 16-byte loads (a leading odd element is taken alone)
  double dot = 0.0;
  for (int rr0 = (g * 4 + wave) * 2; rr0 < m; rr0 += G * 8) {
    const bool two = rr0 + 1 < m;
    const long r0 = j + 1 + rr0;
    const double* row0 = A + r0 * ld + (j + 1);
    const double* row1 = two ? row0 + ld : row0;
    double s0 = 0.0, s1 = 0.0;
    const bool pair1 = two && (((uintptr_t)row1 & 15) == ((uintptr_t)row0 & 15));
    int c0 = 0;
    if (((uintptr_t)row0 & 15) != 0) {
      if (lane == 0) {
        s0 = row0[0] * vs[0];
        if (pair1) s1 = row1[0] * vs[0];
      }
      c0 = 1;
    }
    const int nv = (m - c0) >> 1;
    const double2* a0 = reinterpret_cast<const double2*>(row0 + c0);
    const double2* a1 = reinterpret_cast<const double2*>(row1 + c0);
#pragma unroll 4
    for (int t = lane; t < nv; t += 64) {
      const double2 x0 = a0[t];
      const double va = vs[c0 + 2 * t], vb = vs[c0 + 2 * t + 1];
      s0 += x0.x * va + x0.y * vb;
      if (pair1) {
        const double2 x1 = a1[t];
        s1 += x1.x * va + x1.y * vb;
      }
    }
    if (two && !pair1)  // ld odd: the second row is not 16-byte aligned with the first
      for (int c = lane; c < m; c += 64) s1 += row1[c] * vs[c];
    if (((m - c0) & 1) && lane == 0) {
      const int c = m - 1;
      s0 += row0[c] * vs[c];
      if (pair1) s1 += row1[c] * vs[c];
    }
    s0 = wave_sum64(s0);
    s1 = wave_sum64(s1);
    // panel corrections sum_i (v_i[r] (w_i . v) + w_i[r] (v_i . v)), one term per lane
    double q0 = 0.0, q1 = 0.0;
    if (lane < k) {
      const double* vi = P + (long)lane * p;
      const double* wi = P + (long)(NBT + lane) * p;
      q0 = vi[r0] * cvw[NBT + lane] + wi[r0] * cvw[lane];
      if (two) q1 = vi[r0 + 1] * cvw[NBT + lane] + wi[r0 + 1] * cvw[lane];
    }
    q0 = wave_sum64(q0);
    q1 = wave_sum64(q1);
    if (lane == 0) {
      const double y0 = tau * (s0 - q0);
      y[rr0] = y0;
      dot += y0 * vs[rr0];
      if (two) {
        const double y1 = tau * (s1 - q1);
        y[rr0 + 1] = y1;
        dot += y1 * vs[rr0 + 1];
      }
    }
  }
  dot = block_sum(dot, red);
  if (tid == 0) partB[g] = dot;
}

// w of the panel's last column j (rows >= j + 1), index k in the panel.
__global__ __launch_bounds__(256) void trd_panel_wlast_kernel(int p, int j, int k, double* __restrict__ P,
                                                              double* __restrict__ Q, const double* __restrict__ y,
                                                              const double* __restrict__ partB, int nB,
                                                              const double* __restrict__ tauv) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nB; i += 256) s += partB[i];
  s = block_sum(s, red);
  const double hs = 0.5 * tauv[j] * s;
  const int m = p - j - 1;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < m; c += gridDim.x * 256) {
    const long r = j + 1 + c;
    const double wn = y[c] - hs * P[(long)k * p + r];
    P[(long)(NBT + k) * p + r] = wn;
    Q[(long)k * p + r] = wn;
  }
}

// ---- the same column step for large trailing blocks, reading only the lower
// triangle of the symmetric trailing block (half the bytes of a full pass).
// The block is cut into SYT x SYT tiles on the global grid (multiples of SYT,
// so 16-byte row loads stay aligned); a lower tile (I > J) contributes
// A_IJ v_J to the rows of block I and A_IJ^T v_I to the rows of block J, a
// diagonal tile (read whole) A_II v_I.  Each contribution goes to its own slot
// S[b][t] (row block b, partner block t), and the reduction sums the slots in
// a fixed order, so the result does not depend on scheduling.
constexpr int SYT = 64;

// tau, alpha and the reflector scale of column j from the sum of squares ss
// of x_1.. and x_0 (every kernel that needs them computes them this way)
__device__ __forceinline__ void trd_reflector(double x0, double ss, double& tau, double& alpha, double& scale) {
  tau = 0.0;
  alpha = x0;
  scale = 0.0;
  if (ss > 0.0) {
    const double nrm = sqrt(x0 * x0 + ss);
    alpha = (x0 >= 0.0) ? -nrm : nrm;
    tau = (alpha - x0) / alpha;
    scale = 1.0 / (x0 - alpha);
  }
}

// sum of squares of x_1.. from the row kernel's partial sums, reduced by the
// calling wave (every lane gets the same bits; both kernels below use it, so
// they agree on tau and the reflector scale)
__device__ __forceinline__ double trd_ss(const double* __restrict__ partA, int nA) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = lane; b < nA; b += 64) s += partA[(long)b * (1 + 2 * NBT)];
  return wave_sum64(s);
}

// Grid: nt (nt + 1) / 2 workgroups, one per lower tile (I >= J) of the tile
// range [a0, a0 + nt) that covers rows and columns j + 1 .. p - 1.
// S: nt x nt slots of SYT doubles.  Every wave issues its tile loads first and
// forms the reflector entries it needs itself (no barrier before the product).
__global__ __launch_bounds__(256) void trd_symv_tile_kernel(const double* __restrict__ A, long ld, int p, int j,
                                                            const double* __restrict__ xbuf,
                                                            const double* __restrict__ partA, int nA, int a0, int nt,
                                                            double* __restrict__ S) {
  __shared__ double rs[SYT];
  __shared__ double cs[8][SYT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  int I = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= b) ++I;
  while (I * (I + 1) / 2 > b) --I;
  const int J = b - I * (I + 1) / 2;
  const long r1 = j + 1;  // first row and column of the trailing block
  const long R0 = (long)(a0 + I) * SYT, C0 = (long)(a0 + J) * SYT;
  // half-wave h of wave `wave` takes rows wave * 16 + 2 it + h; lane hl the
  // column pair 2 hl, 2 hl + 1.  Rows go in batches of four: the batch's
  // loads are issued before the previous batch is reduced.
  const int hl = lane & 31, h = lane >> 5;
  const long c = C0 + 2 * hl;
  const bool ok0 = c >= r1 && c < p, ok1 = c + 1 >= r1 && c + 1 < p;
  // the tile's columns all inside the block and its rows 16-byte aligned:
  // one vector load per lane and row
  const bool tvec = C0 >= r1 && C0 + SYT <= p && (ld & 1) == 0 && ((uintptr_t)A & 15) == 0;
  const long rb = R0 + wave * 16 + h;  // row of it = 0
  const double* base = A + rb * ld + c;
  constexpr int NIT = SYT / 8, NB = 4;
  double2 a[NB];
  double xr[NB];
  auto load = [&](int it0) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const long r = rb + 2 * (it0 + u);
      a[u] = make_double2(0.0, 0.0);
      xr[u] = 0.0;
      if (r >= r1 && r < p) {
        const double* q = base + (long)(2 * (it0 + u)) * ld;
        if (tvec) {
          a[u] = *reinterpret_cast<const double2*>(q);
        } else {
          if (ok0) a[u].x = q[0];
          if (ok1) a[u].y = q[1];
        }
        if (r > r1) xr[u] = xbuf[r - r1];
      }
    }
  };
  load(0);
  const double xc0 = (ok0 && c > r1) ? xbuf[c - r1] : 0.0;
  const double xc1 = (ok1 && c + 1 > r1) ? xbuf[c + 1 - r1] : 0.0;
  double tau, alpha, scale;
  trd_reflector(xbuf[0], trd_ss(partA, nA), tau, alpha, scale);
  // v at global index g: 1 at r1, x scaled above it, 0 outside the block
  const double v0 = !ok0 ? 0.0 : (c == r1 ? 1.0 : xc0 * scale);
  const double v1 = !ok1 ? 0.0 : (c + 1 == r1 ? 1.0 : xc1 * scale);
  double cs0 = 0.0, cs1 = 0.0;
#pragma unroll
  for (int it0 = 0; it0 < NIT; it0 += NB) {
    double2 b[NB];
    double xb[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      b[u] = a[u];
      xb[u] = xr[u];
    }
    if (it0 + NB < NIT) load(it0 + NB);
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int rl = wave * 16 + 2 * (it0 + u) + h;
      const long r = R0 + rl;
      double s = b[u].x * v0 + b[u].y * v1;
      s += __shfl_xor(s, 16);
      s += __shfl_xor(s, 8);
      s += __shfl_xor(s, 4);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 1);
      if (hl == 0) rs[rl] = s;
      const double w = (r < r1 || r >= p) ? 0.0 : (r == r1 ? 1.0 : xb[u] * scale);
      cs0 += b[u].x * w;
      cs1 += b[u].y * w;
    }
  }
  cs[wave * 2 + h][2 * hl] = cs0;
  cs[wave * 2 + h][2 * hl + 1] = cs1;
  __syncthreads();
  if (tid < SYT) {
    S[((long)I * nt + J) * SYT + tid] = rs[tid];
  } else if (I != J && tid < 2 * SYT) {
    const int cl = tid - SYT;
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += cs[q][cl];
    S[((long)J * nt + I) * SYT + cl] = t;
  }
}

// Rows of column j's product from the tile slots, the panel corrections, y_j
// and the partial sums of y_j . v_j (one per workgroup); every workgroup
// stores v_j for its rows (below the diagonal of column j and in P, Q),
// workgroup 0 e_j and tau_j.  Eight lanes per row, each summing every eighth
// slot and correction term, then combined in a fixed order.
// Grid: ceil(m / 32) workgroups; nA <= 64.
__global__ __launch_bounds__(256) void trd_symv_reduce_kernel(double* __restrict__ A, long ld, int p, int j, int k,
                                                              double* __restrict__ P, double* __restrict__ Q,
                                                              const double* __restrict__ xbuf,
                                                              const double* __restrict__ partA, int nA, int a0,
                                                              int nt, const double* __restrict__ S,
                                                              double* __restrict__ y, double* __restrict__ partB,
                                                              double* __restrict__ ev, double* __restrict__ tauv) {
  __shared__ double red[4];
  __shared__ double sums[2 * NBT + 1];
  __shared__ double cvw[2 * NBT];  // (v_i . v_j), (w_i . v_j)
  const int m = p - j - 1, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nq = 1 + 2 * NBT;
  const int g = tid & 7;
  const int c = blockIdx.x * 32 + (tid >> 3);
  const long r = j + 1 + c;
  // loads that need no reduction result go first: this row's tile slots (the
  // first 128 partners), its panel entries and x
  constexpr int NPQ = NBT / 8;
  double u[16], pv[NPQ], pw[NPQ], xc = 0.0;
  const double* Sb = S + (r / SYT - a0) * nt * SYT + (r % SYT);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int t = g + 8 * q;
    u[q] = (c < m && t < nt) ? Sb[(long)t * SYT] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < NPQ; ++q) {
    const int i = g + 8 * q;
    pv[q] = (c < m && i < k) ? P[(long)i * p + r] : 0.0;
    pw[q] = (c < m && i < k) ? P[(long)(NBT + i) * p + r] : 0.0;
  }
  if (c < m && c > 0) xc = xbuf[c];
  // sums of (v_i . x) and (w_i . x) over the row kernel's workgroups: wave w
  // reduces q = 1 + w, 5 + w, ... (one lane per workgroup partial)
  {
    double vals[(2 * NBT + 3) / 4];
#pragma unroll
    for (int uu = 0; uu < (2 * NBT + 3) / 4; ++uu) {
      const int q = 1 + wave + 4 * uu;
      vals[uu] = (lane < nA && q < nq && (q - 1) % NBT < k) ? partA[(long)lane * nq + q] : 0.0;
    }
#pragma unroll
    for (int uu = 0; uu < (2 * NBT + 3) / 4; ++uu) {
      const double t = wave_sum64(vals[uu]);
      const int q = 1 + wave + 4 * uu;
      if (lane == 0 && q < nq) sums[q] = t;
    }
  }
  double tau, alpha, scale;
  trd_reflector(xbuf[0], trd_ss(partA, nA), tau, alpha, scale);
  __syncthreads();
  if (tid < k) {
    const long r1 = j + 1;
    cvw[tid] = sums[1 + tid] * scale + P[(long)tid * p + r1];
    cvw[NBT + tid] = sums[1 + NBT + tid] * scale + P[(long)(NBT + tid) * p + r1];
  }
  if (blockIdx.x == 0 && tid == 0) {
    ev[j] = alpha;
    tauv[j] = tau;
  }
  __syncthreads();
  double s = 0.0, qv = 0.0, vcur = 0.0;
  if (c < m) {
    vcur = (c == 0) ? 1.0 : xc * scale;
#pragma unroll
    for (int q = 0; q < 16; ++q) s += u[q];
    for (int t0 = 128; t0 < nt; t0 += 128) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int t = t0 + g + 8 * q;
        u[q] = t < nt ? Sb[(long)t * SYT] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) s += u[q];
    }
#pragma unroll
    for (int q = 0; q < NPQ; ++q) {
      const int i = g + 8 * q;
      if (i < k) qv += pv[q] * cvw[NBT + i] + pw[q] * cvw[i];
    }
  }
  s += __shfl_xor(s, 4);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 1);
  qv += __shfl_xor(qv, 4);
  qv += __shfl_xor(qv, 2);
  qv += __shfl_xor(qv, 1);
  double dot = 0.0;
  if (c < m && g == 0) {
    A[r * ld + j] = vcur;
    P[(long)k * p + r] = vcur;
    Q[(long)(NBT + k) * p + r] = vcur;
    const double yv = tau * (s - qv);
    y[c] = yv;
    dot = yv * vcur;
  }
  dot = block_sum(dot, red);
  if (tid == 0) partB[blockIdx.x] = dot;
}

// ---- the panel's rank-64 update of the trailing block, lower triangle only:
// A[r, c] -= sum_i (v_i[r] w_i[c] + w_i[r] v_i[c]) for the 64 x 64 tiles on
// the global grid with I >= J (diagonal tiles whole), rows and columns from
// j1 on.  P = [V; W] and Q = [W; V] (2 NBT rows of length p), so the sum is
// P^T Q over K = 2 NBT.  One workgroup per tile: both K x 64 panels in LDS,
// FP64 MFMA 16x16x4 (each wave a 32 x 32 quarter), then one read-modify-write
// of the tile.  Grid: nt (nt + 1) / 2 tiles of the range [a0, a0 + nt).
typedef double trd_d4 __attribute__((ext_vector_type(4)));
constexpr int SYK = 2 * NBT;
__global__ __launch_bounds__(256) void trd_syr2k_lower_kernel(double* __restrict__ A, long ld, int p, int j1,
                                                              const double* __restrict__ P,
                                                              const double* __restrict__ Q, int a0) {
  constexpr int KH = SYK / 2;  // the K range in two halves (33 KB of LDS, four workgroups per CU)
  __shared__ double Pr[KH][SYT + 1];
  __shared__ double Qc[KH][SYT + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  int I = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= b) ++I;
  while (I * (I + 1) / 2 > b) --I;
  const int J = b - I * (I + 1) / 2;
  const long R0 = (long)(a0 + I) * SYT, C0 = (long)(a0 + J) * SYT;
  const int wm = wave >> 1, wn = wave & 1;
  trd_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) acc[i][jj] = trd_d4{0.0, 0.0, 0.0, 0.0};
  for (int h0 = 0; h0 < SYK; h0 += KH) {
    // panels: Pr[kk][x] = P[h0 + kk][R0 + x], Qc[kk][y] = Q[h0 + kk][C0 + y] (zero outside [j1, p))
    if (h0 > 0) __syncthreads();
    for (int e = tid; e < KH * SYT; e += 256) {
      const int kk = e / SYT, x = e % SYT;
      const long r = R0 + x, c = C0 + x;
      Pr[kk][x] = (r >= j1 && r < p) ? P[(long)(h0 + kk) * p + r] : 0.0;
      Qc[kk][x] = (c >= j1 && c < p) ? Q[(long)(h0 + kk) * p + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KH; kk += 4) {
      const int kr = kk + (lane >> 4);
      double fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i] = Pr[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) fb[jj] = Qc[kr][wn * 32 + jj * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
          acc[i][jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[jj], acc[i][jj], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long r = R0 + wm * 32 + i * 16 + (lane >> 4) + 4 * q;
        const long c = C0 + wn * 32 + jj * 16 + (lane & 15);
        if (r >= j1 && r < p && c >= j1 && c < p) A[r * ld + c] -= acc[i][jj][q];
      }
}

// A[c, r] = A[r, c] for j0 <= c < r < p: the upper triangle of the trailing
// block rebuilt from the lower one, once, before kernels that read whole rows
// take over.  Grid: nt (nt + 1) / 2 tiles; diagonal tiles mirror in place.
__global__ __launch_bounds__(256) void trd_mirror_lower_kernel(double* __restrict__ A, long ld, int p, int j0,
                                                               int a0) {
  __shared__ double T[SYT][SYT + 1];
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  int I = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= b) ++I;
  while (I * (I + 1) / 2 > b) --I;
  const int J = b - I * (I + 1) / 2;
  const long R0 = (long)(a0 + I) * SYT, C0 = (long)(a0 + J) * SYT;
  for (int e = tid; e < SYT * SYT; e += 256) {
    const int x = e / SYT, y = e % SYT;
    const long r = R0 + x, c = C0 + y;
    T[x][y] = (r >= j0 && r < p && c >= j0 && c < p) ? A[r * ld + c] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < SYT * SYT; e += 256) {
    const int y = e / SYT, x = e % SYT;  // write row C0 + y, column R0 + x
    const long r = R0 + x, c = C0 + y;
    if (r >= j0 && r < p && c >= j0 && c < p && r > c) A[c * ld + r] = T[x][y];
  }
}

// ---- the last columns in LDS: the trailing block A[j0:, j0:] (m0 = p - j0
// <= TAIL rows) is reduced by one workgroup without leaving LDS; reflectors,
// d, e and tau are written exactly as by the per-column kernels.
constexpr int TAIL = 136;  // 136 x 137 doubles + two vectors: 148 KB of LDS
__global__ __launch_bounds__(256) void trd_tail_lds_kernel(double* __restrict__ A, long ld, int p, int j0,
                                                           double* __restrict__ dv, double* __restrict__ ev,
                                                           double* __restrict__ tauv) {
  extern __shared__ double lds[];
  __shared__ double red[4];
  const int m0 = p - j0, L = m0 + 1;
  double* M = lds;               // m0 x L
  double* vs = M + (long)m0 * L; // m0
  double* ys = vs + m0;          // m0
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {  // load the block in batches of 8 elements per thread (loads before stores)
    const int tot = m0 * m0;
    for (int base = 0; base < tot; base += 256 * 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + 256 * u + tid;
        t[u] = (idx < tot) ? A[(long)(j0 + idx / m0) * ld + (j0 + idx % m0)] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + 256 * u + tid;
        if (idx < tot) M[(idx / m0) * L + idx % m0] = t[u];
      }
    }
  }
  __syncthreads();
  for (int jj = 0; jj < m0 - 2; ++jj) {
    const int j = j0 + jj, m = m0 - jj - 1;
    const double* x = M + jj * L + (jj + 1);
    double ss = 0.0;
    for (int c = tid; c < m; c += 256) {
      const double t = x[c];
      vs[c] = t;
      if (c > 0) ss += t * t;
    }
    ss = block_sum(ss, red);
    const double x0 = vs[0];
    double tau = 0.0, alpha = x0, scale = 0.0;
    if (ss > 0.0) {
      const double nrm = sqrt(x0 * x0 + ss);
      alpha = (x0 >= 0.0) ? -nrm : nrm;
      tau = (alpha - x0) / alpha;
      scale = 1.0 / (x0 - alpha);
    }
    __syncthreads();
    for (int c = tid; c < m; c += 256) {
      const double v = (c == 0) ? 1.0 : vs[c] * scale;
      vs[c] = v;
      A[(long)(j + 1 + c) * ld + j] = v;
    }
    if (tid == 0) {
      dv[j] = M[jj * L + jj];
      ev[j] = alpha;
      tauv[j] = tau;
    }
    __syncthreads();
    if (tau == 0.0) continue;  // uniform
    // y = tau M22 v, one thread per row (odd row stride: no bank conflicts)
    double dot = 0.0;
    if (tid < m) {
      const double* row = M + (jj + 1 + tid) * L + (jj + 1);
      double sacc = 0.0;
      for (int c = 0; c < m; ++c) sacc += row[c] * vs[c];
      const double yr = tau * sacc;
      ys[tid] = yr;
      dot = yr * vs[tid];
    }
    dot = block_sum(dot, red);
    const double hs = 0.5 * tau * dot;
    for (int c = tid; c < m; c += 256) ys[c] = ys[c] - hs * vs[c];  // w
    __syncthreads();
    for (int c = lane; c < m; c += 64) {
      const double vc = vs[c], wc = ys[c];
      for (int r = wave; r < m; r += 4) M[(jj + 1 + r) * L + (jj + 1 + c)] -= vs[r] * wc + ys[r] * vc;
    }
    __syncthreads();
  }
  if (tid == 0) {
    dv[p - 2] = M[(m0 - 2) * L + (m0 - 2)];
    ev[p - 2] = M[(m0 - 2) * L + (m0 - 1)];
    dv[p - 1] = M[(m0 - 1) * L + (m0 - 1)];
  }
}

// last 2 x 2 corner and the explicit reflector matrix V (p x p, column j =
// v_j with zeros above row j + 1; columns j >= p - 2 are zero)
__global__ void trd_tail_kernel(const double* __restrict__ A, long ld, int p, double* dv, double* ev) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (p >= 2) {
    dv[p - 2] = A[(long)(p - 2) * ld + (p - 2)];
    ev[p - 2] = A[(long)(p - 2) * ld + (p - 1)];
  }
  dv[p - 1] = A[(long)(p - 1) * ld + (p - 1)];
}
__global__ __launch_bounds__(256) void trd_extract_v_kernel(const double* __restrict__ A, long ld, int p,
                                                            const double* __restrict__ tauv, double* __restrict__ V) {
  const long tot = (long)p * p;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const int r = (int)(i / p), c = (int)(i % p);
    V[i] = (c < p - 2 && r > c && tauv[c] != 0.0) ? A[(long)r * ld + c] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// 1z. Hermitian tridiagonalisation (complex128, interleaved) to a REAL
// symmetric tridiagonal: reflectors H = I - tau v v^H with v[0] = 1 chosen so
// that H^H x = beta e_1 with beta real (the zlarfg convention); the update of
// step j is G22 <- H^H G22 H = G22 - v w^H - w v^H with y = tau G22 v,
// alpha = -(tau/2) (y^H v) (real) and w = y + alpha v.  Reflectors are kept
// below the diagonal; p - 1 of them (the last one only removes a phase).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 zinv(double2 a) {
  const double d = a.x * a.x + a.y * a.y;
  return make_double2(a.x / d, -a.y / d);
}
__device__ double2 block_sum_z(double2 v, double* red) {
  const double re = block_sum(v.x, red);
  const double im = block_sum(v.y, red);
  return make_double2(re, im);
}
// reflector from x (length m, x[0] = alpha): tau, beta and the scale of x[1:]
struct ZRefl {
  double2 tau, scale;
  double beta;
};
__device__ ZRefl zlarfg(double2 alpha, double xnorm2) {
  ZRefl r;
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    r.tau = make_double2(0.0, 0.0);
    r.beta = alpha.x;
    r.scale = make_double2(0.0, 0.0);
    return r;
  }
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
  r.beta = (alpha.x >= 0.0) ? -nrm : nrm;
  r.tau = make_double2((r.beta - alpha.x) / r.beta, -alpha.y / r.beta);
  r.scale = zinv(make_double2(alpha.x - r.beta, alpha.y));
  return r;
}

// step j: x = column j below the diagonal = conj(row j right of it)
__global__ __launch_bounds__(256) void trdz_reflect_kernel(double2* __restrict__ A, long ld, int p, int j,
                                                           double2* __restrict__ y, double2* __restrict__ part,
                                                           double* __restrict__ dv, double* __restrict__ ev,
                                                           double2* __restrict__ tauv) {
  extern __shared__ double2 zv[];
  __shared__ double red[4];
  const int m = p - j - 1, tid = threadIdx.x;
  const double2* xr = A + (long)j * ld + (j + 1);
  double ss = 0.0;
  for (int c = tid; c < m; c += 256) {
    const double2 t = zconj(xr[c]);
    zv[c] = t;
    if (c > 0) ss += t.x * t.x + t.y * t.y;
  }
  ss = block_sum(ss, red);
  const double2 alpha = zv[0];
  const ZRefl R = zlarfg(alpha, ss);
  __syncthreads();
  for (int c = tid; c < m; c += 256) zv[c] = (c == 0) ? make_double2(1.0, 0.0) : zmul(zv[c], R.scale);
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int c = tid; c < m; c += 256) A[(long)(j + 1 + c) * ld + j] = zv[c];
    if (tid == 0) {
      dv[j] = A[(long)j * ld + j].x;
      ev[j] = R.beta;
      tauv[j] = R.tau;
    }
  }
  if (R.tau.x == 0.0 && R.tau.y == 0.0) {
    if (tid == 0) part[blockIdx.x] = make_double2(0.0, 0.0);
    return;
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int r0 = blockIdx.x * HR;
  double2 dot = make_double2(0.0, 0.0);
  for (int rr = wave; rr < HR; rr += 4) {
    const int r = r0 + rr;
    if (r >= m) break;
    const double2* row = A + (long)(j + 1 + r) * ld + (j + 1);
    double sre = 0.0, sim = 0.0;
    for (int c = lane; c < m; c += 64) {
      const double2 t = zmul(row[c], zv[c]);
      sre += t.x;
      sim += t.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
      sre += __shfl_down(sre, o, 64);
      sim += __shfl_down(sim, o, 64);
    }
    if (lane == 0) {
      const double2 yr = zmul(R.tau, make_double2(sre, sim));
      y[r] = yr;
      dot = zadd(dot, zmul(zconj(yr), zv[r]));  // y^H v
    }
  }
  dot = block_sum_z(dot, red);
  if (tid == 0) part[blockIdx.x] = dot;
}

__global__ __launch_bounds__(256) void trdz_update_kernel(double2* __restrict__ A, long ld, int p, int j,
                                                          const double2* __restrict__ y,
                                                          const double2* __restrict__ part, int nparts,
                                                          const double2* __restrict__ tauv) {
  const double2 tau = tauv[j];
  if (tau.x == 0.0 && tau.y == 0.0) return;
  __shared__ double red[4];
  __shared__ double2 vr[64], wr[64], vc[64], wc[64];
  const int m = p - j - 1, tid = threadIdx.x;
  double2 s = make_double2(0.0, 0.0);
  for (int i = tid; i < nparts; i += 256) s = zadd(s, part[i]);
  s = block_sum_z(s, red);
  const double2 al = zscale(zmul(tau, s), -0.5);  // alpha (real up to rounding)
  const int tr0 = blockIdx.y * 64, tc0 = blockIdx.x * 64;
  if (tid < 64) {
    const int r = tr0 + tid;
    if (r < m) {
      const double2 v = A[(long)(j + 1 + r) * ld + j];
      vr[tid] = v;
      wr[tid] = zadd(y[r], zmul(al, v));
    } else {
      vr[tid] = make_double2(0.0, 0.0);
      wr[tid] = make_double2(0.0, 0.0);
    }
  } else if (tid < 128) {
    const int c = tc0 + tid - 64;
    if (c < m) {
      const double2 v = A[(long)(j + 1 + c) * ld + j];
      vc[tid - 64] = v;
      wc[tid - 64] = zadd(y[c], zmul(al, v));
    } else {
      vc[tid - 64] = make_double2(0.0, 0.0);
      wc[tid - 64] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  const int cc = tid & 63;
  const int c = tc0 + cc;
  if (c >= m) return;
  const double2 cwc = zconj(wc[cc]), cvc = zconj(vc[cc]);
  for (int rr = tid >> 6; rr < 64; rr += 4) {
    const int r = tr0 + rr;
    if (r >= m) break;
    double2* a = A + (long)(j + 1 + r) * ld + (j + 1 + c);
    *a = zsub(*a, zadd(zmul(vr[rr], cwc), zmul(wr[rr], cvc)));
  }
}

// the last columns in LDS (m0 = p - j0 <= TAILZ), including the final
// one-element reflector that makes e_{p-2} real
constexpr int TAILZ = 90;
__global__ __launch_bounds__(256) void trdz_tail_lds_kernel(double2* __restrict__ A, long ld, int p, int j0,
                                                            double* __restrict__ dv, double* __restrict__ ev,
                                                            double2* __restrict__ tauv) {
  extern __shared__ double2 zl[];
  __shared__ double red[4];
  const int m0 = p - j0, L = m0 + 1;
  double2* M = zl;                // m0 x L
  double2* vs = M + (long)m0 * L; // m0
  double2* ys = vs + m0;          // m0
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int r = wave; r < m0; r += 4)
    for (int c = lane; c < m0; c += 64) M[r * L + c] = A[(long)(j0 + r) * ld + (j0 + c)];
  __syncthreads();
  for (int jj = 0; jj < m0 - 1; ++jj) {
    const int j = j0 + jj, m = m0 - jj - 1;
    double ss = 0.0;
    for (int c = tid; c < m; c += 256) {
      const double2 t = zconj(M[jj * L + (jj + 1 + c)]);
      vs[c] = t;
      if (c > 0) ss += t.x * t.x + t.y * t.y;
    }
    ss = block_sum(ss, red);
    const ZRefl R = zlarfg(vs[0], ss);
    __syncthreads();
    for (int c = tid; c < m; c += 256) {
      const double2 v = (c == 0) ? make_double2(1.0, 0.0) : zmul(vs[c], R.scale);
      vs[c] = v;
      A[(long)(j + 1 + c) * ld + j] = v;
    }
    if (tid == 0) {
      dv[j] = M[jj * L + jj].x;
      ev[j] = R.beta;
      tauv[j] = R.tau;
    }
    __syncthreads();
    if (R.tau.x == 0.0 && R.tau.y == 0.0) continue;  // uniform
    double2 dot = make_double2(0.0, 0.0);
    if (tid < m) {
      const double2* row = M + (jj + 1 + tid) * L + (jj + 1);
      double sre = 0.0, sim = 0.0;
      for (int c = 0; c < m; ++c) {
        const double2 t = zmul(row[c], vs[c]);
        sre += t.x;
        sim += t.y;
      }
      const double2 yr = zmul(R.tau, make_double2(sre, sim));
      ys[tid] = yr;
      dot = zmul(zconj(yr), vs[tid]);
    }
    dot = block_sum_z(dot, red);
    const double2 al = zscale(zmul(R.tau, dot), -0.5);
    for (int c = tid; c < m; c += 256) ys[c] = zadd(ys[c], zmul(al, vs[c]));  // w
    __syncthreads();
    for (int c = lane; c < m; c += 64) {
      const double2 cvc = zconj(vs[c]), cwc = zconj(ys[c]);
      for (int r = wave; r < m; r += 4) {
        double2* a = M + (jj + 1 + r) * L + (jj + 1 + c);
        *a = zsub(*a, zadd(zmul(vs[r], cwc), zmul(ys[r], cvc)));
      }
    }
    __syncthreads();
  }
  if (tid == 0) dv[p - 1] = M[(m0 - 1) * L + (m0 - 1)].x;
}

// explicit complex reflectors (p x p): column j = v_j below row j, zero elsewhere
__global__ __launch_bounds__(256) void trdz_extract_v_kernel(const double2* __restrict__ A, long ld, int p,
                                                             const double2* __restrict__ tauv,
                                                             double2* __restrict__ V) {
  const long tot = (long)p * p;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const int r = (int)(i / p), c = (int)(i % p);
    const bool live = c < p - 1 && r > c && (tauv[c].x != 0.0 || tauv[c].y != 0.0);
    V[i] = live ? A[(long)r * ld + c] : make_double2(0.0, 0.0);
  }
}

// W <- T W (complex), T^{-1} = diag(1/tau) + striu(S), S = V^H V; one thread per column
__global__ __launch_bounds__(256) void wyz_tsolve_kernel(const double2* __restrict__ S, const double2* __restrict__ tauv,
                                                         int j0, int nb, double2* __restrict__ W, long n) {
  constexpr int NB = 32;  // reflectors per complex block
  __shared__ double2 Ss[NB][NB + 1];
  __shared__ double2 tt[NB];
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) Ss[idx / NB][idx % NB] = S[idx];
  if (threadIdx.x < NB) tt[threadIdx.x] = (threadIdx.x < nb) ? tauv[j0 + threadIdx.x] : make_double2(0.0, 0.0);
  __syncthreads();
  const long col = blockIdx.x * 256L + threadIdx.x;
  if (col >= n) return;
  double2 x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i < nb) ? W[(long)i * n + col] : make_double2(0.0, 0.0);
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
    if (i < nb) {
      double2 acc = x[i];
#pragma unroll
      for (int jx = i + 1; jx < NB; ++jx)
        if (jx < nb) acc = zsub(acc, zmul(Ss[i][jx], x[jx]));
      x[i] = zmul(tt[i], acc);
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < nb) W[(long)i * n + col] = x[i];
}

// real p x p (ld p) -> complex (ld p), and Vt[i][r] = conj(E[r][i]) (ld ldv)
__global__ __launch_bounds__(256) void real_to_z_kernel(const double* __restrict__ Q, long n, double2* __restrict__ E) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) E[i] = make_double2(Q[i], 0.0);
}
__global__ __launch_bounds__(256) void adjoint_z_kernel(const double2* __restrict__ E, int p, double2* __restrict__ Vt,
                                                        long ldv) {
  __shared__ double2 t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    t[k][tx] = (r < p && c < p) ? E[(long)r * p + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int i = c0 + k, r = r0 + tx;  // Vt[i][r] = conj(E[r][i])
    if (i < p && r < p) Vt[(long)i * ldv + r] = zconj(t[tx][k]);
  }
}

// ---------------------------------------------------------------------------
// 2. divide and conquer
// ---------------------------------------------------------------------------
// Leaf Jacobi: block [a, a + n) of the torn tridiagonal (n <= LEAF = 32),
// one wavefront per leaf; eigenvectors into Q[a:a+n, a:a+n] (columns),
// values into lam[a:a+n].  Lane t owns rotation pair t/4 of a step and 8 of
// its columns (rows).
__global__ __launch_bounds__(64) void dc_leaf_kernel(const double* __restrict__ d, const double* __restrict__ e,
                                                     const int* __restrict__ leaf_a, const int* __restrict__ leaf_n,
                                                     double tol_abs, double* __restrict__ Q, long ldq,
                                                     double* __restrict__ lam) {
  constexpr int P = LEAF, LD = LEAF + 1;
  __shared__ double Gs[P][LD];
  __shared__ double Js[P][LD];
  __shared__ double cs[P / 2], sn[P / 2];
  __shared__ int pi_[P / 2], pj_[P / 2], fl[P / 2];
  const int a = leaf_a[blockIdx.x], n = leaf_n[blockIdx.x], tid = threadIdx.x;
  for (int idx = tid; idx < P * P; idx += 64) {
    const int r = idx / P, c = idx % P;
    double g = 0.0;
    if (r < n && c < n) {
      if (r == c) g = d[a + r];
      else if (c == r + 1) g = e[a + r];
      else if (r == c + 1) g = e[a + c];
    }
    Gs[r][c] = g;
    Js[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double tol_rel = 64.0 * EPS;
  const int pp = tid >> 2, sub = tid & 3;
  // the tournament runs over the leaf's own size rounded up to even (Pe - 1
  // rounds of Pe / 2 pairs), not over all P slots: a padding index has no
  // coupling and never rotates
  const int Pe = (n + 1) & ~1;
  for (int sweep = 0; sweep < 60; ++sweep) {
    int nsweep = 0;
    for (int step = 0; step < Pe - 1; ++step) {
      int f = 0;
      if (tid < P / 2 && tid >= Pe / 2) {
        pi_[tid] = 0; pj_[tid] = 0; cs[tid] = 1.0; sn[tid] = 0.0; fl[tid] = 0;
      } else if (tid < P / 2) {
        const int pa = tid, pb = Pe - 1 - tid;
        const int i = (pa == 0) ? 0 : 1 + ((pa - 1 + step) % (Pe - 1));
        const int jj = 1 + ((pb - 1 + step) % (Pe - 1));
        const double x = Gs[i][i], z = Gs[jj][jj], g = Gs[i][jj];
        double c = 1.0, s = 0.0;
        if (fabs(g) > fmax(tol_rel * sqrt(fabs(x)) * sqrt(fabs(z)), tol_abs)) {
          const double zeta = (z - x) / (2.0 * g);
          double t;
          if (fabs(zeta) > 1e150) t = 0.5 / zeta;
          else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          c = 1.0 / sqrt(1.0 + t * t);
          s = c * t;
          f = 1;
        }
        pi_[tid] = i; pj_[tid] = jj; cs[tid] = c; sn[tid] = s; fl[tid] = f;
      }
      const int nrot = __syncthreads_count(f);
      if (nrot == 0) continue;
      nsweep += nrot;
      if (fl[pp]) {
        const int i = pi_[pp], jj = pj_[pp];
        const double c = cs[pp], s = sn[pp];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int col = sub * 8 + u;
          const double gi = Gs[i][col], gj = Gs[jj][col];
          Gs[i][col] = c * gi - s * gj;
          Gs[jj][col] = s * gi + c * gj;
        }
      }
      __syncthreads();
      if (fl[pp]) {
        const int i = pi_[pp], jj = pj_[pp];
        const double c = cs[pp], s = sn[pp];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int row = sub * 8 + u;
          const double gi = Gs[row][i], gj = Gs[row][jj];
          Gs[row][i] = c * gi - s * gj;
          Gs[row][jj] = s * gi + c * gj;
          const double ji = Js[row][i], jv = Js[row][jj];
          Js[row][i] = c * ji - s * jv;
          Js[row][jj] = s * ji + c * jv;
        }
      }
      __syncthreads();
    }
    if (nsweep == 0) break;
  }
  // padding indices never rotate with real ones (zero coupling), so the
  // first n columns of J restricted to the first n rows are the eigenvectors
  for (int idx = tid; idx < n * n; idx += 64) {
    const int r = idx / n, c = idx % n;
    Q[(long)(a + r) * ldq + (a + c)] = Js[r][c];
  }
  if (tid < n) lam[a + tid] = Gs[tid][tid];
}

// Zero the two off-diagonal sub-blocks of each merge block, so that every
// column of the block is explicit over all n rows.
struct MergeDev {
  int a, n1, n;    // block [a, a + n), split at a + n1
  int k;           // non-deflated roots
  int koff;        // offset into the packed root arrays
  int roff, nrot;  // deflation rotations
  long uoff;       // offset of this merge's U^T (k x k) in the packed buffer
  double rho;      // > 0
};

__global__ __launch_bounds__(256) void dc_zero_off_kernel(double* __restrict__ Q, long ldq,
                                                          const MergeDev* __restrict__ md) {
  const MergeDev m = md[blockIdx.y];
  const long tot = (long)m.n * m.n;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const int r = (int)(i / m.n), c = (int)(i % m.n);
    if ((r < m.n1) != (c < m.n1)) Q[(long)(m.a + r) * ldq + (m.a + c)] = 0.0;
  }
}

struct Rot {
  int ca, cb;  // block-local columns
  double c, s;
};
// deflation rotations, in order, on the rows of each merge block
__global__ __launch_bounds__(256) void dc_givens_kernel(double* __restrict__ Q, long ldq,
                                                        const MergeDev* __restrict__ md,
                                                        const Rot* __restrict__ rots) {
  const MergeDev m = md[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m.n || m.nrot == 0) return;
  double* row = Q + (long)(m.a + r) * ldq + m.a;
  for (int t = 0; t < m.nrot; ++t) {
    const Rot R = rots[m.roff + t];
    const double x = row[R.ca], y = row[R.cb];
    row[R.ca] = R.c * x + R.s * y;
    row[R.cb] = R.c * y - R.s * x;
  }
}

// secular equation 1 + rho sum_q z_q^2 / (d_q - lambda) = 0 (rho > 0, d
// increasing) for root i of its merge: returns the origin pole o (the pole of
// the interval (d_i, d_{i+1}) nearer to the root) and tau = lambda - d_o.
// Iteration: Li's "middle way" -- the poles on the origin side of the root
// and those on the far side are each replaced by one rational term at their
// nearest pole matching value and slope, and the resulting quadratic is
// solved for tau; bisection of the sign bracket guards every step.
// One wavefront per root, the sums over the poles split over its lanes.
__global__ __launch_bounds__(256) void dc_secular_kernel(const MergeDev* __restrict__ md,
                                                         const int* __restrict__ root_merge,
                                                         const double* __restrict__ dK,
                                                         const double* __restrict__ zK, int nroots,
                                                         int* __restrict__ org, double* __restrict__ tau) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= nroots) return;  // the whole wavefront leaves
  const MergeDev m = md[root_merge[g]];
  const int i = g - m.koff, k = m.k;
  const double* dd = dK + m.koff;
  const double* zz = zK + m.koff;
  const double rho = m.rho;
  int o;
  double lo, hi, df;  // bracket for tau, far pole relative to the origin (0: none)
  if (i < k - 1) {
    const double mid = 0.5 * (dd[i] + dd[i + 1]);
    double f = 0.0;
    for (int q = lane; q < k; q += 64) f += rho * zz[q] * zz[q] / (dd[q] - mid);
    f = 1.0 + wave_sum64(f);
    if (f >= 0.0) {
      o = i; lo = 0.0; hi = mid - dd[i]; df = dd[i + 1] - dd[i];
    } else {
      o = i + 1; lo = mid - dd[i + 1]; hi = 0.0; df = dd[i] - dd[i + 1];
    }
  } else {
    o = i;
    lo = 0.0;
    df = 0.0;
    double z2 = 0.0;
    for (int q = lane; q < k; q += 64) z2 += zz[q] * zz[q];
    hi = rho * wave_sum64(z2);
    for (int it = 0; it < 64; ++it) {  // make sure the bracket holds the root
      double gh = 0.0;
      for (int q = lane; q < k; q += 64) gh += rho * zz[q] * zz[q] / ((dd[q] - dd[o]) - hi);
      gh = 1.0 + wave_sum64(gh);
      if (gh >= 0.0) break;
      hi *= 2.0;
    }
  }
  const double dO = dd[o];
  // poles q with (q <= i) == (o == i) are on the origin side
  const bool left_origin = (o == i);
  double t = 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    double so = 0.0, dso = 0.0, sf = 0.0, dsf = 0.0, sabs = 0.0;
    for (int q = lane; q < k; q += 64) {
      const double inv = 1.0 / ((dd[q] - dO) - t);
      const double r = rho * zz[q] * zz[q] * inv;
      sabs += fabs(r);
      if ((q <= i) == left_origin) { so += r; dso += r * inv; }
      else { sf += r; dsf += r * inv; }
    }
    so = wave_sum64(so);
    dso = wave_sum64(dso);
    sf = wave_sum64(sf);
    dsf = wave_sum64(dsf);
    sabs = wave_sum64(sabs);
    const double gv = 1.0 + so + sf;
    // converged once g is within its own rounding error
    if (fabs(gv) <= 8.0 * EPS * (1.0 + sabs)) break;
    if (gv < 0.0) lo = t;
    else hi = t;
    // model: c + bo / (0 - x) + bf / (df - x), matching value and slope at t
    const double bo = dso * t * t;
    const double ao = so - bo / (-t);
    double bf = 0.0, af = 0.0;
    if (df != 0.0) {
      bf = dsf * (df - t) * (df - t);
      af = sf - bf / (df - t);
    } else {
      af = sf;
    }
    const double cc = 1.0 + ao + af;
    // c x^2 - (c df + bo + bf) x + bo df = 0  (df = 0: c x - bo = 0)
    double tn;
    if (df == 0.0) {
      tn = (cc > 0.0) ? bo / cc : 0.5 * (lo + hi);
    } else {
      const double A = cc, B = -(cc * df + bo + bf), Cq = bo * df;
      if (A == 0.0) {
        tn = -Cq / B;
      } else {
        const double disc = fmax(B * B - 4.0 * A * Cq, 0.0);
        const double qq = -0.5 * (B + copysign(sqrt(disc), B));
        const double t1 = qq / A, t2 = (qq != 0.0) ? Cq / qq : t1;
        tn = (t2 > lo && t2 < hi) ? t2 : t1;
      }
    }
    if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
    const bool done = fabs(tn - t) <= 4.0 * EPS * fabs(tn) || (hi - lo) <= 4.0 * EPS * fmax(fabs(lo), fabs(hi));
    t = tn;
    if (done) break;
  }
  if (lane == 0) {
    org[g] = o;
    tau[g] = t;
  }
}

// Gu-Eisenstat: z-hat_j so that the computed roots are exact eigenvalues;
// one wavefront per j, the product over the roots split over its lanes
__global__ __launch_bounds__(256) void dc_zhat_kernel(const MergeDev* __restrict__ md,
                                                      const int* __restrict__ root_merge,
                                                      const double* __restrict__ dK,
                                                      const double* __restrict__ zK,
                                                      const int* __restrict__ org, const double* __restrict__ tau,
                                                      int nroots, double* __restrict__ zh) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= nroots) return;  // the whole wavefront leaves
  const MergeDev m = md[root_merge[g]];
  const int j = g - m.koff, k = m.k;
  const double* dd = dK + m.koff;
  const int* oo = org + m.koff;
  const double* tt = tau + m.koff;
  const double dj = dd[j];
  // lambda_i - d_j = (d_o(i) - d_j) + tau_i; by interlacing every factor
  // below lies in (0, 1)
  double prod = 1.0;
  for (int i = lane; i < k - 1; i += 64) {
    const double den = (i < j) ? (dd[i] - dj) : (dd[i + 1] - dj);
    prod *= ((dd[oo[i]] - dj) + tt[i]) / den;
  }
  prod = wave_prod64(prod) * (((dd[oo[k - 1]] - dj) + tt[k - 1]) / m.rho);
  if (lane == 0) zh[g] = copysign(sqrt(fmax(prod, 0.0)), zK[m.koff + j]);
}


// row i of U^T: u_i[j] = zh_j / (d_j - lambda_i), normalised; one workgroup per root
__global__ __launch_bounds__(256) void dc_evec_kernel(const MergeDev* __restrict__ md,
                                                      const int* __restrict__ root_merge,
                                                      const double* __restrict__ dK,
                                                      const double* __restrict__ zh,
                                                      const int* __restrict__ org, const double* __restrict__ tau,
                                                      double* __restrict__ Ut) {
  __shared__ double red[4];
  const int g = blockIdx.x;
  const MergeDev m = md[root_merge[g]];
  const int i = g - m.koff, k = m.k;
  const double* dd = dK + m.koff;
  const double* zz = zh + m.koff;
  const double dO = dd[org[g]], t = tau[g];
  double* u = Ut + m.uoff + (long)i * k;
  double ss = 0.0;
  for (int j = threadIdx.x; j < k; j += 256) {
    const double v = zz[j] / ((dd[j] - dO) - t);
    u[j] = v;
    ss += v * v;
  }
  ss = block_sum(ss, red);
  const double inv = 1.0 / sqrt(ss);
  for (int j = threadIdx.x; j < k; j += 256) u[j] *= inv;
}

// z of each merge: last row of its left block, first row of its right block
__global__ __launch_bounds__(256) void dc_zgather_kernel(const double* __restrict__ Q, long ldq,
                                                         const int* __restrict__ zl, double* __restrict__ z) {
  const int a = zl[3 * blockIdx.y], n1 = zl[3 * blockIdx.y + 1], n = zl[3 * blockIdx.y + 2];
  for (int t = threadIdx.x; t < n; t += 256) {
    const long r = (t < n1) ? (a + n1 - 1) : (a + n1);
    z[a + t] = Q[r * ldq + (a + t)];
  }
}

// new column order of a merge block: tmp[:, t] = Q[:, a + src[t]] (n x n, ld n)
__global__ __launch_bounds__(256) void dc_gather_cols_kernel(const double* __restrict__ Q, long ldq, int a, int n,
                                                             const int* __restrict__ src, double* __restrict__ tmp) {
  const long tot = (long)n * n;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const int r = (int)(i / n), t = (int)(i % n);
    tmp[i] = Q[(long)(a + r) * ldq + (a + src[t])];
  }
}
// Q[a + r][a + t] = tmp[r][t] for t in [t0, n)
__global__ __launch_bounds__(256) void dc_put_cols_kernel(double* __restrict__ Q, long ldq, int a, int n, int t0,
                                                          const double* __restrict__ tmp) {
  const int w = n - t0;
  const long tot = (long)n * w;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const int r = (int)(i / w), t = t0 + (int)(i % w);
    Q[(long)(a + r) * ldq + (a + t)] = tmp[(long)r * n + t];
  }
}

// ---------------------------------------------------------------------------
// 3. back-transformation helpers
// ---------------------------------------------------------------------------
// W <- T W for the compact WY form, through T^{-1} = diag(1/tau) + striu(S),
// S = V^T V (Joffrain et al.): back-substitution, one thread per column of
// W (nb x n, ld n).  A reflector with tau = 0 has a zero column in V.
__global__ __launch_bounds__(256) void wy_tsolve_kernel(const double* __restrict__ S, const double* __restrict__ tauv,
                                                        int j0, int nb, double* __restrict__ W, long n) {
  __shared__ double Ss[NBW][NBW + 1];
  __shared__ double dinv[NBW];
  for (int idx = threadIdx.x; idx < NBW * NBW; idx += 256) Ss[idx / NBW][idx % NBW] = S[idx];
  if (threadIdx.x < NBW) {
    const double t = (threadIdx.x < nb) ? tauv[j0 + threadIdx.x] : 0.0;
    dinv[threadIdx.x] = (t != 0.0) ? t : 0.0;  // T_ii = tau_i (T^{-1}_ii = 1/tau_i)
  }
  __syncthreads();
  const long col = blockIdx.x * 256L + threadIdx.x;
  if (col >= n) return;
  double x[NBW];
#pragma unroll
  for (int i = 0; i < NBW; ++i) x[i] = (i < nb) ? W[(long)i * n + col] : 0.0;
  // solve (D^{-1} + U) y = x from the bottom: y_i = tau_i (x_i - sum_{j>i} S_ij y_j),
  // column-oriented: once y_i is final it is subtracted from every x_j, j < i,
  // so the dependent chain is one step per row, not one per product
#pragma unroll
  for (int i = NBW - 1; i >= 0; --i) {
    if (i < nb) {
      const double yi = dinv[i] * x[i];
      x[i] = yi;
#pragma unroll
      for (int j = 0; j < i; ++j) x[j] -= Ss[j][i] * yi;
    }
  }
#pragma unroll
  for (int i = 0; i < NBW; ++i)
    if (i < nb) W[(long)i * n + col] = x[i];
}

unsigned grid_for(long tot, int cap = 4096) { return (unsigned)std::max<long>(1, std::min<long>(cap, (tot + 255) / 256)); }

}  // namespace

// host deflation of one merge (block-local indices); see the file comment
struct HostMerge {
  int a, n1, n;
  double sgn;                  // eigenvalues of the solved problem are sgn * lambda
  std::vector<int> K;          // non-deflated columns, by increasing solved d
  std::vector<double> dK, zK;  // solved problem (rho > 0)
  std::vector<int> defl;       // deflated columns
  std::vector<double> ddefl;   // their eigenvalues (true sign)
  std::vector<Rot> rots;
  double rho;
};

static void deflate(HostMerge& hm, std::vector<double> D, std::vector<double> z, double beta) {
  const int n = hm.n;
  double rho = 2.0 * beta;
  hm.sgn = 1.0;
  if (rho < 0.0) {
    hm.sgn = -1.0;
    rho = -rho;
    for (double& v : D) v = -v;
  }
  hm.rho = rho;
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return D[x] < D[y]; });
  double dmax = 0.0, zmax = 0.0;
  for (int i = 0; i < n; ++i) {
    dmax = std::max(dmax, std::fabs(D[i]));
    zmax = std::max(zmax, std::fabs(z[i]));
  }
  const double tol = 8.0 * EPS * std::max(dmax, rho * zmax);
  std::vector<char> dfl(n, 0);
  int pj = -1;
  for (int c : perm) {
    if (rho * std::fabs(z[c]) <= tol) {
      dfl[c] = 1;
      continue;
    }
    if (pj < 0) {
      pj = c;
      continue;
    }
    const double s = z[pj], cc = z[c];
    const double t = std::hypot(cc, s);
    const double cth = cc / t, sth = -s / t;
    if (std::fabs((D[c] - D[pj]) * cth * sth) <= tol) {
      hm.rots.push_back(Rot{pj, c, cth, sth});
      const double dp = D[pj] * cth * cth + D[c] * sth * sth;
      const double dj = D[pj] * sth * sth + D[c] * cth * cth;
      D[pj] = dp;
      D[c] = dj;
      z[c] = t;
      z[pj] = 0.0;
      dfl[pj] = 1;
    }
    pj = c;
  }
  for (int c = 0; c < n; ++c) {
    if (dfl[c]) {
      hm.defl.push_back(c);
      hm.ddefl.push_back(hm.sgn * D[c]);
    } else {
      hm.K.push_back(c);
    }
  }
  std::stable_sort(hm.K.begin(), hm.K.end(), [&](int x, int y) { return D[x] < D[y]; });
  for (int c : hm.K) {
    hm.dK.push_back(D[c]);
    hm.zK.push_back(z[c]);
  }
}

// TNHIP_DEBUG_EIG: stage times (each stage synchronised) on stderr
struct EigStages {
  tn_ctx* c;
  long p;
  bool on;
  std::chrono::steady_clock::time_point t;
  EigStages(tn_ctx* ctx, long n) : c(ctx), p(n), on(std::getenv("TNHIP_DEBUG_EIG") != nullptr),
                                   t(std::chrono::steady_clock::now()) {}
  void operator()(const char* name) {
    if (!on) return;
    TN_HIP(hipStreamSynchronize(c->stream));
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[tnhip dc] p=%ld %s %.3f ms\n", p, name,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

static size_t level_bytes(long p) {
  return (size_t)p * (5 * 8 + 6 * 4) + (size_t)p * sizeof(MergeDev) + (size_t)p * sizeof(Rot) + 1024;
}

// Divide and conquer on the symmetric tridiagonal (d, e), p >= 2: Q (p x p,
// ld p) receives the eigenvectors as columns; returns their eigenvalues.
static std::vector<double> dc_tridiag(tn_ctx* c, long p, const std::vector<double>& d, const std::vector<double>& e,
                                      double* Q, EigStages& stage) {
  const int P = (int)p;
  const size_t lvb = level_bytes(p);
  char* hs = (char*)c->host_stage(lvb + 8 * (size_t)p * sizeof(double) + 256);
  double* hx = (double*)(hs + ((lvb + 255) & ~(size_t)255));
  // ---- 2. divide and conquer on (d, e)
  double tnorm = 0.0;
  for (long i = 0; i < p; ++i) tnorm = std::max(tnorm, std::fabs(d[i]));
  for (long i = 0; i + 1 < p; ++i) tnorm = std::max(tnorm, 2.0 * std::fabs(e[i]));
  struct Blk { int a, n; };
  std::vector<Blk> blocks;
  {  // ceil(P / LEAF) leaves of nearly equal size (the slowest leaf sets the time)
    const int nl = (P + LEAF - 1) / LEAF;
    for (int b = 0, a = 0; b < nl; ++b) {
      const int nb = P / nl + (b < P % nl ? 1 : 0);
      blocks.push_back(Blk{a, nb});
      a += nb;
    }
  }
  std::vector<double> dt = d;  // torn diagonal
  for (size_t b = 1; b < blocks.size(); ++b) {
    const int k = blocks[b].a;
    dt[k - 1] -= e[k - 1];
    dt[k] -= e[k - 1];
  }
  const long ldq = p;
  double* small = (double*)c->workspace(WS_DC_SMALL, (size_t)(8 * p + 64) * sizeof(double));
  double* d_dt = small;
  double* d_e = d_dt + p;
  double* d_lam = d_e + p;
  int* d_la = (int*)(d_lam + p);
  int* d_ln = d_la + p;
  {
    // mirror of the `small` layout: dt, e, lam (p doubles each), leaf starts and sizes (p ints each)
    std::copy(dt.begin(), dt.end(), hx);
    std::copy(e.begin(), e.end(), hx + p);
    int* hla = (int*)(hx + 3 * p);
    int* hln = hla + p;
    for (size_t b = 0; b < blocks.size(); ++b) { hla[b] = blocks[b].a; hln[b] = blocks[b].n; }
    TN_HIP(hipMemcpyAsync(small, hx, (size_t)3 * p * sizeof(double) + 2 * (size_t)p * sizeof(int),
                          hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(dc_leaf_kernel, dim3((unsigned)blocks.size()), dim3(64), 0, c->stream, d_dt, d_e, d_la,
                       d_ln, EPS * tnorm, Q, ldq, d_lam);
    TN_HIP(hipGetLastError());
  }
  TN_HIP(hipMemcpyAsync(hx + 2 * p, d_lam, p * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  stage("leaves");
  TN_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> D(hx + 2 * p, hx + 3 * p);  // eigenvalues per block, block-local column order

  double* tmp = (double*)c->workspace(WS_DC_TMP, (size_t)p * p * sizeof(double));
  double* Utb = (double*)c->workspace(WS_DC_UT, (size_t)p * p * sizeof(double));
  // packed per-level device arrays: dK, zK, zh, tau (p doubles each), org,
  // root_merge, src (p ints each), merge descriptors and rotations
  char* lv = (char*)c->workspace(WS_DC_LEVEL, lvb);
  auto H = [&](const void* dev) { return hs + ((const char*)dev - lv); };  // host mirror of a level table
  double* d_dK = (double*)lv;
  double* d_zK = d_dK + p;
  double* d_zh = d_zK + p;
  double* d_tau = d_zh + p;
  int* d_org = (int*)(d_tau + p);
  int* d_rm = d_org + p;
  int* d_src = d_rm + p;
  MergeDev* d_md = (MergeDev*)(((uintptr_t)(d_src + p) + 15) & ~(uintptr_t)15);
  Rot* d_rot = (Rot*)(((uintptr_t)(d_md + p) + 15) & ~(uintptr_t)15);
  double* d_zrow = (double*)(((uintptr_t)(d_rot + p) + 15) & ~(uintptr_t)15);
  int* d_zl = (int*)(d_zrow + p);  // 3 ints per merge

  while (blocks.size() > 1) {
    std::vector<Blk> nb;
    std::vector<HostMerge> hms;
    // z of every merge: last row of the left block, first row of the right block
    int* hzl = (int*)H(d_zl);  // (a, n1, n) per merge
    int nzm = 0;
    for (size_t i = 0; i + 1 < blocks.size(); i += 2, ++nzm) {
      hzl[3 * nzm] = blocks[i].a;
      hzl[3 * nzm + 1] = blocks[i].n;
      hzl[3 * nzm + 2] = blocks[i].n + blocks[i + 1].n;
    }
    TN_HIP(hipMemcpyAsync(d_zl, hzl, 3 * nzm * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(dc_zgather_kernel, dim3(1, nzm), dim3(256), 0, c->stream, Q, ldq, d_zl, d_zrow);
    TN_HIP(hipMemcpyAsync(H(d_zrow), d_zrow, p * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    const double* zrow = (const double*)H(d_zrow);
    std::vector<MergeDev> mds;
    std::vector<Rot> allrot;
    std::vector<double> hdK, hzK;
    std::vector<int> hrm, hsrc;
    long uoff = 0;
    for (size_t i = 0; i + 1 < blocks.size(); i += 2) {
      const Blk b1 = blocks[i], b2 = blocks[i + 1];
      HostMerge hm;
      hm.a = b1.a;
      hm.n1 = b1.n;
      hm.n = b1.n + b2.n;
      std::vector<double> Dm(D.begin() + hm.a, D.begin() + hm.a + hm.n);
      std::vector<double> zm(zrow + hm.a, zrow + hm.a + hm.n);
      const double is2 = 1.0 / std::sqrt(2.0);
      for (double& v : zm) v *= is2;
      deflate(hm, Dm, zm, e[b2.a - 1]);
      MergeDev md;
      md.a = hm.a; md.n1 = hm.n1; md.n = hm.n;
      md.k = (int)hm.K.size();
      md.koff = (int)hdK.size();
      md.roff = (int)allrot.size();
      md.nrot = (int)hm.rots.size();
      md.uoff = uoff;
      md.rho = hm.rho;
      uoff += (long)md.k * md.k;
      for (int t = 0; t < md.k; ++t) hrm.push_back((int)mds.size());
      hdK.insert(hdK.end(), hm.dK.begin(), hm.dK.end());
      hzK.insert(hzK.end(), hm.zK.begin(), hm.zK.end());
      allrot.insert(allrot.end(), hm.rots.begin(), hm.rots.end());
      mds.push_back(md);
      hms.push_back(std::move(hm));
      nb.push_back(Blk{b1.a, b1.n + b2.n});
    }
    if (blocks.size() % 2) nb.push_back(blocks.back());
    const int nroots = (int)hdK.size();
    const int nm = (int)mds.size();
    // all tables of the level in one upload through the pinned mirror
    std::copy(mds.begin(), mds.end(), (MergeDev*)H(d_md));
    std::copy(allrot.begin(), allrot.end(), (Rot*)H(d_rot));
    std::copy(hdK.begin(), hdK.end(), (double*)H(d_dK));
    std::copy(hzK.begin(), hzK.end(), (double*)H(d_zK));
    std::copy(hrm.begin(), hrm.end(), (int*)H(d_rm));
    {
      int* hsrc = (int*)H(d_src);  // new column order of every merge block: [roots..., deflated...]
      for (int mi = 0; mi < nm; ++mi) {
        const HostMerge& hm = hms[mi];
        std::copy(hm.K.begin(), hm.K.end(), hsrc + hm.a);
        std::copy(hm.defl.begin(), hm.defl.end(), hsrc + hm.a + hm.K.size());
      }
    }
    TN_HIP(hipMemcpyAsync(lv, hs, (size_t)((char*)d_zrow - lv), hipMemcpyHostToDevice, c->stream));
    int maxn = 0;
    for (auto& m : mds) maxn = std::max(maxn, m.n);
    // explicit zeros off the two diagonal sub-blocks, then the deflation rotations
    hipLaunchKernelGGL(dc_zero_off_kernel, dim3(grid_for((long)maxn * maxn, 1024), nm), dim3(256), 0, c->stream, Q,
                       ldq, d_md);
    if (!allrot.empty())
      hipLaunchKernelGGL(dc_givens_kernel, dim3((maxn + 255) / 256, nm), dim3(256), 0, c->stream, Q, ldq, d_md,
                         d_rot);
    if (nroots > 0) {
      hipLaunchKernelGGL(dc_secular_kernel, dim3((unsigned)((nroots + 3) / 4)), dim3(256), 0, c->stream, d_md, d_rm,
                         d_dK, d_zK, nroots, d_org, d_tau);
      hipLaunchKernelGGL(dc_zhat_kernel, dim3((unsigned)((nroots + 3) / 4)), dim3(256), 0, c->stream, d_md, d_rm,
                         d_dK, d_zK, d_org, d_tau, nroots, d_zh);
      hipLaunchKernelGGL(dc_evec_kernel, dim3((unsigned)nroots), dim3(256), 0, c->stream, d_md, d_rm, d_dK, d_zh,
                         d_org, d_tau, Utb);
    }
    TN_HIP(hipGetLastError());
    // per merge: Q_blk[:, :k] = Q_K U, deflated columns after the roots
    for (int mi = 0; mi < nm; ++mi) {
      const HostMerge& hm = hms[mi];
      const MergeDev& md = mds[mi];
      int* d_s = d_src + hm.a;
      hipLaunchKernelGGL(dc_gather_cols_kernel, dim3(grid_for((long)hm.n * hm.n)), dim3(256), 0, c->stream, Q, ldq,
                         hm.a, hm.n, d_s, tmp);
      if (md.k > 0) {
        GemmDesc g;
        g.M = hm.n; g.N = md.k; g.K = md.k;
        g.A = tmp; g.lda = hm.n;
        g.B = Utb + md.uoff; g.ldb = md.k; g.tB = 1;
        g.C = Q + (long)hm.a * ldq + hm.a; g.ldc = ldq;
        gemm(c, Prec::F64, g);
      }
      if (md.k < hm.n)
        hipLaunchKernelGGL(dc_put_cols_kernel, dim3(grid_for((long)hm.n * (hm.n - md.k))), dim3(256), 0, c->stream,
                           Q, ldq, hm.a, hm.n, md.k, tmp);
      TN_HIP(hipGetLastError());
    }
    // eigenvalues of the merged blocks (host), in the new column order
    // tau (p doubles) and org (p ints) are adjacent: one download
    if (nroots > 0)
      TN_HIP(hipMemcpyAsync(H(d_tau), d_tau, (size_t)p * sizeof(double) + (size_t)nroots * sizeof(int),
                            hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    const int* horg = (const int*)H(d_org);
    const double* htau = (const double*)H(d_tau);
    for (int mi = 0; mi < nm; ++mi) {
      const HostMerge& hm = hms[mi];
      const MergeDev& md = mds[mi];
      for (int t = 0; t < md.k; ++t)
        D[hm.a + t] = hm.sgn * (hm.dK[horg[md.koff + t]] + htau[md.koff + t]);
      for (size_t t = 0; t < hm.defl.size(); ++t) D[hm.a + md.k + t] = hm.ddefl[t];
    }
    blocks.swap(nb);
    stage("merge level");
  }

  return D;
}

void wy_tsolve(tn_ctx* c, const double* S, const double* tauv, int j0, int nb, double* W, long n) {
  TN_REQUIRE(nb >= 1 && nb <= NBW, TN_E_INVALID, "wy_tsolve: block size");
  hipLaunchKernelGGL(wy_tsolve_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, S, tauv, j0, nb, W,
                     n);
  TN_HIP(hipGetLastError());
}

// smallest p reduced in panels: below it one fused launch per column is
// faster (measured: p = 2048 takes 45 ms fused against 57 ms in panels,
// p = 4096 245 ms against 171 ms)
constexpr int TRD_BLOCKED_MIN = 3072;
// smallest trailing block whose product reads only the lower triangle (two
// launches per column instead of one)
constexpr int TRD_TILE_MIN = 1024;

void sym_eig_dc(tn_ctx* c, double* G, long p, long ldg, double* Vt, long ldv, std::vector<double>& w) {
  TN_REQUIRE(p >= 3 && p <= 16384, TN_E_INVALID, "sym_eig_dc: size out of range");
  const int P = (int)p;
  EigStages stage(c, p);
  // ---- 1. tridiagonalisation
  double* vecs = (double*)c->workspace(WS_DC_VEC, (size_t)(12 * p + 2048) * sizeof(double));
  double* dv = vecs;
  double* ev = dv + p;
  double* tauv = ev + p;
  double* y = tauv + p;
  double* part = y + p;  // <= ceil(p / HR) entries
  // panels of NBT columns (one pass over the trailing block per column) from
  // p = TRD_BLOCKED_MIN on; below it one fused launch per column
  const int jt = std::max(0, P - TAIL);  // columns jt.. are reduced in LDS
  bool done = false;
  if (jt >= 1 && P >= TRD_BLOCKED_MIN) {
    const int nAmax = (P + RCHUNK - 1) / RCHUNK;
    const int nBmax = (P + 7) / 8;
    const long ntmax = (p + SYT - 1) / SYT + 1;
    double* Pp = (double*)c->workspace(WS_DC_PANEL, ((size_t)4 * NBT * p + p + (size_t)nAmax * (2 * NBT + 1) +
                                                     nBmax + 64 + (size_t)ntmax * ntmax * SYT) * sizeof(double));
    double* Qp = Pp + (size_t)2 * NBT * p;
    double* xb = Qp + (size_t)2 * NBT * p;
    double* partA = xb + p;
    double* partB = partA + (size_t)nAmax * (2 * NBT + 1);
    double* Sy = partB + nBmax + 64;  // tile slots of the lower-triangle product
    raise_lds_limit((const void*)trd_panel_symv_kernel, (size_t)p * sizeof(double));
    raise_lds_limit((const void*)trd_panel_row_kernel, TRD_ROW_LDS);
    // panel updates touch only the lower triangle while the next panel's
    // columns all take the lower-triangle product
    bool lower_only = P - NBT >= TRD_TILE_MIN;
    for (int j0 = 0; j0 < jt; j0 += NBT) {
      const int j1 = std::min(jt, j0 + NBT);
      fill_f64(c, Pp, 0.0, (long)4 * NBT * p);
      int nBprev = 0;
      for (int j = j0; j < j1; ++j) {
        const int k = j - j0, m = P - j - 1;
        const int nA = (m + RCHUNK - 1) / RCHUNK;
        hipLaunchKernelGGL(trd_panel_row_kernel, dim3(nA), dim3(256), TRD_ROW_LDS, c->stream, (const double*)G, ldg, P, j, k,
                           Pp, Qp, (const double*)y, (const double*)partB, nBprev, (const double*)tauv, xb, partA,
                           dv);
        if (m >= TRD_TILE_MIN) {
          // lower-triangle tiles, then the fixed-order reduction
          const int a0 = (j + 1) / SYT;
          const int nt = (P + SYT - 1) / SYT - a0;
          const int nB = (m + 31) / 32;  // eight lanes per row
          TN_REQUIRE(nA <= 64 && nB <= nBmax && nt <= ntmax, TN_E_INVALID, "tridiagonalisation: tile plan");
          PhaseTimer pt(c, PH_TRD_SYMV);
          if (c->timing) c->phase_bytes[PH_TRD_SYMV] += 4.0 * (double)m * (double)(m + SYT);  // lower half of G22
          hipLaunchKernelGGL(trd_symv_tile_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, c->stream,
                             (const double*)G, ldg, P, j, (const double*)xb, (const double*)partA, nA, a0, nt, Sy);
          hipLaunchKernelGGL(trd_symv_reduce_kernel, dim3(nB), dim3(256), 0, c->stream, G, ldg, P, j, k, Pp, Qp,
                             (const double*)xb, (const double*)partA, nA, a0, nt, (const double*)Sy, y, partB, ev,
                             tauv);
          nBprev = nB;
        } else {
          const int nB = (m + 7) / 8;  // two rows per wave
          PhaseTimer pt(c, PH_TRD_SYMV);
          if (c->timing) c->phase_bytes[PH_TRD_SYMV] += 8.0 * (double)m * (double)m;  // one read of G22
          hipLaunchKernelGGL(trd_panel_symv_kernel, dim3(nB), dim3(256), (size_t)m * sizeof(double), c->stream, G,
                             ldg, P, j, k, Pp, Qp, (const double*)xb, (const double*)partA, nA, y, partB, ev, tauv);
          nBprev = nB;
        }
      }
      const int jl = j1 - 1, ml = P - jl - 1;
      hipLaunchKernelGGL(trd_panel_wlast_kernel, dim3((unsigned)std::min(64, (ml + 255) / 256)), dim3(256), 0,
                         c->stream, P, jl, jl - j0, Pp, Qp, (const double*)y, (const double*)partB, nBprev,
                         (const double*)tauv);
      TN_HIP(hipGetLastError());
      const long m1 = p - j1;
      if (m1 > 0 && lower_only) {
        // G[j1:, j1:] -= V W^T + W V^T on the lower tiles; while every column
        // of the next panel reads only the lower triangle that is all that is
        // needed, and before the first one that reads whole rows (or the LDS
        // tail) the upper triangle is rebuilt from it once
        const int a0 = j1 / SYT;
        const int nt = (P + SYT - 1) / SYT - a0;
        hipLaunchKernelGGL(trd_syr2k_lower_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, c->stream, G,
                           ldg, P, j1, (const double*)Pp, (const double*)Qp, a0);
        if (P - j1 - NBT < TRD_TILE_MIN) {
          hipLaunchKernelGGL(trd_mirror_lower_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, c->stream,
                             G, ldg, P, j1, a0);
          lower_only = false;
        }
        TN_HIP(hipGetLastError());
      } else if (m1 > 0) {  // G[j1:, j1:] -= V W^T + W V^T (K = 2 NBT, unused rows of P, Q zero)
        GemmDesc g;
        g.M = m1; g.N = m1; g.K = 2 * NBT;
        g.A = Pp + j1; g.lda = p; g.tA = 1;
        g.B = Qp + j1; g.ldb = p;
        g.C = G + (long)j1 * ldg + j1; g.ldc = ldg;
        g.alpha = -1.0; g.beta = 1.0;
        gemm(c, Prec::F64, g);
      }
    }
    const int m0 = P - jt;
    const size_t tshm = ((size_t)m0 * (m0 + 1) + 2 * (size_t)m0) * sizeof(double);
    raise_lds_limit((const void*)trd_tail_lds_kernel, ((size_t)TAIL * (TAIL + 1) + 2 * TAIL) * sizeof(double));
    hipLaunchKernelGGL(trd_tail_lds_kernel, dim3(1), dim3(256), tshm, c->stream, G, ldg, P, jt, dv, ev, tauv);
    done = true;
  }
  if (!done) {
    const size_t lds_bytes = (size_t)p * sizeof(double);
    raise_lds_limit((const void*)trd_reflect_kernel, (size_t)lds_bytes);
    const int j0 = std::max(0, P - TAIL);  // columns j0.. are reduced in LDS
    if (j0 >= 1 && P <= 6000) {
      // fused path: reflector 0, then one launch per column, then the update of step j0 - 1
      double* vA = vecs + 6 * p + 512;
      double* vB = vA + p;
      double* yB = vB + p;
      double* pB = yB + p;  // >= 512 + p / 4 entries
      const size_t step_lds_max = 3 * (size_t)(P - 1) * sizeof(double);
      raise_lds_limit((const void*)trd_step_kernel, step_lds_max);
      const int m = P - 1;
      const int np0 = (m + HR - 1) / HR;
      hipLaunchKernelGGL(trd_reflect_kernel, dim3(np0), dim3(256), (size_t)m * sizeof(double), c->stream, G, ldg, P,
                         0, y, part, dv, ev, tauv, vA);
      double *cv = vA, *cy = y, *cp = part, *nv = vB, *ny = yB, *npart = pB;
      int cnp = np0;
      for (int j = 0; j + 2 <= j0; ++j) {
        const int mj = P - j - 1, mn = mj - 1;
        const int nwg = std::min(256, (mn + 3) / 4);
        hipLaunchKernelGGL(trd_step_kernel, dim3(nwg), dim3(256), 3 * (size_t)mj * sizeof(double), c->stream, G, ldg,
                           P, j, cv, cy, cp, cnp, nv, ny, npart, dv, ev, tauv);
        std::swap(cv, nv);
        std::swap(cy, ny);
        std::swap(cp, npart);
        cnp = nwg;
      }
      const int mlast = P - j0;
      const unsigned tiles = (unsigned)((mlast + 63) / 64);
      hipLaunchKernelGGL(trd_update_kernel, dim3(tiles, tiles), dim3(256), 0, c->stream, G, ldg, P, j0 - 1, cy, cp,
                         cnp, tauv);
    } else {
      for (int j = 0; j < j0; ++j) {
        const int m = P - j - 1;
        const int nparts = (m + HR - 1) / HR;
        hipLaunchKernelGGL(trd_reflect_kernel, dim3(nparts), dim3(256), (size_t)m * sizeof(double), c->stream, G,
                           ldg, P, j, y, part, dv, ev, tauv, (double*)nullptr);
        const unsigned tiles = (unsigned)((m + 63) / 64);
        hipLaunchKernelGGL(trd_update_kernel, dim3(tiles, tiles), dim3(256), 0, c->stream, G, ldg, P, j, y, part,
                           nparts, tauv);
      }
    }
    const int m0 = P - j0;
    const size_t tshm = ((size_t)m0 * (m0 + 1) + 2 * (size_t)m0) * sizeof(double);
    raise_lds_limit((const void*)trd_tail_lds_kernel, ((size_t)TAIL * (TAIL + 1) + 2 * TAIL) * sizeof(double));
    hipLaunchKernelGGL(trd_tail_lds_kernel, dim3(1), dim3(256), tshm, c->stream, G, ldg, P, j0, dv, ev, tauv);
  }
  TN_HIP(hipGetLastError());
  // d, e to the host through the pinned staging buffer
  double* hx = (double*)c->host_stage(2 * (size_t)p * sizeof(double));
  TN_HIP(hipMemcpyAsync(hx, dv, (2 * p - 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));  // d, e
  TN_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> d(hx, hx + p), e(hx + p, hx + 2 * p - 1);
  stage("tridiagonal");
  double* Q = (double*)c->workspace(WS_DC_Q, (size_t)p * p * sizeof(double));
  const long ldq = p;
  std::vector<double> D = dc_tridiag(c, p, d, e, Q, stage);
  double* tmp = (double*)c->workspace(WS_DC_TMP, (size_t)p * p * sizeof(double));

  // ---- 3. back-transformation Q <- H Q, H = H_0 ... H_{p-3}
  double* V = tmp;  // explicit reflectors (p x p)
  hipLaunchKernelGGL(trd_extract_v_kernel, dim3(grid_for(p * p)), dim3(256), 0, c->stream, G, ldg, P, tauv, V);
  double* W = (double*)c->workspace(WS_DC_W, (size_t)(NBW * p + NBW * NBW) * sizeof(double));
  double* S = W + (long)NBW * p;
  const int nrefl = P - 2;
  for (int j0 = ((nrefl - 1) / NBW) * NBW; j0 >= 0; j0 -= NBW) {
    const int nbk = std::min(NBW, nrefl - j0);
    const long rows = p - j0 - 1;  // rows touched by the block
    const double* Vs = V + (long)(j0 + 1) * p + j0;
    fill_f64(c, S, 0.0, NBW * NBW);
    {  // S = Vs^T Vs (nbk x nbk, ld NBW)
      GemmDesc g;
      g.M = nbk; g.N = nbk; g.K = rows;
      g.A = Vs; g.lda = p; g.tA = 1;
      g.B = Vs; g.ldb = p;
      g.C = S; g.ldc = NBW;
      gemm(c, Prec::F64, g);
    }
    {  // W = Vs^T Q[j0+1:, :]
      GemmDesc g;
      g.M = nbk; g.N = p; g.K = rows;
      g.A = Vs; g.lda = p; g.tA = 1;
      g.B = Q + (long)(j0 + 1) * ldq; g.ldb = ldq;
      g.C = W; g.ldc = p;
      gemm(c, Prec::F64, g);
    }
    // W <- T W
    hipLaunchKernelGGL(wy_tsolve_kernel, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, c->stream, S, tauv, j0, nbk,
                       W, p);
    {  // Q[j0+1:, :] -= Vs W
      GemmDesc g;
      g.M = rows; g.N = p; g.K = nbk;
      g.A = Vs; g.lda = p;
      g.B = W; g.ldb = p;
      g.C = Q + (long)(j0 + 1) * ldq; g.ldc = ldq;
      g.alpha = -1.0; g.beta = 1.0;
      gemm(c, Prec::F64, g);
    }
    TN_HIP(hipGetLastError());
  }
  // eigenvectors as rows: Vt[i][r] = Q[r][i]
  transpose_f64(c, Q, p, p, tmp);
  TN_HIP(hipMemcpy2DAsync(Vt, ldv * sizeof(double), tmp, p * sizeof(double), p * sizeof(double), p,
                          hipMemcpyDeviceToDevice, c->stream));
  stage("back-transform");
  w.assign(D.begin(), D.end());
  TN_HIP(hipStreamSynchronize(c->stream));
}


void sym_eig_dc_z(tn_ctx* c, double* Gd, long p, long ldg, double* Vtd, long ldv, std::vector<double>& w) {
  TN_REQUIRE(p >= 2 && p <= 16384, TN_E_INVALID, "sym_eig_dc_z: size out of range");
  double2* G = reinterpret_cast<double2*>(Gd);
  double2* Vt = reinterpret_cast<double2*>(Vtd);
  const int P = (int)p;
  EigStages stage(c, p);
  // ---- 1. Hermitian -> real tridiagonal
  double* vecs = (double*)c->workspace(WS_DC_VEC, (size_t)(12 * p + 2048) * sizeof(double));
  double* dv = vecs;
  double* ev = dv + p;
  double2* tauz = reinterpret_cast<double2*>(ev + p);  // p
  double2* yz = tauz + p;                              // p
  double2* partz = yz + p;                             // <= ceil(p / HR)
  const int j0 = std::max(0, P - TAILZ);
  {
    const size_t lds_max = (size_t)p * sizeof(double2);
    raise_lds_limit((const void*)trdz_reflect_kernel, lds_max);
  }
  for (int j = 0; j < j0; ++j) {
    const int m = P - j - 1;
    const int nparts = (m + HR - 1) / HR;
    hipLaunchKernelGGL(trdz_reflect_kernel, dim3(nparts), dim3(256), (size_t)m * sizeof(double2), c->stream, G, ldg,
                       P, j, yz, partz, dv, ev, tauz);
    const unsigned tiles = (unsigned)((m + 63) / 64);
    hipLaunchKernelGGL(trdz_update_kernel, dim3(tiles, tiles), dim3(256), 0, c->stream, G, ldg, P, j, yz, partz,
                       nparts, tauz);
  }
  {
    const int m0 = P - j0;
    const size_t tshm = ((size_t)m0 * (m0 + 1) + 2 * (size_t)m0) * sizeof(double2);
    raise_lds_limit((const void*)trdz_tail_lds_kernel, ((size_t)TAILZ * (TAILZ + 1) + 2 * TAILZ) * sizeof(double2));
    hipLaunchKernelGGL(trdz_tail_lds_kernel, dim3(1), dim3(256), tshm, c->stream, G, ldg, P, j0, dv, ev, tauz);
  }
  TN_HIP(hipGetLastError());
  double* hx = (double*)c->host_stage(2 * (size_t)p * sizeof(double));
  TN_HIP(hipMemcpyAsync(hx, dv, (2 * p - 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> d(hx, hx + p), e(hx + p, hx + 2 * p - 1);
  stage("tridiagonal");
  // ---- 2. real divide and conquer
  double* Q = (double*)c->workspace(WS_DC_Q, (size_t)p * p * sizeof(double));
  std::vector<double> D = dc_tridiag(c, p, d, e, Q, stage);
  // ---- 3. E = H_0 ... H_{p-2} Q (complex), blocks of 32 reflectors in compact WY form
  double2* V = (double2*)c->workspace(WS_DC_TMP, (size_t)p * p * sizeof(double2));
  double2* E = (double2*)c->workspace(WS_DC_UT, (size_t)p * p * sizeof(double2));
  hipLaunchKernelGGL(real_to_z_kernel, dim3(grid_for(p * p)), dim3(256), 0, c->stream, Q, p * p, E);
  hipLaunchKernelGGL(trdz_extract_v_kernel, dim3(grid_for(p * p)), dim3(256), 0, c->stream, G, ldg, P, tauz, V);
  constexpr int NBZ = 32;
  double2* W = (double2*)c->workspace(WS_DC_W, (size_t)(NBZ * p + NBZ * NBZ) * sizeof(double2));
  double2* S = W + (long)NBZ * p;
  const int nrefl = P - 1;
  for (int jb = ((nrefl - 1) / NBZ) * NBZ; jb >= 0; jb -= NBZ) {
    const int nbk = std::min(NBZ, nrefl - jb);
    const long rows = p - jb - 1;
    const double2* Vs = V + (long)(jb + 1) * p + jb;
    fill_f64(c, (double*)S, 0.0, 2 * NBZ * NBZ);
    {  // S = Vs^H Vs
      GemmDesc g;
      g.M = nbk; g.N = nbk; g.K = rows;
      g.A = Vs; g.lda = p; g.tA = 1; g.cA = 1;
      g.B = Vs; g.ldb = p;
      g.C = S; g.ldc = NBZ;
      gemm(c, Prec::Z64, g);
    }
    {  // W = Vs^H E[jb+1:, :]
      GemmDesc g;
      g.M = nbk; g.N = p; g.K = rows;
      g.A = Vs; g.lda = p; g.tA = 1; g.cA = 1;
      g.B = E + (long)(jb + 1) * p; g.ldb = p;
      g.C = W; g.ldc = p;
      gemm(c, Prec::Z64, g);
    }
    hipLaunchKernelGGL(wyz_tsolve_kernel, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, c->stream, S, tauz, jb,
                       nbk, W, p);
    {  // E[jb+1:, :] -= Vs W
      GemmDesc g;
      g.M = rows; g.N = p; g.K = nbk;
      g.A = Vs; g.lda = p;
      g.B = W; g.ldb = p;
      g.C = E + (long)(jb + 1) * p; g.ldc = p;
      g.alpha = -1.0; g.beta = 1.0;
      gemm(c, Prec::Z64, g);
    }
    TN_HIP(hipGetLastError());
  }
  // eigenvectors as conjugated rows: Vt = E^H
  hipLaunchKernelGGL(adjoint_z_kernel, dim3((unsigned)((p + 31) / 32), (unsigned)((p + 31) / 32)), dim3(256), 0,
                     c->stream, E, P, Vt, ldv);
  TN_HIP(hipGetLastError());
  stage("back-transform");
  w.assign(D.begin(), D.end());
  TN_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace tn
