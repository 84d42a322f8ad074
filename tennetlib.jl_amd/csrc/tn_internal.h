// Internal declarations shared by the libtnhip sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/tnhip.h"

namespace tn {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TN_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::tn::Error(TN_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TN_REQUIRE(cond, code, msg)                 \
  do {                                              \
    if (!(cond)) throw ::tn::Error((code), (msg));  \
  } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  // Kernels queued on any stream may still use the buffer: wait for them
  // before the memory goes back to the allocator, where another context
  // (another stream, another host thread) could receive it at once.
  void release() {
    if (p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  // grow-only; contents are NOT preserved.  A new buffer starts zeroed:
  // padding and channels that no kernel writes (and that are only ever
  // multiplied by zero) then hold zeros, not whatever the allocator returns
  // (memory another context freed, which may encode NaN).  The fill is
  // complete before the buffer is returned; allocation happens outside any
  // captured region.
  void ensure(size_t nbytes) {
    if (nbytes <= bytes && p) return;
    release();
    size_t nb = nbytes ? nbytes : 256;
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(TN_E_NOMEM, "hipMalloc(" + std::to_string(nb) + ") failed: " + hipGetErrorString(e));
    }
    e = hipMemset(p, 0, nb);
    // the fill runs on the null stream, which does not order with the
    // library's non-blocking streams: wait for it before any kernel writes
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      (void)hipFree(p);
      p = nullptr;
      throw Error(TN_E_HIP, std::string("hipMemset of a new buffer failed: ") + hipGetErrorString(e));
    }
    bytes = nb;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// timed phases; PH_MV_GEMM brackets each GEMM launch of a matvec on its own
// (nested inside PH_MATVEC), the others do not nest
// PH_TRD_SYMV brackets each launch of the tridiagonalisation's product
// kernel (nested inside PH_SVD) and counts its algorithmic bytes
enum Phase { PH_MATVEC = 0, PH_ENV = 1, PH_KRYLOV = 2, PH_SVD = 3, PH_MV_GEMM = 4, PH_TRD_SYMV = 5, PH_COUNT = 6 };

}  // namespace tn

struct tn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  // named workspace slots (grow-only)
  std::vector<tn::DevBuf> ws;
  // host-pinned staging: `pinned` (PINNED_N doubles) for the Krylov loop's
  // scalars, `hstage` (grow-only) for the eigensolver's per-level tables
  static constexpr size_t PINNED_N = 4096;
  double* pinned = nullptr;
  double* pinned_dev = nullptr;  // the same memory as seen from kernels
  // completion events of the Krylov steps in flight (two slots)
  hipEvent_t kry_ev[2] = {nullptr, nullptr};
  void* hstage = nullptr;
  size_t hstage_bytes = 0;
  void* host_stage(size_t bytes) {
    if (hstage && bytes <= hstage_bytes) return hstage;
    if (hstage) {
      (void)hipStreamSynchronize(stream);
      (void)hipHostFree(hstage);
      hstage = nullptr;
      hstage_bytes = 0;
    }
    const size_t nb = bytes > ((size_t)1 << 20) ? bytes : ((size_t)1 << 20);
    if (hipHostMalloc(&hstage, nb, hipHostMallocDefault) != hipSuccess) {
      hstage = nullptr;
      throw tn::Error(TN_E_NOMEM, "hipHostMalloc failed for the staging buffer");
    }
    hstage_bytes = nb;
    return hstage;
  }
  // per-phase timing: phases whose bit is set in timing_mask
  bool timing = false;
  unsigned timing_mask = 0;
  double phase_ms[tn::PH_COUNT] = {0, 0, 0, 0, 0, 0};
  long phase_n[tn::PH_COUNT] = {0, 0, 0, 0, 0, 0};
  double phase_bytes[tn::PH_COUNT] = {0, 0, 0, 0, 0, 0};  // algorithmic bytes of timed launches
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<int> ev_phase;
  size_t ev_used = 0;
  int open_timers = 0;
  const double* ones = nullptr;  // device scalar 1.0 (edge environments)
  tn::DevBuf ones_buf;
  // device address of this device's fixed scratch array (allocation-free
  // product), taken once per device at context creation
  void* scratch_dev = nullptr;

  void* workspace(int slot, size_t bytes) {
    if ((int)ws.size() <= slot) ws.resize(slot + 1);
    ws[slot].ensure(bytes);
    return ws[slot].p;
  }
};

namespace tn {

// ---------------------------------------------------------------------------
// wavefront sums in registers
// ---------------------------------------------------------------------------
// x plus the value of the lane that the DPP control CTRL selects (row-local
// permutations; every source lane must be active)
template <int CTRL>
__device__ __forceinline__ double dpp_add(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return x + __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

__device__ __forceinline__ double read_lane(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// sum over the 64 lanes of a wavefront without LDS round trips: lane pairs
// xor 1 and xor 2 inside each quad, then the half-row and row mirrors (both
// partners hold equal partial sums, so every lane of a row ends with the same
// bits), then the four row sums read from lanes 0, 16, 32, 48.  Every lane
// gets the same value; every lane of the wavefront must be active.
__device__ __forceinline__ double wave_sum64(double x) {
  x = dpp_add<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x = dpp_add<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x = dpp_add<0x141>(x);  // row_half_mirror
  x = dpp_add<0x140>(x);  // row_mirror
  return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// workspace slot ids (each slot is an independent grow-only buffer)
enum WsSlot {
  WS_T1 = 0, WS_T2, WS_PSI32, WS_OUT32, WS_KRY_BASIS, WS_KRY_W, WS_KRY_PART, WS_KRY_COEF,
  WS_SVD_X, WS_SVD_Q, WS_SVD_GRAM, WS_SVD_J, WS_SVD_CNT, WS_SVD_PAIRS, WS_SVD_S, WS_SVD_PERM,
  WS_SVD_SCALE, WS_PHI, WS_MISC, WS_ENV_X, WS_ENV_Y, WS_KRY_TMP,
  WS_DC_VEC, WS_DC_Q, WS_DC_SMALL, WS_DC_TMP, WS_DC_UT, WS_DC_LEVEL, WS_DC_W, WS_PROJ_X, WS_PROJ_Y, WS_PROJ_C,
  WS_QR_X, WS_QR_V, WS_QR_Q, WS_QR_R, WS_QR_S, WS_QR_W, WS_TTN_A, WS_TTN_S, WS_TTN_G, WS_TTN_P, WS_TTN_F,
  WS_GEMM_SPLIT, WS_BOND, WS_PHI2, WS_TTNP_0, WS_TTNP_1, WS_TTNP_2, WS_TTNP_3, WS_DC_PANEL,
  WS_QN_M, WS_QN_V, WS_QN_W, WS_QN_IDX, WS_QN_T, WS_QN_RHO, WS_QN_PLAN, WS_SLOTS
};

// scoped phase timer (HIP events on the context stream)
struct PhaseTimer {
  tn_ctx* c;
  int ph;
  size_t idx;
  PhaseTimer(tn_ctx* ctx, int phase);
  ~PhaseTimer();
};
void flush_timers(tn_ctx* c);

// ---------------------------------------------------------------------------
// GEMM (MFMA) -- gemm.hip
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], row-major, batch over b.
// op(A) is M x K: element (m,k) = A[m*lda + k] (tA = 0) or A[k*lda + m] (tA = 1).
// op(B) is K x N: element (k,n) = B[k*ldb + n] (tB = 0) or B[n*ldb + k] (tB = 1).
// ---------------------------------------------------------------------------
enum class Prec { F64 = 0, F32 = 1, BF16X3 = 2, Z64 = 3 };

struct GemmDesc {
  long M = 0, N = 0, K = 0;
  const void* A = nullptr; long lda = 0; long sA = 0; int tA = 0; int cA = 0;
  const void* B = nullptr; long ldb = 0; long sB = 0; int tB = 0; int cB = 0;
  void* C = nullptr; long ldc = 0; long sC = 0;
  int batch = 1;
  double alpha = 1.0, beta = 0.0;
  // split-K partial products: a caller-owned device buffer of split_cap bytes
  // (see gemm_split_bytes) used instead of the context workspace when set
  void* split_ws = nullptr;
  size_t split_cap = 0;
  // zpA / zpB: the rows of A (not transposed) / B (not transposed) hold finite
  // values from their valid width (K / N) up to the next 16-byte boundary, so
  // the fast kernel may read whole vectors there (A's extra columns meet B's
  // rows past K, which load as zeros; B's extra columns land in discarded
  // output columns)
  int zpA = 0, zpB = 0;
};
void gemm(tn_ctx* c, Prec p, const GemmDesc& g);
// bytes of split-K partial products gemm() needs for g (0: no split)
size_t gemm_split_bytes(Prec p, const GemmDesc& g);
// FP64 grouped GEMM: for every output group g < ngroups and batch z < batch,
// C_g[z] (M x N, row stride ldc) = sum over terms t in [off[g], off[g+1]) of
// w[t] * A_{ia[t]}[z] (M x K, lda) * B_{ib[t]}[z] (K x N, ldb), with
// A_t = A + ia[t] tsA + z sA, B_t = B + ib[t] tsB + z sB, C_g = C + g gsC + z sC
// (device term arrays); a group without terms is written as zeros.
struct GroupedDesc {
  long M = 0, N = 0, K = 0;
  const double* A = nullptr; long lda = 0, sA = 0, tsA = 0;
  const double* B = nullptr; long ldb = 0, sB = 0, tsB = 0;
  double* C = nullptr; long ldc = 0, sC = 0, gsC = 0;
  const int* off = nullptr; const int* ia = nullptr; const int* ib = nullptr; const double* w = nullptr;
  int ngroups = 0, batch = 1;
};
void gemm_grouped_f64(tn_ctx* c, const GroupedDesc& d);
// Block-sparse GEMM over uneven blocks (gemm.hip): task i writes the block
// C + c_off (M x N, ld ldc) = alpha * sum over its segments s of
// (A + a_off_s)(M x K_s, ld lda) * (B + b_off_s)(K_s x N, ld ldb); a task
// without segments writes zeros.  tiles (device) lists {task, m0, n0, 0} of
// the BS_TILE x BS_TILE output tiles in XCD-major order: entry (i % 8) *
// (ntiles / 8) + i / 8 runs as workgroup i; ntiles is a multiple of 8 and
// padding entries have task -1.  Offsets are in elements of the precision.
constexpr int BS_TILE = 64;
struct BsTask {
  long c_off;
  int M, N, seg0, nseg;
};
struct BsSeg {
  long a_off, b_off, K;
};
struct BsDesc {
  const void* A = nullptr; long lda = 0;
  const void* B = nullptr; long ldb = 0;
  int tB = 0;  // 1: segment operand B_s is stored transposed (N x K_s, ld ldb)
  void* C = nullptr; long ldc = 0;
  const BsTask* tasks = nullptr;
  const BsSeg* segs = nullptr;
  const int4* tiles = nullptr;
  long ntiles = 0;
  double alpha = 1.0;
};
void gemm_blocksparse(tn_ctx* c, Prec p, const BsDesc& d);
// double-buffered FP64/FP32 kernels (gemm_fast.hip); gemm() uses them when
// gemm_fast_ok() holds (shapes >= one tile, 16-byte aligned operand rows)
bool gemm_fast_ok(Prec p, const GemmDesc& g);
void gemm_fast(tn_ctx* c, Prec p, const GemmDesc& g);
// C (M x N, float) = A * B^T with A (M x K) and B (N x K) as bf16 hi/lo
// pairs, rows padded to multiples of 8 elements
void gemm_bf16x3_nt(tn_ctx* c, long M, long N, long K, const unsigned short* Ah, const unsigned short* Al,
                    long lda, const unsigned short* Bh, const unsigned short* Bl, long ldb, float* C, long ldc);
// bf16 hi/lo split of a rows x cols matrix (row stride lds, FP64 or FP32);
// transpose = false: dst is rows x ldd; true: dst is cols x ldd (ldd >= rows);
// padding columns are zero
void split_bf16(tn_ctx* c, const void* src, bool src_f64, long rows, long cols, long lds, bool transpose,
                unsigned short* hi, unsigned short* lo, long ldd);

// ---------------------------------------------------------------------------
// small-MPO-operator contraction ("W step") -- tensor_ops.hip
// out[g', A, S', B] = sum_{g,S} Op[(g' S'), (g S)] * in[g, A, S, B]
// Op is given as CSR over output rows o = g'*Dp + S' with columns i = g*D + S.
// Element type: double or float; complex data is handled by the caller as
// interleaved real pairs (B doubled, strides doubled).
// ---------------------------------------------------------------------------
struct SparseOp {
  int rows = 0, cols = 0, nnz = 0;
  DevBuf rowptr, colidx, vals;  // int, int, double
  // row-major dense copies (FP64, FP32) for the MFMA W step, kept when the
  // operator has at least 8 entries per row and at most 128 rows and columns
  bool dense = false;
  DevBuf d64, d32;
};
struct WStepDesc {
  int G = 0, Gp = 0, D = 0, Dp = 0;
  long nA = 0, nB = 0;
  const void* in = nullptr; long in_sg = 0, in_sA = 0, in_sS = 0;
  void* out = nullptr; long out_sg = 0, out_sA = 0, out_sS = 0;
  const SparseOp* op = nullptr;
};
void wstep(tn_ctx* c, bool fp32, const WStepDesc& w);
// W step from FP32 input writing the result as bf16 hi/lo pairs at the
// positions w.out_* would address (w.out is ignored)
void wstep_split(tn_ctx* c, const WStepDesc& w, unsigned short* hi, unsigned short* lo);
void build_sparse_op(SparseOp& op, int rows, int cols, const std::vector<double>& dense);

// ---------------------------------------------------------------------------
// elementwise / Krylov vector kernels -- tensor_ops.hip
// ---------------------------------------------------------------------------
void cvt_f64_to_f32(tn_ctx* c, const double* x, float* y, long n);
void cvt_f32_to_f64(tn_ctx* c, const float* x, double* y, long n);
// rows x cols FP64 block (row stride lds) to FP32 with row stride ldd
// (columns cols .. ldd - 1 of the destination set to zero)
void cvt_f64_to_f32_2d(tn_ctx* c, const double* x, long rows, long cols, long lds, float* y, long ldd);
// columns c0 .. c1 - 1 of a rows x ld FP32 matrix set to zero
void zero_cols_f32(tn_ctx* c, float* y, long rows, long c0, long c1, long ld);
// FP64: rows x cols block to row stride ldd with zero padding; zero columns
void pad_rows_f64(tn_ctx* c, const double* x, long rows, long cols, long lds, double* y, long ldd);
void zero_cols_f64(tn_ctx* c, double* y, long rows, long c0, long c1, long ld);
// y[r * ld + j] = 0 for r < rows, c0 <= j < c1 (bf16 hi/lo planes)
void zero_cols_u16(tn_ctx* c, unsigned short* y, long rows, long c0, long c1, long ld);
void fill_f64(tn_ctx* c, double* x, double v, long n);
// FP32 layouts of interleaved complex FP64 matrices (rows x cols complex):
// split rows: dst (rows x 2 cols) = [re | im]; block: dst (2 rows x 2 cols) =
// [[re, im], [-im, re]]; join rows: interleaved dst from [re | im] rows
void cplx_split_rows_f32(tn_ctx* c, const double* src, long rows, long cols, float* dst);
void cplx_block_f32(tn_ctx* c, const double* src, long rows, long cols, float* dst);
void cplx_join_rows_f64(tn_ctx* c, const float* src, long rows, long cols, double* dst);
void copy_f64(tn_ctx* c, double* dst, const double* src, long n);
// out[i] = <basis_i, x> for i < count (basis vectors contiguous with stride ld);
// cplx: complex128 data of n complex elements, out holds (re, im) pairs.
// out_host (optional): a device-side address of mapped pinned host memory
// that receives the same values, readable on the host after the stream has
// been synchronised (no separate copy).
void multi_dot(tn_ctx* c, const double* basis, long ld, int count, const double* x, long n,
               bool cplx, double* out_dev, double* out_host = nullptr);
// x -= sum_i coef[i] * basis_i ; coef on device (real or complex pairs)
void multi_axpy_sub(tn_ctx* c, const double* basis, long ld, int count, const double* coef_dev,
                    double* x, long n, bool cplx);
// out = sum_i coef[i] * basis_i ; coef on device
void lincomb(tn_ctx* c, const double* basis, long ld, int count, const double* coef_dev,
             double* out, long n, bool cplx);
// x *= s (s real, on host)
void scale_f64(tn_ctx* c, double* x, double s, long n);
// y = a*x (a real), n doubles
void scaled_copy(tn_ctx* c, double* y, const double* x, double a, long n);
// one Krylov orthogonalisation step (two passes against K basis vectors, then
// |w|^2) in a single launch for short vectors; coefficients at coef_dev[0..),
// [128..), |w|^2 at [256], mirrored to hout.  Returns false (nothing queued)
// when n or K is outside its range
bool krylov_ortho_small(tn_ctx* c, const double* basis, long ld, int K, double* w, long n, bool cplx,
                        double* coef_dev, double* hout);
// y = x / sqrt(max(0, *r2_dev)) (r2_dev: a device scalar)
void scale_rnorm(tn_ctx* c, double* y, const double* x, const double* r2_dev, long n);
// row sum of squares of a (rows x L) row-major matrix into out[rows]
void row_norms2(tn_ctx* c, const double* X, long rows, long L, long ld, double* out);
// y += scale * coef_dev[0] * x (complex coefficient pair when cplx)
void add_scaled_dev(tn_ctx* c, double* y, const double* x, const double* coef_dev, double scale, long n, bool cplx);
// out (rows x cols) = in^T  (in: cols x rows), doubles
void transpose_f64(tn_ctx* c, const double* in, long in_rows, long in_cols, double* out);
// out[k][j] = src[perm[k]][j] * scale[k]   (k < kcount, j < L)
void gather_rows(tn_ctx* c, const double* src, long ld_src, const int* perm, const double* scale,
                 long kcount, long L, double* out);
// out[j][k] = src[perm[k]][j] * scale[k]   (transposed gather)
void gather_rows_T(tn_ctx* c, const double* src, long ld_src, const int* perm,
                   const double* scale, long kcount, long L, double* out);

// ---------------------------------------------------------------------------
// density-matrix eigen-decomposition (two-sided block Jacobi) -- svd.hip
// ---------------------------------------------------------------------------
struct GramEig {
  long p = 0, ppad = 0;     // Gram size (m if left, n otherwise), padded to 64
  bool left = true;         // G = M M^H (left) or M^H M (right)
  bool cplx = false;        // complex128 data (interleaved)
  void* Vt = nullptr;       // ppad x ppad workspace: V^T (real) / V^H (complex), first p columns
  std::vector<double> w;    // eigenvalues (unsorted, length p) = weights s^2
  int sweeps = 0;
};
// G = M M^H (left) or M^H M (right), plus xs * X X^H (left, X: p x kx with
// row stride ldx, default kx) or xs * X^H X (right, X: kx x p) when X is
// given, then its eigen-decomposition
void gram_eig(tn_ctx* c, const void* M, long m, long n, bool left, bool cplx, GramEig& out,
              const void* X = nullptr, long kx = 0, double xs = 0.0, long ldx = 0);
// eigen-decomposition of a Hermitian p x p device matrix H (row stride ldh,
// left unchanged): out.w the eigenvalues, out.Vt rows the conjugated
// eigenvectors (ld out.ppad), as gram_eig gives them
void herm_eig(tn_ctx* c, const void* H, long p, long ldh, bool cplx, GramEig& out);
// eigen-decomposition of a symmetric / Hermitian p x p device matrix G (ld
// ldg, padding rows/columns zero); Vt (ld = padded size) receives V^T / V^H
void sym_eig_jacobi(tn_ctx* c, void* G, long p, long ldg, void* Vt, long ldv, bool cplx, std::vector<double>& w,
                    int* sweeps_out);
// real symmetric p x p (3 <= p <= 16384) by Householder tridiagonalisation and
// divide and conquer (eig_dc.hip); same contract as sym_eig_jacobi (G is
// overwritten, Vt rows are the eigenvectors, w their eigenvalues)
// One library call at a time per process: every C entry point holds this
// lock for its duration (calls on different contexts from different host
// threads are serialised; their GPU work still overlaps where it is queued
// asynchronously).
std::recursive_mutex& api_mutex();
// Raise a kernel's dynamic-LDS limit to at least `bytes` (above the 64 KiB
// default).  Serialised and applied once per kernel and size: changing the
// attribute while another host thread launches the same kernel is not safe.
void raise_lds_limit(const void* kernel, size_t bytes);
// compact-WY helper: W (nb x n, ld n) <- T W for reflectors j0..j0+nb-1 with
// T^{-1} = diag(1/tau) + striu(S), S = V^T V (nb x nb, ld 64), nb <= 64
constexpr int WY_NB = 64;
void wy_tsolve(tn_ctx* c, const double* S, const double* tauv, int j0, int nb, double* W, long n);
// Householder QR (LAPACK dgeqrf convention: beta = -sign(alpha)|x|,
// tau = (beta - alpha)/beta, v_0 = 1) of M (m x n) given column by column:
// X (n x m, row c = column c of M) is overwritten.  With k = min(m, n):
// Qt (k x m) receives the rows of Q^T (orthonormal), R (k x n) the upper
// trapezoidal factor, M = Q R.
void householder_qr(tn_ctx* c, double* X, long m, long n, double* Qt, double* R);
// permute a row-major 3-leg tensor A[d0][d1][d2] to out[d1][d0][d2]
void swap01_f64(tn_ctx* c, const double* A, long d0, long d1, long d2, double* out);
// out = A concatenated with P along axis ax of a 3-leg tensor (dims of A;
// P has dims with axis ax replaced by p)
void concat3_f64(tn_ctx* c, const double* A, const long* dims, int ax, const double* P, long p, double* out);

// Lanczos ground state (eigsolve(A, x0, 1, :SR), src/base/solver.jl:36) of the
// symmetric operator y = A x given by `apply`; phi (n elements, device) is
// the start vector and receives the Ritz vector; returns the Ritz value
double krylov_eig_op(tn_ctx* c, bool cx, long n, const std::function<void(const double*, double*)>& apply,
                     double* phi, int kdim, int maxiter, double tol, tn_krylov_info* info);
void sym_eig_dc(tn_ctx* c, double* G, long p, long ldg, double* Vt, long ldv, std::vector<double>& w);
// Hermitian p x p complex128 (interleaved, 2 <= p <= 16384): complex Householder
// reduction to a real tridiagonal, the same divide and conquer, complex
// back-transformation; Vt rows receive the conjugated eigenvectors
void sym_eig_dc_z(tn_ctx* c, double* G, long p, long ldg, double* Vt, long ldv, std::vector<double>& w);
// complex: out[j][k] = conj?(src[perm[k]][j]) * scale[k]  (src rows of complex)
void gather_rows_T_z(tn_ctx* c, const double* src, long ld_src, const int* perm, const double* scale,
                     long kcount, long L, double* out, bool conj);

// QN blocks of the factorisation (qn.hip).  qn_block_eig: nblk density-matrix
// blocks, each at most pmax <= QN_JACOBI_MAX rows, stacked in R (ld
// QN_RHO_LD); blk_dev[4b..4b+3] = {row offset in R, size p, offset of the p
// eigenvalues in w, offset of the p eigenvectors (rows of length p) in V}.
constexpr int QN_JACOBI_MAX = 96;
constexpr int QN_RHO_LD = 96;
void qn_block_eig(tn_ctx* c, const double* R, const int* blk_dev, int nblk, int pmax, double* w, double* V);
// qn_scatter: column j of out (rows x k, ld ldo; its transpose when trans)
// receives eigenvector col[4j+2] (stored at V + col[4j] as rows of length
// col[4j+3]) on the rows rows[col[4j+1] ..]; out must be zero elsewhere
void qn_scatter(tn_ctx* c, const double* V, const int* col_dev, const int* rows_dev, int k, long ldo, bool trans,
                double* out);

// truncation rule of NDTensors truncate!!
struct TruncRule {
  long n;
  double truncerr;
  double docut;
};
TruncRule truncate_weights(const std::vector<double>& P_desc, long maxdim, long mindim,
                           double cutoff);

}  // namespace tn
