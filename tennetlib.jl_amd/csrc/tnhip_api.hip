// C ABI of libtnhip: StateEnvs{ProjMPO} on one MI355X.
//
// Storage in HBM (row-major, one buffer per tensor, FP64 or complex128):
//   site k        A_k[a, s, b]               (chi_{k-1}, d_k, chi_k)
//   left env k    Lt_k[w, a'(bra), a(ket)]   (w_k, chi_k, chi_k)       sites 1..k
//   right env k   Rt_k[w, b(ket), b'(bra)]   (w_{k-1}, chi_{k-1}, ...) sites k..N
// With these layouts every large contraction is a plain (batched) GEMM:
//   matvec (two-site, pos p, D = d_p d_{p+1}):
//     T1[(w a'), (S b)]   = Lt[(w a'), a] * phi[a, (S b)]            GEMM  M=w chil, N=D chir, K=chil
//     T2[a', S', w2, b]   = sum_{w,S} Op2[(w2 S'), (w S)] T1          W step (sparse MPO pair)
//     out[(a' S'), b']    = T2[(a' S'), (w2 b)] * Rt[(w2 b), b']      GEMM  M=chil D, N=chir, K=w chir
//   left env update (site k):  X = Lt*A ; Y = W-step ; Lt'[w'] = A^H * Y[w']   (batched over w')
//   right env update (site k): X[w2] = B*Rt[w2] ; Y = W-step ; Rt'[w] = Y[w] * B^H (batched over w)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <array>
#include <map>
#include <numeric>

#include "tn_internal.h"

using namespace tn;
using cplx = std::complex<double>;

namespace {
thread_local std::string g_thread_error;

int set_error(tn_ctx* c, int code, const std::string& msg) {
  g_thread_error = msg;
  if (c) c->last_error = msg;
  return code;
}

template <class F>
int guarded(tn_ctx* c, F&& f) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  try {
    f();
    return TN_OK;
  } catch (const tn::Error& e) {
    return set_error(c, e.code, e.what());
  } catch (const std::bad_alloc&) {
    return set_error(c, TN_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(c, TN_E_INVALID, e.what());
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// phase timers
// ---------------------------------------------------------------------------
namespace tn {
PhaseTimer::PhaseTimer(tn_ctx* ctx, int phase) : c(ctx), ph(phase), idx((size_t)-1) {
  if (!c->timing || !((c->timing_mask >> phase) & 1u)) return;
  // flush only when no timer is open: an enclosing timer's stop event is
  // not recorded yet and must keep its slot
  if (c->ev_used >= 4096 && c->open_timers == 0) flush_timers(c);
  if (c->ev_used >= c->ev_pool.size()) {
    hipEvent_t a, b;
    TN_HIP(hipEventCreate(&a));
    TN_HIP(hipEventCreate(&b));
    c->ev_pool.push_back({a, b});
    c->ev_phase.push_back(0);
  }
  idx = c->ev_used++;
  c->ev_phase[idx] = ph;
  c->open_timers++;
  TN_HIP(hipEventRecord(c->ev_pool[idx].first, c->stream));
}
PhaseTimer::~PhaseTimer() {
  if (idx == (size_t)-1) return;
  c->open_timers--;
  if (hipEventRecord(c->ev_pool[idx].second, c->stream) != hipSuccess) (void)hipGetLastError();
}
void flush_timers(tn_ctx* c) {
  if (c->ev_used == 0) return;
  TN_HIP(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < c->ev_used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_pool[i].first, c->ev_pool[i].second) == hipSuccess) {
      c->phase_ms[c->ev_phase[i]] += ms;
      c->phase_n[c->ev_phase[i]] += 1;
    } else {
      (void)hipGetLastError();  // do not leave a sticky error for the next launch check
    }
  }
  c->ev_used = 0;
}
}  // namespace tn

// ---------------------------------------------------------------------------
// environment object
// ---------------------------------------------------------------------------
struct SiteT {
  DevBuf buf;
  long chil = 0, chir = 0;
};
struct EnvT {
  DevBuf buf;
  long chi = 0;
  int w = 0;
};

// ProjMPS2 (src/mps/projmps2.jl:34-224): an MPS M held constant, its overlap
// environments with psi and the local projection m of M at the current
// position; the projected operator adds weight * m (m^H v) to H v
// (ProjMPO_MPS2, src/mps/projmpo_mps2.jl:94-101).
struct ProjT {
  std::vector<SiteT> M;   // site k at [k-1]
  std::vector<EnvT> LR;   // LR[k]: overlap environment (chi_psi x chi_M) of sites 1..k (left) or k..N (right)
  std::vector<long> chiM; // bond of M of the environment LR[k]
  int lpos = 0, rpos = 0;
  DevBuf m;               // local projection at position m_pos for m_nsite sites
  bool m_ok = false;
  int m_pos = 0, m_nsite = 0;
};

// Block-sparse two-site product of a QN-conserving environment at one
// position: the task lists of the two GEMMs (uploaded once per position and
// link sectors), see qn_plan.
struct QnMvPlan {
  long key[4] = {-1, -1, -1, -1};  // pos, versions of links pos-1 and pos+1, MPO version
  bool ok = false;                 // false: links not sector-contiguous, use the dense product
  std::vector<BsTask> t1, t3;
  std::vector<BsSeg> s1, s3;
  std::vector<int4> tl1, tl3;
  DevBuf dev;
  const BsTask* dt1 = nullptr;
  const BsTask* dt3 = nullptr;
  const BsSeg* ds1 = nullptr;
  const BsSeg* ds3 = nullptr;
  const int4* dtl1 = nullptr;
  const int4* dtl3 = nullptr;
  double flops = 0.0;  // block-sparse GEMM flops of one product
};

struct tn_env {
  tn_ctx* ctx = nullptr;
  int N = 0;
  std::vector<int> d, w;  // d[k-1] for site k; w[k] for MPO bond k (0..N)
  int dtype = TN_F64;
  size_t esz = 8;
  std::vector<std::vector<double>> Wh;  // host MPO copies, site k at [k-1]
  std::vector<char> W_set;
  std::vector<SiteT> A;                 // site k at [k-1]
  std::vector<EnvT> LR;                 // 0..N+1
  int lpos = 0, rpos = 0;
  int nsite = 2;
  int ortho = 0;
  std::vector<SparseOp> opL, opR, op2;
  std::vector<char> opL_ok, opR_ok, op2_ok;
  // FP32 shadows of the environments used by mixed-precision matvecs
  DevBuf L32, R32;
  long shadow_key_l = -1, shadow_key_r = -1;
  // FP64 copies with even row strides of environments of odd bond dimension
  DevBuf L64, R64;
  long pad64_key_l = -1, pad64_key_r = -1;
  // bf16 hi/lo shadows (BF16x3 matvec)
  DevBuf Lh, Ll, Rh, Rl;
  long split_key_l = -1, split_key_r = -1;
  long env_version = 0;
  std::vector<long> LR_version;
  // penalty projectors (excited states): H v + weight sum_k m_k (m_k^H v)
  std::vector<ProjT> proj;
  double proj_weight = 0.0;
  // channels of bond k left out of the noise factor: the first noise_skip_l[k]
  // (identity on sites <= k) for the left factor, the last noise_skip_r[k]
  // (identity on sites > k) for the right factor
  std::vector<int> noise_skip_l, noise_skip_r;
  // quantum numbers (U(1) or U(1)xU(1), nq charges; nq = 0: none): per site
  // the QN of each basis state, per link 0..N the QN of each link element
  // (link k carries the QN of sites 1..k)
  int nq = 0;
  std::vector<std::vector<std::array<int, 2>>> site_qn, link_qn;
  std::vector<long> link_ver;  // bumped whenever link_qn[k] changes
  long mpo_ver = 0;            // bumped by every set_mpo
  // QN flux of every MPO link element: bra QN - ket QN of the environments
  // on that channel (mpo_flux[k][w], k = 0..N); flux_known[k][w] = 0 for
  // channels no nonzero path from the left edge reaches
  long flux_ver = -1;
  std::vector<std::vector<std::array<int, 2>>> mpo_flux;
  std::vector<std::vector<char>> flux_known;
  QnMvPlan qplan;

  bool cplx() const { return dtype == TN_C128; }
};

namespace {

tn_ctx* C(tn_env* e) { return e ? e->ctx : nullptr; }

void check_env(tn_env* e) { TN_REQUIRE(e != nullptr, TN_E_INVALID, "null environment"); }

void require_mpo(tn_env* e) {
  for (int k = 0; k < e->N; ++k)
    TN_REQUIRE(e->W_set[k], TN_E_INVALID, "MPO tensor of site " + std::to_string(k + 1) + " not set");
}
void require_sites(tn_env* e) {
  for (int k = 0; k < e->N; ++k)
    TN_REQUIRE(e->A[k].buf.p != nullptr, TN_E_INVALID, "MPS tensor of site " + std::to_string(k + 1) + " not set");
}

// edge environment: a 1x1x1 tensor holding 1
void set_edge(tn_env* e, int k) {
  EnvT& E = e->LR[k];
  E.buf.ensure(16);
  E.chi = 1;
  E.w = 1;
  double one[2] = {1.0, 0.0};
  TN_HIP(hipMemcpyAsync(E.buf.p, one, e->esz, hipMemcpyHostToDevice, e->ctx->stream));
  TN_HIP(hipStreamSynchronize(e->ctx->stream));
}

// sparse W-step operators built from the host MPO copies
const SparseOp& op_left(tn_env* e, int k) {  // Op[(w' s'), (w s)] = W_k[w, s', s, w']
  if (!e->opL_ok[k - 1]) {
    const int wl = e->w[k - 1], wr = e->w[k], d = e->d[k - 1];
    const auto& W = e->Wh[k - 1];
    std::vector<double> M((size_t)wr * d * wl * d, 0.0);
    for (int a = 0; a < wl; ++a)
      for (int sp = 0; sp < d; ++sp)
        for (int s = 0; s < d; ++s)
          for (int b = 0; b < wr; ++b)
            M[(size_t)(b * d + sp) * (wl * d) + (a * d + s)] = W[((size_t)(a * d + sp) * d + s) * wr + b];
    build_sparse_op(e->opL[k - 1], wr * d, wl * d, M);
    e->opL_ok[k - 1] = 1;
  }
  return e->opL[k - 1];
}
const SparseOp& op_right(tn_env* e, int k) {  // Op[(w t'), (w2 t)] = W_k[w, t', t, w2]
  if (!e->opR_ok[k - 1]) {
    const int wl = e->w[k - 1], wr = e->w[k], d = e->d[k - 1];
    const auto& W = e->Wh[k - 1];
    std::vector<double> M((size_t)wl * d * wr * d, 0.0);
    for (int a = 0; a < wl; ++a)
      for (int tp = 0; tp < d; ++tp)
        for (int t = 0; t < d; ++t)
          for (int b = 0; b < wr; ++b)
            M[(size_t)(a * d + tp) * (wr * d) + (b * d + t)] = W[((size_t)(a * d + tp) * d + t) * wr + b];
    build_sparse_op(e->opR[k - 1], wl * d, wr * d, M);
    e->opR_ok[k - 1] = 1;
  }
  return e->opR[k - 1];
}
// fused pair (p, p+1): Op2[(w2, s', t'), (w, s, t)] = sum_w1 W_p[w,s',s,w1] W_{p+1}[w1,t',t,w2]
const SparseOp& op_two(tn_env* e, int p) {
  if (!e->op2_ok[p - 1]) {
    const int w0 = e->w[p - 1], w1 = e->w[p], w2 = e->w[p + 1];
    const int d1 = e->d[p - 1], d2 = e->d[p];
    const int D = d1 * d2;
    const auto& Wa = e->Wh[p - 1];
    const auto& Wb = e->Wh[p];
    std::vector<double> M((size_t)w2 * D * w0 * D, 0.0);
    for (int a = 0; a < w0; ++a)
      for (int sp = 0; sp < d1; ++sp)
        for (int s = 0; s < d1; ++s)
          for (int m = 0; m < w1; ++m) {
            const double x = Wa[((size_t)(a * d1 + sp) * d1 + s) * w1 + m];
            if (x == 0.0) continue;
            for (int tp = 0; tp < d2; ++tp)
              for (int t = 0; t < d2; ++t)
                for (int b = 0; b < w2; ++b) {
                  const double y = Wb[((size_t)(m * d2 + tp) * d2 + t) * w2 + b];
                  if (y == 0.0) continue;
                  const size_t row = (size_t)b * D + (sp * d2 + tp);
                  const size_t col = (size_t)a * D + (s * d2 + t);
                  M[row * (w0 * D) + col] += x * y;
                }
          }
    build_sparse_op(e->op2[p - 1], w2 * D, w0 * D, M);
    e->op2_ok[p - 1] = 1;
  }
  return e->op2[p - 1];
}

Prec gemm_prec_full(tn_env* e) { return e->cplx() ? Prec::Z64 : Prec::F64; }

// environment updates -------------------------------------------------------
void update_left(tn_env* e, int k) {  // LR[k] from LR[k-1] and site k
  PhaseTimer pt(e->ctx, PH_ENV);
  tn_ctx* c = e->ctx;
  const EnvT& Lin = e->LR[k - 1];
  const SiteT& S = e->A[k - 1];
  const long chil = S.chil, chir = S.chir;
  const int d = e->d[k - 1], wl = e->w[k - 1], wr = e->w[k];
  TN_REQUIRE(Lin.chi == chil && Lin.w == wl, TN_E_SHAPE,
             "left environment " + std::to_string(k - 1) + " does not match site " + std::to_string(k));
  const size_t es = e->esz;
  void* X = c->workspace(WS_ENV_X, (size_t)wl * chil * d * chir * es);
  void* Y = c->workspace(WS_ENV_Y, (size_t)wr * chil * d * chir * es);
  const Prec P = gemm_prec_full(e);
  GemmDesc g1;
  g1.M = (long)wl * chil; g1.N = (long)d * chir; g1.K = chil;
  g1.A = Lin.buf.p; g1.lda = chil;
  g1.B = S.buf.p; g1.ldb = (long)d * chir;
  g1.C = X; g1.ldc = (long)d * chir;
  gemm(c, P, g1);
  const int m2 = e->cplx() ? 2 : 1;
  WStepDesc ws;
  ws.G = wl; ws.Gp = wr; ws.D = d; ws.Dp = d;
  ws.nA = chil; ws.nB = chir * m2;
  ws.in = X; ws.in_sg = chil * d * chir * m2; ws.in_sA = d * chir * m2; ws.in_sS = chir * m2;
  ws.out = Y; ws.out_sg = chil * d * chir * m2; ws.out_sA = d * chir * m2; ws.out_sS = chir * m2;
  ws.op = &op_left(e, k);
  wstep(c, false, ws);
  EnvT& Lo = e->LR[k];
  Lo.buf.ensure((size_t)wr * chir * chir * es);
  Lo.chi = chir;
  Lo.w = wr;
  GemmDesc g3;
  g3.M = chir; g3.N = chir; g3.K = chil * d;
  g3.A = S.buf.p; g3.lda = chir; g3.tA = 1; g3.sA = 0; g3.cA = e->cplx();
  g3.B = Y; g3.ldb = chir; g3.sB = chil * d * chir;
  g3.C = Lo.buf.p; g3.ldc = chir; g3.sC = chir * chir;
  g3.batch = wr;
  gemm(c, P, g3);
  e->LR_version[k] = ++e->env_version;
}

void update_right(tn_env* e, int k) {  // LR[k] from LR[k+1] and site k
  PhaseTimer pt(e->ctx, PH_ENV);
  tn_ctx* c = e->ctx;
  const EnvT& Rin = e->LR[k + 1];
  const SiteT& S = e->A[k - 1];
  const long chil = S.chil, chir = S.chir;
  const int d = e->d[k - 1], wl = e->w[k - 1], wr = e->w[k];
  TN_REQUIRE(Rin.chi == chir && Rin.w == wr, TN_E_SHAPE,
             "right environment " + std::to_string(k + 1) + " does not match site " + std::to_string(k));
  const size_t es = e->esz;
  void* X = c->workspace(WS_ENV_X, (size_t)wr * chil * d * chir * es);
  void* Y = c->workspace(WS_ENV_Y, (size_t)wl * chil * d * chir * es);
  const Prec P = gemm_prec_full(e);
  GemmDesc g1;
  g1.M = chil * d; g1.N = chir; g1.K = chir;
  g1.A = S.buf.p; g1.lda = chir; g1.sA = 0;
  g1.B = Rin.buf.p; g1.ldb = chir; g1.sB = chir * chir;
  g1.C = X; g1.ldc = chir; g1.sC = chil * d * chir;
  g1.batch = wr;
  gemm(c, P, g1);
  const int m2 = e->cplx() ? 2 : 1;
  WStepDesc ws;
  ws.G = wr; ws.Gp = wl; ws.D = d; ws.Dp = d;
  ws.nA = chil; ws.nB = chir * m2;
  ws.in = X; ws.in_sg = chil * d * chir * m2; ws.in_sA = d * chir * m2; ws.in_sS = chir * m2;
  ws.out = Y; ws.out_sg = chil * d * chir * m2; ws.out_sA = d * chir * m2; ws.out_sS = chir * m2;
  ws.op = &op_right(e, k);
  wstep(c, false, ws);
  EnvT& Ro = e->LR[k];
  Ro.buf.ensure((size_t)wl * chil * chil * es);
  Ro.chi = chil;
  Ro.w = wl;
  GemmDesc g3;
  g3.M = chil; g3.N = chil; g3.K = d * chir;
  g3.A = Y; g3.lda = d * chir; g3.sA = chil * d * chir;
  g3.B = S.buf.p; g3.ldb = d * chir; g3.tB = 1; g3.sB = 0; g3.cB = e->cplx();
  g3.C = Ro.buf.p; g3.ldc = chil; g3.sC = chil * chil;
  g3.batch = wl;
  gemm(c, P, g3);
  e->LR_version[k] = ++e->env_version;
}

// ProjMPO makeL!/makeR!/position! (mirror src/mps/projmps2.jl:77-149)
void makeL(tn_env* e, int k) {
  int ll = e->lpos;
  if (ll >= k) { e->lpos = k; return; }
  ll = std::max(ll, 0);
  while (ll < k) {
    update_left(e, ll + 1);
    ++ll;
  }
  e->lpos = k;
}
void makeR(tn_env* e, int k) {
  int rl = e->rpos;
  if (rl <= k) { e->rpos = k; return; }
  rl = std::min(rl, e->N + 1);
  while (rl > k) {
    update_right(e, rl - 1);
    --rl;
  }
  e->rpos = k;
}
void proj_position(tn_env* e, int pos);
void position(tn_env* e, int pos) {
  TN_REQUIRE(pos >= 1 && pos + e->nsite - 1 <= e->N, TN_E_POSITION,
             "position " + std::to_string(pos) + " out of range for nsite=" + std::to_string(e->nsite));
  require_mpo(e);
  require_sites(e);
  makeL(e, pos - 1);
  makeR(e, pos + e->nsite);
  if (!e->proj.empty()) proj_position(e, pos);
  // the W-step operator of this position is built here, so that products at
  // this position (including the allocation-free one) upload nothing
  if (e->nsite == 2) (void)op_two(e, pos);
  else if (e->nsite == 1) (void)op_left(e, pos);
}

// overlap environments of a projector (mirror projmps2.jl:77-149):
//   L_k[b, b'] = sum L_{k-1}[a, a'] psi_k[a, s, b] conj(M_k[a', s, b'])
//   R_k[a, a'] = sum psi_k[a, s, b] conj(M_k[a', s, b']) R_{k+1}[b, b']
void proj_edge(tn_env* e, ProjT& P, int k) {
  EnvT& E = P.LR[k];
  E.buf.ensure(16);
  TN_HIP(hipMemcpyAsync(E.buf.p, e->ctx->ones, e->esz, hipMemcpyDeviceToDevice, e->ctx->stream));
  E.chi = 1;
  P.chiM[k] = 1;
}
void proj_update_left(tn_env* e, ProjT& P, int k) {
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const Prec pr = cx ? Prec::Z64 : Prec::F64;
  if (k == 1) proj_edge(e, P, 0);
  const EnvT& L = P.LR[k - 1];
  const SiteT& A = e->A[k - 1];
  const SiteT& M = P.M[k - 1];
  const int d = e->d[k - 1];
  TN_REQUIRE(L.chi == A.chil && P.chiM[k - 1] == M.chil, TN_E_SHAPE, "projector: bond dimensions disagree");
  double* X = (double*)c->workspace(WS_PROJ_X, (size_t)A.chil * d * M.chir * e->esz);
  GemmDesc g1;  // X[a, (s b')] = L[a, a'] conj(M)[a', (s b')]
  g1.M = A.chil; g1.N = (long)d * M.chir; g1.K = M.chil;
  g1.A = L.buf.p; g1.lda = M.chil;
  g1.B = M.buf.p; g1.ldb = (long)d * M.chir; g1.cB = cx;
  g1.C = X; g1.ldc = (long)d * M.chir;
  gemm(c, pr, g1);
  EnvT& Ln = P.LR[k];
  Ln.buf.ensure((size_t)A.chir * M.chir * e->esz);
  GemmDesc g2;  // L'[b, b'] = psi[(a s), b]^T X[(a s), b']
  g2.M = A.chir; g2.N = M.chir; g2.K = A.chil * d;
  g2.A = A.buf.p; g2.lda = A.chir; g2.tA = 1;
  g2.B = X; g2.ldb = M.chir;
  g2.C = Ln.buf.p; g2.ldc = M.chir;
  gemm(c, pr, g2);
  Ln.chi = A.chir;
  P.chiM[k] = M.chir;
}
void proj_update_right(tn_env* e, ProjT& P, int k) {
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const Prec pr = cx ? Prec::Z64 : Prec::F64;
  if (k == e->N) proj_edge(e, P, e->N + 1);
  const EnvT& R = P.LR[k + 1];
  const SiteT& A = e->A[k - 1];
  const SiteT& M = P.M[k - 1];
  const int d = e->d[k - 1];
  TN_REQUIRE(R.chi == A.chir && P.chiM[k + 1] == M.chir, TN_E_SHAPE, "projector: bond dimensions disagree");
  double* X = (double*)c->workspace(WS_PROJ_X, (size_t)M.chil * d * A.chir * e->esz);
  GemmDesc g1;  // X[(a' s), b] = conj(M)[(a' s), b'] R[b, b']^T
  g1.M = M.chil * d; g1.N = A.chir; g1.K = M.chir;
  g1.A = M.buf.p; g1.lda = M.chir; g1.cA = cx;
  g1.B = R.buf.p; g1.ldb = M.chir; g1.tB = 1;
  g1.C = X; g1.ldc = A.chir;
  gemm(c, pr, g1);
  EnvT& Rn = P.LR[k];
  Rn.buf.ensure((size_t)A.chil * M.chil * e->esz);
  GemmDesc g2;  // R'[a, a'] = psi[a, (s b)] X[a', (s b)]^T
  g2.M = A.chil; g2.N = M.chil; g2.K = (long)d * A.chir;
  g2.A = A.buf.p; g2.lda = (long)d * A.chir;
  g2.B = X; g2.ldb = (long)d * A.chir; g2.tB = 1;
  g2.C = Rn.buf.p; g2.ldc = M.chil;
  gemm(c, pr, g2);
  Rn.chi = A.chil;
  P.chiM[k] = M.chil;
}
void proj_position(tn_env* e, int pos) {
  for (ProjT& P : e->proj) {
    const int kl = pos - 1, kr = pos + e->nsite;
    if (P.lpos >= kl) P.lpos = kl;
    for (int ll = std::max(P.lpos, 0); ll < kl; ++ll) proj_update_left(e, P, ll + 1);
    if (kl == 0) proj_edge(e, P, 0);
    P.lpos = kl;
    if (P.rpos <= kr) P.rpos = kr;
    for (int rl = std::min(P.rpos, e->N + 1); rl > kr; --rl) proj_update_right(e, P, rl - 1);
    if (kr == e->N + 1) proj_edge(e, P, e->N + 1);
    P.rpos = kr;
    P.m_ok = false;
  }
}

void invalidate_site(tn_env* e, int k) {
  if (k <= e->lpos) e->lpos = k - 1;
  if (k >= e->rpos) e->rpos = k + 1;
  for (ProjT& P : e->proj) {
    if (k <= P.lpos) P.lpos = k - 1;
    if (k >= P.rpos) P.rpos = k + 1;
    P.m_ok = false;
  }
}

// shape of the operator domain at the current position
struct Dom {
  int pos;
  long chil, chir;
  int d1, d2;
  long n;  // elements
};
Dom domain(tn_env* e) {
  const int pos = e->lpos + 1;
  TN_REQUIRE(e->rpos == pos + e->nsite, TN_E_POSITION,
             "environment not positioned: call position! first (lpos=" + std::to_string(e->lpos) +
                 ", rpos=" + std::to_string(e->rpos) + ", nsite=" + std::to_string(e->nsite) + ")");
  Dom D;
  D.pos = pos;
  if (e->nsite == 2) {
    D.chil = e->A[pos - 1].chil;
    D.chir = e->A[pos].chir;
    D.d1 = e->d[pos - 1];
    D.d2 = e->d[pos];
  } else if (e->nsite == 1) {
    D.chil = e->A[pos - 1].chil;
    D.chir = e->A[pos - 1].chir;
    D.d1 = e->d[pos - 1];
    D.d2 = 1;
  } else if (e->nsite == 0) {  // bond matrix between sites pos-1 and pos
    D.chil = pos >= 2 ? e->A[pos - 2].chir : 1;
    D.chir = pos <= e->N ? e->A[pos - 1].chil : 1;
    D.d1 = 1;
    D.d2 = 1;
  } else {
    throw Error(TN_E_UNSUPPORTED, "nsite must be 0, 1 or 2");
  }
  D.n = D.chil * D.d1 * D.d2 * D.chir;
  return D;
}

inline long pad4(long x) { return ((x + 3) / 4) * 4; }

// FP32 shadow of an environment (mixed-precision matvecs): rows (w, chi) of
// chi entries, with row stride *ld = chi, or chi rounded up to 4 when pad
// (16-byte aligned rows for the vector-load GEMM)
const float* shadow(tn_env* e, int k, bool left, bool pad, long* ld) {
  DevBuf& B = left ? e->L32 : e->R32;
  long& key = left ? e->shadow_key_l : e->shadow_key_r;
  const EnvT& E = e->LR[k];
  const long want = (e->LR_version[k] * 8192 + k) * 2 + (pad ? 1 : 0);
  const long rows = (long)E.w * E.chi, ldd = pad ? pad4(E.chi) : E.chi;
  if (key != want) {
    B.ensure((size_t)rows * ldd * sizeof(float));
    cvt_f64_to_f32_2d(e->ctx, E.buf.as<double>(), rows, E.chi, E.chi, B.as<float>(), ldd);
    key = want;
  }
  *ld = ldd;
  return B.as<float>();
}

inline long pad8(long x) { return ((x + 7) / 8) * 8; }

// A real FP64 environment with 16-byte aligned rows: itself when chi is
// even, else a copy with row stride chi + 1 and a zero last column
const double* pad64(tn_env* e, int k, bool left, long* ld) {
  const EnvT& E = e->LR[k];
  if (E.chi % 2 == 0) {
    *ld = E.chi;
    return E.buf.as<double>();
  }
  DevBuf& B = left ? e->L64 : e->R64;
  long& key = left ? e->pad64_key_l : e->pad64_key_r;
  const long want = e->LR_version[k] * 8192 + k;
  const long rows = (long)E.w * E.chi, ldd = E.chi + 1;
  if (key != want) {
    B.ensure((size_t)rows * ldd * sizeof(double));
    pad_rows_f64(e->ctx, E.buf.as<double>(), rows, E.chi, E.chi, B.as<double>(), ldd);
    key = want;
  }
  *ld = ldd;
  return B.as<double>();
}

// FP32 shadows of complex environments for the complex mixed-precision
// product: left Lt[(w a'), a] as rows [re | im] (w chi x 2 chi); right
// Rt[(w2 b), b'] as the real block [[re, im], [-im, re]] (2 w2 chi x 2 chi)
const float* cshadow(tn_env* e, int k, bool left) {
  DevBuf& B = left ? e->L32 : e->R32;
  long& key = left ? e->shadow_key_l : e->shadow_key_r;
  const EnvT& E = e->LR[k];
  const long want = e->LR_version[k] * 4096 + k;
  const long rows = (long)E.w * E.chi, cols = E.chi;
  if (key != want) {
    B.ensure((size_t)rows * cols * (left ? 2 : 4) * sizeof(float));
    if (left) cplx_split_rows_f32(e->ctx, E.buf.as<double>(), rows, cols, B.as<float>());
    else cplx_block_f32(e->ctx, E.buf.as<double>(), rows, cols, B.as<float>());
    key = want;
  }
  return B.as<float>();
}

// bf16 hi/lo shadows of the environments for the BF16x3 matvec:
//   left  Lt[(w a'), a]   -> rows w*chi, k = a contiguous, ld = pad8(chi)
//   right Rt[(w2 b), b']  -> transposed: rows b', k = (w2 b) contiguous,
//                            ld = pad8(w2 chi)
struct SplitView {
  const unsigned short* hi;
  const unsigned short* lo;
  long ld;
};
SplitView split_shadow(tn_env* e, int k, bool left) {
  DevBuf& H = left ? e->Lh : e->Rh;
  DevBuf& Lo = left ? e->Ll : e->Rl;
  long& key = left ? e->split_key_l : e->split_key_r;
  const EnvT& E = e->LR[k];
  const long want = e->LR_version[k] * 4096 + k;
  long ld, rows;
  if (left) { rows = (long)E.w * E.chi; ld = pad8(E.chi); }
  else { rows = E.chi; ld = pad8((long)E.w * E.chi); }
  if (key != want) {
    H.ensure((size_t)rows * ld * 2);
    Lo.ensure((size_t)rows * ld * 2);
    if (left)
      split_bf16(e->ctx, E.buf.p, true, rows, E.chi, E.chi, false, H.as<unsigned short>(), Lo.as<unsigned short>(), ld);
    else  // source [(w2 b)][b'] is (w2 chi) x chi; transposed to chi x (w2 chi)
      split_bf16(e->ctx, E.buf.p, true, (long)E.w * E.chi, E.chi, E.chi, true, H.as<unsigned short>(),
                 Lo.as<unsigned short>(), ld);
    key = want;
  }
  return SplitView{H.as<unsigned short>(), Lo.as<unsigned short>(), ld};
}

// zero-site product on the bond matrix C[a, b] between sites pos-1 and pos
// (the backward step of a one-site update, update_site.jl:176-183):
//   out[a', b'] = sum_w sum_{a,b} L[w][a'][a] C[a][b] R[w][b][b']
// T[(w a'), b] = L C, permuted to T'[a'][(w b)], then out = T' R[(w b), b'].
// Always FP64 (complex when the environment is).
void matvec_zero(tn_env* e, const Dom& D, const void* v, void* out) {
  tn_ctx* c = e->ctx;
  const int pos = D.pos;
  const EnvT& L = e->LR[pos - 1];
  const EnvT& R = e->LR[pos];
  const long chil = D.chil, chir = D.chir;
  const int w = e->w[pos - 1];
  TN_REQUIRE(L.chi == chil && L.w == w && R.chi == chir && R.w == w, TN_E_SHAPE,
             "The order of the ProjMPO-ITensor product P*v is not equal to the order of the ITensor v: "
             "environment and bond-matrix shapes disagree");
  const bool cx = e->cplx();
  const Prec P = cx ? Prec::Z64 : Prec::F64;
  const int m2 = cx ? 2 : 1;
  double* T1 = (double*)c->workspace(WS_T1, (size_t)w * chil * chir * e->esz);
  double* T2 = (double*)c->workspace(WS_T2, (size_t)w * chil * chir * e->esz);
  GemmDesc g1;
  g1.M = (long)w * chil; g1.N = chir; g1.K = chil;
  g1.A = L.buf.p; g1.lda = chil;
  g1.B = v; g1.ldb = chir;
  g1.C = T1; g1.ldc = chir;
  gemm(c, P, g1);
  swap01_f64(c, T1, w, chil, chir * m2, T2);
  GemmDesc g2;
  g2.M = chil; g2.N = chir; g2.K = (long)w * chir;
  g2.A = T2; g2.lda = (long)w * chir;
  g2.B = R.buf.p; g2.ldb = chir;
  g2.C = out; g2.ldc = chir;
  gemm(c, P, g2);
}

// ---------------------------------------------------------------------------
// QN-conserving products: block-sparse GEMMs over the link sectors
// ---------------------------------------------------------------------------
using QN2 = std::array<int, 2>;
inline QN2 qplus(const QN2& a, const QN2& b) { return QN2{a[0] + b[0], a[1] + b[1]}; }
inline QN2 qminus(const QN2& a, const QN2& b) { return QN2{a[0] - b[0], a[1] - b[1]}; }

// flux of every MPO channel, propagated from the left edge (flux 0): a
// nonzero W_k[wl, s', s, wr] carries flux(wr) = flux(wl) + q(s') - q(s).
// An entry that contradicts it, or a right edge with flux != 0, means the
// MPO does not conserve the QNs.
void infer_mpo_flux(tn_env* e) {
  if (e->flux_ver == e->mpo_ver) return;
  const int N = e->N;
  e->mpo_flux.assign(N + 1, {});
  e->flux_known.assign(N + 1, {});
  for (int k = 0; k <= N; ++k) {
    e->mpo_flux[k].assign(e->w[k], QN2{0, 0});
    e->flux_known[k].assign(e->w[k], 0);
  }
  e->flux_known[0][0] = 1;
  for (int k = 1; k <= N; ++k) {
    const int wl = e->w[k - 1], wr = e->w[k], d = e->d[k - 1];
    const auto& W = e->Wh[k - 1];
    const auto& q = e->site_qn[k - 1];
    for (int a = 0; a < wl; ++a) {
      if (!e->flux_known[k - 1][a]) continue;
      for (int sp = 0; sp < d; ++sp)
        for (int s = 0; s < d; ++s)
          for (int b = 0; b < wr; ++b) {
            if (W[((size_t)(a * d + sp) * d + s) * wr + b] == 0.0) continue;
            const QN2 f = qplus(e->mpo_flux[k - 1][a], qminus(q[sp], q[s]));
            if (!e->flux_known[k][b]) {
              e->flux_known[k][b] = 1;
              e->mpo_flux[k][b] = f;
            } else {
              TN_REQUIRE(e->mpo_flux[k][b] == f, TN_E_INVALID,
                         "QN conservation: the MPO tensor of site " + std::to_string(k) +
                             " does not conserve the QNs");
            }
          }
    }
  }
  TN_REQUIRE(!e->flux_known[N][0] || e->mpo_flux[N][0] == (QN2{0, 0}), TN_E_INVALID,
             "QN conservation: the MPO changes the total QN");
  e->flux_ver = e->mpo_ver;
}

struct SectorRange {
  QN2 q;
  long off, n;
};
// sectors of a link as contiguous ranges; false if some QN occurs in two runs
bool link_sectors(const std::vector<QN2>& lq, std::vector<SectorRange>& out, std::map<QN2, int>& index) {
  out.clear();
  index.clear();
  for (long a = 0; a < (long)lq.size(); ++a) {
    if (!out.empty() && out.back().q == lq[a]) {
      ++out.back().n;
      continue;
    }
    if (index.count(lq[a])) return false;
    index[lq[a]] = (int)out.size();
    out.push_back({lq[a], a, 1});
  }
  return true;
}

// append the BS_TILE x BS_TILE tiles of task t to tl
void push_tiles(std::vector<int4>& tl, int t, int M, int N) {
  for (int m0 = 0; m0 < M; m0 += BS_TILE)
    for (int n0 = 0; n0 < N; n0 += BS_TILE) tl.push_back(int4{t, m0, n0, 0});
}
void pad_tiles(std::vector<int4>& tl) {
  while (tl.size() % 8) tl.push_back(int4{-1, 0, 0, 0});
}

// Task lists of the two-site product at position pos (nsite 2) with
// T1[(w a'), (S b)] = Lt[(w a'), a] phi[a, (S b)]  and
// out[(a' S'), b'] = T2[(a' S'), (w2 b)] Rt[(w2 b), b']:
//  * GEMM 1: one task per (w, sector Qa' of link pos-1, S) whose ket sector
//    Qa = Qa' - flux(w) exists, output block columns b in Qb = Qa + q(S)
//    (tasks without Qa write zeros, so every T1 block the W step can carry
//    into a block GEMM 2 reads is written);
//  * GEMM 2: one task per (Qa', S') with Qb' = Qa' + q(S') on link pos+1,
//    one K segment per channel w2 (ket sector Qb = Qb' - flux(w2)).
// Blocks of T1 and T2 outside these are never read; out is zero-filled.
const QnMvPlan* qn_plan(tn_env* e, int pos) {
  QnMvPlan& P = e->qplan;
  const long key[4] = {pos, e->link_ver[pos - 1], e->link_ver[pos + 1], e->mpo_ver};
  if (std::equal(key, key + 4, P.key)) return P.ok ? &P : nullptr;
  P.ok = false;
  infer_mpo_flux(e);
  std::copy(key, key + 4, P.key);
  std::vector<SectorRange> sa, sb;
  std::map<QN2, int> ia, ib;
  if (!link_sectors(e->link_qn[pos - 1], sa, ia) || !link_sectors(e->link_qn[pos + 1], sb, ib)) return nullptr;
  const long chil = (long)e->link_qn[pos - 1].size(), chir = (long)e->link_qn[pos + 1].size();
  const int d1 = e->d[pos - 1], d2 = e->d[pos], DD = d1 * d2;
  const int w0 = e->w[pos - 1], w2 = e->w[pos + 1];
  std::vector<QN2> qS(DD);
  for (int s1 = 0; s1 < d1; ++s1)
    for (int s2 = 0; s2 < d2; ++s2) qS[s1 * d2 + s2] = qplus(e->site_qn[pos - 1][s1], e->site_qn[pos][s2]);
  P.t1.clear(); P.t3.clear(); P.s1.clear(); P.s3.clear(); P.tl1.clear(); P.tl3.clear();
  double flops = 0.0;
  const long ld1 = (long)DD * chir;
  auto task1 = [&](int w, const SectorRange& ap, int S, const SectorRange& bq, const SectorRange* aq) {
    BsTask t;
    t.c_off = ((long)w * chil + ap.off) * ld1 + (long)S * chir + bq.off;
    t.M = (int)ap.n;
    t.N = (int)bq.n;
    t.seg0 = (int)P.s1.size();
    t.nseg = 0;
    if (aq) {
      BsSeg g;
      g.a_off = ((long)w * chil + ap.off) * chil + aq->off;
      g.b_off = aq->off * ld1 + (long)S * chir + bq.off;
      g.K = aq->n;
      P.s1.push_back(g);
      t.nseg = 1;
      flops += 2.0 * t.M * t.N * g.K;
    }
    push_tiles(P.tl1, (int)P.t1.size(), t.M, t.N);
    P.t1.push_back(t);
  };
  for (int w = 0; w < w0; ++w) {
    for (const SectorRange& ap : sa) {
      if (!e->flux_known[pos - 1][w]) {
        // no nonzero environment on this channel: zero its T1 rows everywhere
        for (int S = 0; S < DD; ++S)
          for (const SectorRange& bq : sb) task1(w, ap, S, bq, nullptr);
        continue;
      }
      const QN2 qa = qminus(ap.q, e->mpo_flux[pos - 1][w]);
      auto ita = ia.find(qa);
      const SectorRange* aq = (ita == ia.end()) ? nullptr : &sa[ita->second];
      for (int S = 0; S < DD; ++S) {
        auto itb = ib.find(qplus(qa, qS[S]));
        if (itb != ib.end()) task1(w, ap, S, sb[itb->second], aq);
      }
    }
  }
  for (const SectorRange& ap : sa) {
    for (int S = 0; S < DD; ++S) {
      auto it = ib.find(qplus(ap.q, qS[S]));
      if (it == ib.end()) continue;
      const SectorRange& bp = sb[it->second];
      BsTask t;
      t.c_off = (ap.off * DD + S) * chir + bp.off;
      t.M = (int)ap.n;
      t.N = (int)bp.n;
      t.seg0 = (int)P.s3.size();
      t.nseg = 0;
      for (int v = 0; v < w2; ++v) {
        if (!e->flux_known[pos + 1][v]) continue;  // T2 is zero on such a channel
        auto jt = ib.find(qminus(bp.q, e->mpo_flux[pos + 1][v]));
        if (jt == ib.end()) continue;
        const SectorRange& kb = sb[jt->second];
        BsSeg g;
        g.a_off = (ap.off * DD + S) * (long)w2 * chir + (long)v * chir + kb.off;
        g.b_off = ((long)v * chir + kb.off) * chir + bp.off;
        g.K = kb.n;
        P.s3.push_back(g);
        ++t.nseg;
        flops += 2.0 * t.M * t.N * g.K;
      }
      push_tiles(P.tl3, (int)P.t3.size(), t.M, t.N);
      P.t3.push_back(t);
    }
  }
  pad_tiles(P.tl1);
  pad_tiles(P.tl3);
  // every block the kernels touch lies inside its buffer
  auto inside = [](long off, long rows, long cols, long ld, long size) {
    return rows <= 0 || cols <= 0 || (off >= 0 && off + (rows - 1) * ld + cols <= size);
  };
  const long nL = (long)w0 * chil * chil, nphi = chil * DD * chir, nT1 = (long)w0 * chil * DD * chir;
  const long nT2 = (long)w2 * chil * DD * chir, nR = (long)w2 * chir * chir;
  for (const BsTask& t : P.t1) {
    TN_REQUIRE(inside(t.c_off, t.M, t.N, ld1, nT1), TN_E_INVALID, "QN product: block outside T1");
    for (int i = t.seg0; i < t.seg0 + t.nseg; ++i)
      TN_REQUIRE(inside(P.s1[i].a_off, t.M, P.s1[i].K, chil, nL) && inside(P.s1[i].b_off, P.s1[i].K, t.N, ld1, nphi),
                 TN_E_INVALID, "QN product: block outside L or phi");
  }
  for (const BsTask& t : P.t3) {
    TN_REQUIRE(inside(t.c_off, t.M, t.N, ld1, nphi), TN_E_INVALID, "QN product: block outside the output");
    for (int i = t.seg0; i < t.seg0 + t.nseg; ++i)
      TN_REQUIRE(inside(P.s3[i].a_off, t.M, P.s3[i].K, (long)DD * w2 * chir, nT2) &&
                     inside(P.s3[i].b_off, P.s3[i].K, t.N, chir, nR),
                 TN_E_INVALID, "QN product: block outside T2 or R");
  }
  // one upload: tasks, segments, tiles of both GEMMs (16-byte aligned parts)
  auto al = [](size_t x) { return (x + 15) / 16 * 16; };
  const size_t b_t1 = al(P.t1.size() * sizeof(BsTask)), b_t3 = al(P.t3.size() * sizeof(BsTask));
  const size_t b_s1 = al(P.s1.size() * sizeof(BsSeg)), b_s3 = al(P.s3.size() * sizeof(BsSeg));
  const size_t b_l1 = al(P.tl1.size() * sizeof(int4)), b_l3 = al(P.tl3.size() * sizeof(int4));
  const size_t total = b_t1 + b_t3 + b_s1 + b_s3 + b_l1 + b_l3 + 16;
  std::vector<char> host(total, 0);
  size_t o = 0;
  auto put = [&](const void* src, size_t n, size_t span) {
    if (n) std::memcpy(host.data() + o, src, n);
    const size_t at = o;
    o += span;
    return at;
  };
  const size_t o_t1 = put(P.t1.data(), P.t1.size() * sizeof(BsTask), b_t1);
  const size_t o_t3 = put(P.t3.data(), P.t3.size() * sizeof(BsTask), b_t3);
  const size_t o_s1 = put(P.s1.data(), P.s1.size() * sizeof(BsSeg), b_s1);
  const size_t o_s3 = put(P.s3.data(), P.s3.size() * sizeof(BsSeg), b_s3);
  const size_t o_l1 = put(P.tl1.data(), P.tl1.size() * sizeof(int4), b_l1);
  const size_t o_l3 = put(P.tl3.data(), P.tl3.size() * sizeof(int4), b_l3);
  P.dev.ensure(total);
  TN_HIP(hipMemcpyAsync(P.dev.p, host.data(), total, hipMemcpyHostToDevice, e->ctx->stream));
  TN_HIP(hipStreamSynchronize(e->ctx->stream));  // host staging goes out of scope
  char* base = (char*)P.dev.p;
  P.dt1 = (const BsTask*)(base + o_t1);
  P.dt3 = (const BsTask*)(base + o_t3);
  P.ds1 = (const BsSeg*)(base + o_s1);
  P.ds3 = (const BsSeg*)(base + o_s3);
  P.dtl1 = (const int4*)(base + o_l1);
  P.dtl3 = (const int4*)(base + o_l3);
  P.flops = flops;
  P.ok = true;
  return &P;
}

// two-site product of a QN-conserving environment on the sector blocks only
// (FP64, or FP32 on the FP32 environment shadows); same result as the dense
// product, whose blocks outside the sectors are zero
void matvec_qn(tn_env* e, const QnMvPlan& qp, const Dom& D, const SparseOp& op, const void* v, void* out,
               int prec) {
  tn_ctx* c = e->ctx;
  // FP32 and BF16x3 take FP32 copies of psi and the environments; BF16x3
  // splits the FP32 fragments into bf16 hi/lo parts inside the block product
  const bool f32 = prec != TN_PREC_FP64;
  const int pos = D.pos;
  const long chil = D.chil, chir = D.chir;
  const int DD = D.d1 * D.d2;
  const int w0 = e->w[pos - 1], w2 = e->w[pos + 1];
  const size_t es = f32 ? 4 : 8;
  // the block GEMM writes only T1's allowed blocks, while a dense operator's
  // W step (an MFMA product) multiplies every entry of T1: the rest must hold
  // zeros, not stale workspace bytes (0 * NaN is NaN).  A sparse operator's W
  // step reads, for every allowed T2 entry, only allowed T1 entries (its
  // nonzero couplings conserve the flux), so it needs no clearing.
  const size_t t1w = ((size_t)w0 * chil * DD * chir * es + 7) / 8;
  void* T1 = c->workspace(WS_T1, t1w * 8);
  if (op.dense) fill_f64(c, (double*)T1, 0.0, (long)t1w);
  void* T2 = c->workspace(WS_T2, (size_t)w2 * chil * DD * chir * es);
  const void* vin = v;
  const void* Lp = e->LR[pos - 1].buf.p;
  const void* Rp = e->LR[pos + 2].buf.p;
  void* outp = out;
  if (f32) {
    float* v32 = (float*)c->workspace(WS_PSI32, (size_t)D.n * 4);
    cvt_f64_to_f32(c, (const double*)v, v32, D.n);
    vin = v32;
    long ldl, ldr;  // unpadded: the block offsets assume row strides chi_l and chi_r
    Lp = shadow(e, pos - 1, true, false, &ldl);
    Rp = shadow(e, pos + 2, false, false, &ldr);
    (void)ldl;
    (void)ldr;
    outp = c->workspace(WS_OUT32, (size_t)(D.n + 1) / 2 * 8);
  }
  fill_f64(c, (double*)outp, 0.0, f32 ? (D.n + 1) / 2 : D.n);
  const Prec P = prec == TN_PREC_FP64 ? Prec::F64 : (prec == TN_PREC_FP32 ? Prec::F32 : Prec::BF16X3);
  BsDesc g1;
  g1.A = Lp; g1.lda = chil;
  g1.B = vin; g1.ldb = (long)DD * chir;
  g1.C = T1; g1.ldc = (long)DD * chir;
  g1.tasks = qp.dt1; g1.segs = qp.ds1; g1.tiles = qp.dtl1; g1.ntiles = (long)qp.tl1.size();
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm_blocksparse(c, P, g1);
  }
  WStepDesc ws;
  ws.G = w0; ws.Gp = w2; ws.D = DD; ws.Dp = DD;
  ws.nA = chil; ws.nB = chir;
  ws.in = T1; ws.in_sg = chil * DD * chir; ws.in_sA = DD * chir; ws.in_sS = chir;
  ws.out = T2; ws.out_sg = chir; ws.out_sA = (long)DD * w2 * chir; ws.out_sS = (long)w2 * chir;
  ws.op = &op;
  wstep(c, f32, ws);
  BsDesc g3;
  g3.A = T2; g3.lda = (long)DD * w2 * chir;
  g3.B = Rp; g3.ldb = chir;
  g3.C = outp; g3.ldc = (long)DD * chir;
  g3.tasks = qp.dt3; g3.segs = qp.ds3; g3.tiles = qp.dtl3; g3.ntiles = (long)qp.tl3.size();
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm_blocksparse(c, P, g3);
  }
  if (f32) cvt_f32_to_f64(c, (const float*)outp, (double*)out, D.n);
}

// complex product in FP32 (nsite 1 or 2): the complex GEMMs as real FP32
// GEMMs on [re | im] rows times [[re, im], [-im, re]] blocks, the W step on
// the real and imaginary planes; FP64 in and out
void matvec_cplx32(tn_env* e, const Dom& D, const SparseOp& op, const void* v, void* out) {
  tn_ctx* c = e->ctx;
  const int pos = D.pos;
  const long chil = D.chil, chir = D.chir;
  const int DD = D.d1 * D.d2;
  const int w0 = e->w[pos - 1];
  const int w2 = e->w[pos - 1 + e->nsite];
  const float* Ls = cshadow(e, pos - 1, true);
  const float* Rs = cshadow(e, pos + e->nsite, false);
  const long N1 = (long)DD * chir;
  float* Bp = (float*)c->workspace(WS_PSI32, (size_t)4 * chil * N1 * sizeof(float));
  cplx_block_f32(c, (const double*)v, chil, N1, Bp);  // [[re, im], [-im, re]] of psi
  float* T1 = (float*)c->workspace(WS_T1, (size_t)w0 * chil * 2 * N1 * sizeof(float));
  float* T2 = (float*)c->workspace(WS_T2, (size_t)chil * DD * 2 * w2 * chir * sizeof(float));
  float* O32 = (float*)c->workspace(WS_OUT32, (size_t)chil * DD * 2 * chir * sizeof(float));
  GemmDesc g1;  // [T1r | T1i][(w a'), .] = [Lr | Li] * psi block
  g1.M = (long)w0 * chil; g1.N = 2 * N1; g1.K = 2 * chil;
  g1.A = Ls; g1.lda = 2 * chil;
  g1.B = Bp; g1.ldb = 2 * N1;
  g1.C = T1; g1.ldc = 2 * N1;
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm(c, Prec::F32, g1);
  }
  for (int part = 0; part < 2; ++part) {  // the W step is real: each plane alone
    WStepDesc ws;
    ws.G = w0; ws.Gp = w2; ws.D = DD; ws.Dp = DD;
    ws.nA = chil; ws.nB = chir;
    ws.in = T1 + part * N1; ws.in_sg = chil * 2 * N1; ws.in_sA = 2 * N1; ws.in_sS = chir;
    ws.out = T2 + part * (long)w2 * chir; ws.out_sg = chir; ws.out_sA = (long)DD * 2 * w2 * chir;
    ws.out_sS = 2L * w2 * chir;
    ws.op = &op;
    wstep(c, true, ws);
  }
  GemmDesc g3;  // [outr | outi][(a' S'), .] = [T2r | T2i] * R block
  g3.M = chil * DD; g3.N = 2 * chir; g3.K = 2L * w2 * chir;
  g3.A = T2; g3.lda = 2L * w2 * chir;
  g3.B = Rs; g3.ldb = 2 * chir;
  g3.C = O32; g3.ldc = 2 * chir;
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm(c, Prec::F32, g3);
  }
  cplx_join_rows_f64(c, O32, chil * DD, chir, (double*)out);
}

// product(sysenv, v) (ProjMPO.product; mirror src/mps/projmps2.jl:159-173)
void matvec_mpo(tn_env* e, const void* v, void* out, int prec) {
  PhaseTimer pt(e->ctx, PH_MATVEC);
  tn_ctx* c = e->ctx;
  const Dom D = domain(e);
  if (e->nsite == 0) {
    matvec_zero(e, D, v, out);
    return;
  }
  const int pos = D.pos;
  const EnvT& L = e->LR[pos - 1];
  const EnvT& R = e->LR[pos + e->nsite];
  const long chil = D.chil, chir = D.chir;
  const int DD = D.d1 * D.d2;
  const int w0 = e->w[pos - 1];
  const int w2 = e->w[pos - 1 + e->nsite];
  TN_REQUIRE(L.chi == chil && L.w == w0 && R.chi == chir && R.w == w2, TN_E_SHAPE,
             "The order of the ProjMPO-ITensor product P*v is not equal to the order of the ITensor v: "
             "environment and wave-function shapes disagree");
  const SparseOp& op = (e->nsite == 2) ? op_two(e, pos) : op_left(e, pos);
  const bool cx = e->cplx();
  if (e->nq > 0 && e->nsite == 2 && !cx &&
      (prec == TN_PREC_FP64 || prec == TN_PREC_FP32 || prec == TN_PREC_BF16X3) &&
      (long)e->link_qn[pos - 1].size() == chil && (long)e->link_qn[pos + 1].size() == chir) {
    const QnMvPlan* qp = qn_plan(e, pos);
    if (qp) {
      matvec_qn(e, *qp, D, op, v, out, prec);
      return;
    }
  }
  if (cx && prec == TN_PREC_FP32) {
    matvec_cplx32(e, D, op, v, out);
    return;
  }
  if (cx && prec != TN_PREC_FP64)
    throw Error(TN_E_UNSUPPORTED, "BF16x3 matvec is implemented for real tensors only");
  if (prec == TN_PREC_BF16X3) {
    // T1 = Lt * psi              (bf16x3, A = Lt split, B^T = psi^T split)
    // T2 = W-step(T1) written as bf16 hi/lo rows (a' S') x (w2 b)
    // out = T2 * Rt              (bf16x3, B^T = Rt^T split)
    const SplitView Ls = split_shadow(e, pos - 1, true);
    const SplitView Rs = split_shadow(e, pos + e->nsite, false);
    const long ldp = pad8(chil);
    const long N1 = (long)DD * chir;
    unsigned short* ph = (unsigned short*)c->workspace(WS_PSI32, (size_t)N1 * ldp * 2 * 2);
    unsigned short* pl = ph + N1 * ldp;
    split_bf16(c, v, true, chil, N1, N1, true, ph, pl, ldp);  // psi^T: N1 x chil
    float* T1 = (float*)c->workspace(WS_T1, (size_t)w0 * chil * N1 * 4);
    {
      PhaseTimer pg(c, PH_MV_GEMM);
      gemm_bf16x3_nt(c, (long)w0 * chil, N1, chil, Ls.hi, Ls.lo, Ls.ld, ph, pl, ldp, T1, N1);
    }
    const long K3 = (long)w2 * chir, ld3 = pad8(K3);
    const long M3 = chil * DD;
    unsigned short* th = (unsigned short*)c->workspace(WS_T2, (size_t)M3 * ld3 * 2 * 2);
    unsigned short* tl = th + M3 * ld3;
    WStepDesc ws;
    ws.G = w0; ws.Gp = w2; ws.D = DD; ws.Dp = DD;
    ws.nA = chil; ws.nB = chir;
    ws.in = T1; ws.in_sg = chil * N1; ws.in_sA = N1; ws.in_sS = chir;
    ws.out = nullptr; ws.out_sg = chir; ws.out_sA = (long)DD * ld3; ws.out_sS = ld3;
    ws.op = &op;
    wstep_split(c, ws, th, tl);
    // the W step writes columns below K3 only; the product reads rows up to
    // pad8(K3), and stale bytes there (possibly Inf/NaN patterns) times the
    // zero padding of Rt would poison the sums
    if (ld3 > K3) zero_cols_u16(c, th, 2 * M3, K3, ld3, ld3);  // th and tl are contiguous
    float* o32 = (float*)c->workspace(WS_OUT32, (size_t)D.n * 4);
    {
      PhaseTimer pg(c, PH_MV_GEMM);
      gemm_bf16x3_nt(c, M3, chir, K3, th, tl, ld3, Rs.hi, Rs.lo, Rs.ld, o32, chir);
    }
    cvt_f32_to_f64(c, o32, (double*)out, D.n);
    return;
  }
  Prec P;
  bool f32 = false;
  if (prec == TN_PREC_FP64) P = cx ? Prec::Z64 : Prec::F64;
  else if (prec == TN_PREC_FP32) { P = Prec::F32; f32 = true; }
  else throw Error(TN_E_INVALID, "unknown precision");
  const size_t es = f32 ? 4 : e->esz;
  const int m2 = cx ? 2 : 1;
  // FP32 operands are copies of ours: their rows are padded to 16 bytes so the
  // GEMMs take vector loads whatever the bond dimensions
  long lda1 = chil, ldb1 = (long)DD * chir, ld2 = (long)w2 * chir, ldb3 = chir;
  const void* vin = v;
  const void* Lp = L.buf.p;
  const void* Rp = R.buf.p;
  void* outp = out;
  if (f32) {
    ldb1 = pad4(ldb1);
    ld2 = pad4(ld2);
    float* v32 = (float*)c->workspace(WS_PSI32, (size_t)chil * ldb1 * 4);
    cvt_f64_to_f32_2d(c, (const double*)v, chil, (long)DD * chir, (long)DD * chir, v32, ldb1);
    vin = v32;
    Lp = shadow(e, pos - 1, true, true, &lda1);
    Rp = shadow(e, pos + e->nsite, false, true, &ldb3);
    outp = c->workspace(WS_OUT32, (size_t)D.n * 4);
  } else if (!cx) {
    // FP64: even row strides for the environments and T2 as well
    ld2 += ld2 & 1;
    Lp = pad64(e, pos - 1, true, &lda1);
    Rp = pad64(e, pos + e->nsite, false, &ldb3);
  }
  const bool zp = f32 || !cx;
  void* T1 = c->workspace(WS_T1, (size_t)w0 * chil * DD * chir * es);
  void* T2 = c->workspace(WS_T2, (size_t)chil * DD * ld2 * es);
  GemmDesc g1;
  g1.M = (long)w0 * chil; g1.N = DD * chir; g1.K = chil;
  g1.A = Lp; g1.lda = lda1; g1.zpA = zp;
  g1.B = vin; g1.ldb = ldb1; g1.zpB = zp;
  g1.C = T1; g1.ldc = DD * chir;
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm(c, P, g1);
  }
  WStepDesc ws;
  ws.G = w0; ws.Gp = w2; ws.D = DD; ws.Dp = DD;
  ws.nA = chil; ws.nB = chir * m2;
  ws.in = T1; ws.in_sg = chil * DD * chir * m2; ws.in_sA = DD * chir * m2; ws.in_sS = chir * m2;
  ws.out = T2; ws.out_sg = chir * m2; ws.out_sA = (long)DD * ld2 * m2; ws.out_sS = ld2 * m2;
  ws.op = &op;
  wstep(c, f32, ws);
  if (ld2 > (long)w2 * chir) {
    if (f32) zero_cols_f32(c, (float*)T2, chil * DD, (long)w2 * chir, ld2, ld2);
    else zero_cols_f64(c, (double*)T2, chil * DD, (long)w2 * chir, ld2, ld2);
  }
  GemmDesc g3;
  g3.M = chil * DD; g3.N = chir; g3.K = (long)w2 * chir;
  g3.A = T2; g3.lda = ld2; g3.zpA = zp;
  g3.B = Rp; g3.ldb = ldb3; g3.zpB = zp;
  g3.C = outp; g3.ldc = chir;
  {
    PhaseTimer pg(c, PH_MV_GEMM);
    gemm(c, P, g3);
  }
  if (f32) cvt_f32_to_f64(c, (const float*)outp, (double*)out, D.n);
}

// ---------------------------------------------------------------------------
// small dense symmetric eigensolver (cyclic Jacobi, FP64, ascending order)
// ---------------------------------------------------------------------------
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) (i == j ? diag : off) += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double zeta = (aqq - app) / (2.0 * apq);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int k = 0; k < n; ++k) {  // columns p, q of A
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = cs * akp - sn * akq;
          A[(size_t)k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = cs * apk - sn * aqk;
          A[(size_t)q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = cs * vkp - sn * vkq;
          V[(size_t)k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  w.resize(n);
  std::vector<double> Vs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)idx[j] * n + idx[j]];
    for (int i = 0; i < n; ++i) Vs[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
  }
  V.swap(Vs);
}

// ---------------------------------------------------------------------------
// Krylov machinery on device vectors
// ---------------------------------------------------------------------------
struct Krylov {
  tn_ctx* c;
  bool cx;
  long n;       // complex or real elements
  long nd;      // doubles per vector
  int slots;
  double* base; // slots x nd
  double* w;    // nd
  double* coef; // device scratch
  Krylov(tn_env* env, long nelem, int nslots) : Krylov(env->ctx, env->cplx(), nelem, nslots) {}
  Krylov(tn_ctx* ctx, bool cplx, long nelem, int nslots) : c(ctx), cx(cplx), n(nelem) {
    nd = cx ? 2 * n : n;
    slots = nslots;
    base = (double*)c->workspace(WS_KRY_BASIS, (size_t)slots * nd * sizeof(double));
    w = (double*)c->workspace(WS_KRY_W, (size_t)nd * sizeof(double));
    coef = (double*)c->workspace(WS_KRY_COEF, 1024 * sizeof(double));
  }
  double* V(int i) { return base + (long)i * nd; }
  // <x, x> (result in its own pinned area, clear of the two step slots)
  double norm2(const double* x) {
    multi_dot(c, x, nd, 1, x, n, cx, coef + 384, c->pinned_dev + 1536);
    TN_HIP(hipStreamSynchronize(c->stream));
    return c->pinned[1536];
  }
  // Krylov steps run one ahead of the host: step results (both passes'
  // coefficients and |w|^2) land in pinned slot s (512 doubles each), and
  // kry_ev[s] marks their completion, so the host reads step k while step
  // k + 1 is already queued.
  void ensure_events() {
    for (hipEvent_t& ev : c->kry_ev)
      if (!ev) TN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  // w <- w - V_{0..K-1} (V^H w), twice (queued; results in slot s)
  void orthogonalize_async(int K, int s) {
    // coefficient scratch: pass 1 at [0, 2K), pass 2 at [128, 128+2K),
    // |w|^2 at 256 (K <= 61 basis vectors, complex pairs fit in 128)
    double* hd = c->pinned_dev + 512 * s;
    if (krylov_ortho_small(c, base, nd, K, w, n, cx, coef, hd)) {
      ensure_events();
      TN_HIP(hipEventRecord(c->kry_ev[s], c->stream));
      return;
    }
    multi_dot(c, base, nd, K, w, n, cx, coef, hd);
    multi_axpy_sub(c, base, nd, K, coef, w, n, cx);
    multi_dot(c, base, nd, K, w, n, cx, coef + 128, hd + 128);
    multi_axpy_sub(c, base, nd, K, coef + 128, w, n, cx);
    multi_dot(c, w, nd, 1, w, n, cx, coef + 256, hd + 256);
    ensure_events();
    TN_HIP(hipEventRecord(c->kry_ev[s], c->stream));
  }
  // the total coefficients (complex pairs when cx) and |w|^2 of slot s
  void orthogonalize_wait(int K, int s, std::vector<cplx>& h, double& nrm2) {
    TN_HIP(hipEventSynchronize(c->kry_ev[s]));
    const double* hb = c->pinned + 512 * s;
    h.resize(K);
    for (int i = 0; i < K; ++i) {
      if (cx) h[i] = cplx(hb[2 * i] + hb[128 + 2 * i], hb[2 * i + 1] + hb[128 + 2 * i + 1]);
      else h[i] = cplx(hb[i] + hb[128 + i], 0.0);
    }
    nrm2 = hb[256];
  }
  void orthogonalize(int K, std::vector<cplx>& h, double& nrm2) {
    orthogonalize_async(K, 0);
    orthogonalize_wait(K, 0, h, nrm2);
  }
  // V(i) = w / |w| with |w|^2 of the last queued step (device side)
  void next_from_w(int i) { scale_rnorm(c, V(i), w, coef + 256, nd); }
  // out = sum_i coefs[i] V_i
  void combine(int K, const std::vector<cplx>& co, double* out) {
    // the pinned upload area may still feed an earlier copy: drain the stream first
    TN_HIP(hipStreamSynchronize(c->stream));
    double* hc = c->pinned + 2048;
    const int nc = cx ? 2 * K : K;
    for (int i = 0; i < K; ++i) {
      if (cx) { hc[2 * i] = co[i].real(); hc[2 * i + 1] = co[i].imag(); }
      else hc[i] = co[i].real();
    }
    double* dc = coef + 512;  // [512, 512 + 2K)
    TN_HIP(hipMemcpyAsync(dc, hc, nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    lincomb(c, base, nd, K, dc, out, n, cx);
  }
};

// local projection of a projector at the current position (proj_mps,
// projmps2.jl:177-194), conjugated: m[a, S, b] = sum conj(L[a, a']) M[a', S, b'] conj(R[b, b'])
void proj_local(tn_env* e, ProjT& P, const Dom& D) {
  if (P.m_ok && P.m_pos == D.pos && P.m_nsite == e->nsite) return;
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const Prec pr = cx ? Prec::Z64 : Prec::F64;
  const int pos = D.pos;
  const EnvT& L = P.LR[pos - 1];
  const EnvT& R = P.LR[pos + e->nsite];
  const long cml = P.chiM[pos - 1], cmr = P.chiM[pos + e->nsite];
  const int DD = (e->nsite == 2) ? D.d1 * D.d2 : D.d1;
  TN_REQUIRE(L.chi == D.chil && R.chi == D.chir, TN_E_SHAPE, "projector: environment shapes disagree");
  if (e->nsite == 0) {  // m[a, b] = sum conj(L[a, c]) conj(R[b, c])
    P.m.ensure((size_t)D.n * e->esz);
    GemmDesc g;
    g.M = D.chil; g.N = D.chir; g.K = cml;
    g.A = L.buf.p; g.lda = cml; g.cA = cx;
    g.B = R.buf.p; g.ldb = cmr; g.tB = 1; g.cB = cx;
    g.C = P.m.p; g.ldc = D.chir;
    gemm(c, pr, g);
    P.m_ok = true;
    P.m_pos = D.pos;
    P.m_nsite = 0;
    return;
  }
  const void* Mphi = P.M[pos - 1].buf.p;
  if (e->nsite == 2) {
    const SiteT& M1 = P.M[pos - 1];
    const SiteT& M2 = P.M[pos];
    double* Y = (double*)c->workspace(WS_PROJ_Y, (size_t)cml * DD * cmr * e->esz);
    GemmDesc g;
    g.M = cml * D.d1; g.N = (long)D.d2 * cmr; g.K = M1.chir;
    g.A = M1.buf.p; g.lda = M1.chir;
    g.B = M2.buf.p; g.ldb = (long)D.d2 * cmr;
    g.C = Y; g.ldc = (long)D.d2 * cmr;
    gemm(c, pr, g);
    Mphi = Y;
  }
  double* X = (double*)c->workspace(WS_PROJ_X, (size_t)D.chil * DD * cmr * e->esz);
  GemmDesc g1;  // X[a, (S b')] = conj(L)[a, a'] Mphi[a', (S b')]
  g1.M = D.chil; g1.N = (long)DD * cmr; g1.K = cml;
  g1.A = L.buf.p; g1.lda = cml; g1.cA = cx;
  g1.B = Mphi; g1.ldb = (long)DD * cmr;
  g1.C = X; g1.ldc = (long)DD * cmr;
  gemm(c, pr, g1);
  P.m.ensure((size_t)D.n * e->esz);
  GemmDesc g2;  // m[(a S), b] = X[(a S), b'] conj(R)[b, b']^T
  g2.M = D.chil * DD; g2.N = D.chir; g2.K = cmr;
  g2.A = X; g2.lda = cmr;
  g2.B = R.buf.p; g2.ldb = cmr; g2.tB = 1; g2.cB = cx;
  g2.C = P.m.p; g2.ldc = D.chir;
  gemm(c, pr, g2);
  P.m_ok = true;
  P.m_pos = D.pos;
  P.m_nsite = e->nsite;
}

// product(ProjMPO_MPS2, v) = PH(v) + weight sum_k m_k (m_k^H v) (projmpo_mps2.jl:94-101)
void matvec(tn_env* e, const void* v, void* out, int prec) {
  matvec_mpo(e, v, out, prec);
  if (e->proj.empty()) return;
  tn_ctx* c = e->ctx;
  const Dom D = domain(e);
  const bool cx = e->cplx();
  const long nd = cx ? 2 * D.n : D.n;
  double* coef = (double*)c->workspace(WS_PROJ_C, 64 * sizeof(double));
  for (ProjT& P : e->proj) {
    proj_local(e, P, D);
    multi_dot(c, P.m.as<double>(), nd, 1, (const double*)v, D.n, cx, coef);
    add_scaled_dev(c, (double*)out, P.m.as<double>(), coef, e->proj_weight, D.n, cx);
  }
}

void apply_op(tn_env* e, const double* x, double* y, int prec) { matvec(e, x, y, prec); }

}  // namespace

// ---------------------------------------------------------------------------
// allocation-free FP64 product
// ---------------------------------------------------------------------------
// One fixed array per GPU holds the product's intermediates (T1, T2 and the
// split-K partial sums).  In host code `scratch` is only a symbol: its device
// address comes from hipGetSymbolAddress, once per device, at context creation.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;  // per device

// the device address of `scratch` on `device`, which must be current
void* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TN_REQUIRE(p != nullptr, TN_E_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_addr[device] = p;
  }
  return g_scratch_addr[device];
}

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// product(sysenv, v) in FP64 with every intermediate in the fixed scratch:
// the same GEMM / W-step / GEMM chain as matvec_mpo, no device allocation,
// no host-device copy.  Needs position! at the current position (which builds
// the W-step operator) and no penalty projectors.
void matvec_noalloc(tn_env* e, const void* v, void* out) {
  tn_ctx* c = e->ctx;
  TN_REQUIRE(c->scratch_dev != nullptr, TN_E_HIP, "device scratch address not initialised");
  TN_REQUIRE(e->proj.empty(), TN_E_UNSUPPORTED,
             "allocation-free product: environments with penalty projectors are not supported");
  TN_REQUIRE(e->nsite == 1 || e->nsite == 2, TN_E_UNSUPPORTED, "allocation-free product: nsite must be 1 or 2");
  TN_REQUIRE(v != nullptr && out != nullptr, TN_E_INVALID, "null vector");
  const Dom D = domain(e);
  const int pos = D.pos;
  const EnvT& L = e->LR[pos - 1];
  const EnvT& R = e->LR[pos + e->nsite];
  const long chil = D.chil, chir = D.chir;
  const int DD = D.d1 * D.d2;
  const int w0 = e->w[pos - 1];
  const int w2 = e->w[pos - 1 + e->nsite];
  TN_REQUIRE(L.chi == chil && L.w == w0 && R.chi == chir && R.w == w2, TN_E_SHAPE,
             "The order of the ProjMPO-ITensor product P*v is not equal to the order of the ITensor v: "
             "environment and wave-function shapes disagree");
  const bool built = (e->nsite == 2) ? e->op2_ok[pos - 1] : e->opL_ok[pos - 1];
  TN_REQUIRE(built, TN_E_POSITION, "allocation-free product: operator of this position not built (call position!)");
  const SparseOp& op = (e->nsite == 2) ? e->op2[pos - 1] : e->opL[pos - 1];
  const bool cx = e->cplx();
  const Prec P = cx ? Prec::Z64 : Prec::F64;
  const int m2 = cx ? 2 : 1;
  const size_t t1 = (size_t)w0 * chil * DD * chir * e->esz;
  const size_t t2 = (size_t)w2 * chil * DD * chir * e->esz;
  char* base = (char*)c->scratch_dev;
  GemmDesc g1;
  g1.M = (long)w0 * chil; g1.N = (long)DD * chir; g1.K = chil;
  g1.A = L.buf.p; g1.lda = chil;
  g1.B = v; g1.ldb = (long)DD * chir;
  g1.C = base; g1.ldc = (long)DD * chir;
  GemmDesc g3;
  g3.M = chil * DD; g3.N = chir; g3.K = (long)w2 * chir;
  g3.A = base + align256(t1); g3.lda = (long)w2 * chir;
  g3.B = R.buf.p; g3.ldb = chir;
  g3.C = out; g3.ldc = chir;
  const size_t split = std::max(gemm_split_bytes(P, g1), gemm_split_bytes(P, g3));
  const size_t need = align256(t1) + align256(t2) + align256(split);
  TN_REQUIRE(need <= NOALLOC_SCRATCH_BYTES, TN_E_INVALID,
             "allocation-free product needs " + std::to_string(need) + " bytes of scratch, more than the fixed " +
                 std::to_string(NOALLOC_SCRATCH_BYTES));
  g1.split_ws = g3.split_ws = base + align256(t1) + align256(t2);
  g1.split_cap = g3.split_cap = split;
  gemm(c, P, g1);
  WStepDesc ws;
  ws.G = w0; ws.Gp = w2; ws.D = DD; ws.Dp = DD;
  ws.nA = chil; ws.nB = chir * m2;
  ws.in = base; ws.in_sg = chil * DD * chir * m2; ws.in_sA = DD * chir * m2; ws.in_sS = chir * m2;
  ws.out = base + align256(t1); ws.out_sg = chir * m2; ws.out_sA = (long)DD * w2 * chir * m2;
  ws.out_sS = (long)w2 * chir * m2;
  ws.op = &op;
  wstep(c, false, ws);
  gemm(c, P, g3);
}
}  // namespace

// eigsolve(A, x0, 1, :SR; krylovdim, maxiter, tol, eager=false) (src/base/solver.jl:36)
// for the symmetric operator y = A x given by `apply`
double tn::krylov_eig_op(tn_ctx* c, bool cx, long n, const std::function<void(const double*, double*)>& apply,
                         double* phi, int kdim, int maxiter, double tol, tn_krylov_info* info) {
  TN_REQUIRE(kdim >= 1 && kdim <= 60 && maxiter >= 1, TN_E_INVALID, "eig_solver: bad krylovdim/maxiter");
  Krylov K(c, cx, n, 2 * kdim + 2);
  tn_krylov_info inf{0, 0, 1, NAN};
  // v0 = x0 / |x0|
  double b0;
  {
    PhaseTimer pt(c, PH_KRYLOV);
    b0 = std::sqrt(K.norm2(phi));
    TN_REQUIRE(std::isfinite(b0) && b0 > 0.0, TN_E_NUMERIC, "eig_solver: zero or non-finite start vector");
    scaled_copy(K.c, K.V(0), phi, 1.0 / b0, K.nd);
  }
  std::vector<double> Hm((size_t)(kdim + 1) * (kdim + 1), 0.0);
  auto H = [&](int i, int j) -> double& { return Hm[(size_t)i * (kdim + 1) + j]; };
  // Kn basis vectors; the step that made V(Kn-1)'s product is queued in slot sl
  int Kn = 1, sl = 0;
  std::vector<cplx> h;
  double r2;
  apply(K.V(0), K.w);
  {
    PhaseTimer pt(c, PH_KRYLOV);
    K.orthogonalize_async(1, sl);
  }
  std::vector<double> theta, U;
  while (true) {
    // queue the next step before reading this one (it is discarded if this
    // step ends the cycle on an invariant subspace)
    const bool ahead = Kn < kdim;
    if (ahead) {
      {
        PhaseTimer pt(c, PH_KRYLOV);
        K.next_from_w(Kn);
      }
      apply(K.V(Kn), K.w);
      PhaseTimer pt(c, PH_KRYLOV);
      K.orthogonalize_async(Kn + 1, sl ^ 1);
    }
    {
      PhaseTimer pt(c, PH_KRYLOV);
      K.orthogonalize_wait(Kn, sl, h, r2);
    }
    inf.numops++;
    H(Kn - 1, Kn - 1) = h[Kn - 1].real();
    // a NaN would pass std::max(0.0, .) as zero and end the cycle as if on
    // an invariant subspace
    TN_REQUIRE(std::isfinite(r2) && std::isfinite(h[Kn - 1].real()), TN_E_NUMERIC,
               "eig_solver: non-finite product or residual");
    const double beta = std::sqrt(std::max(0.0, r2));
    if (Kn == kdim || beta <= tol) {
      std::vector<double> Hs((size_t)Kn * Kn);
      for (int i = 0; i < Kn; ++i)
        for (int j = 0; j < Kn; ++j) Hs[(size_t)i * Kn + j] = 0.5 * (H(i, j) + H(j, i));
      sym_eig(Kn, Hs, theta, U);
      std::vector<double> res(Kn);
      for (int i = 0; i < Kn; ++i) res[i] = std::fabs(beta * U[(size_t)(Kn - 1) * Kn + i]);
      int conv = 0;
      while (conv < Kn && res[conv] <= tol) ++conv;
      inf.converged = conv;
      inf.residual = res[0];
      if (conv >= 1 || beta <= tol || inf.numiter >= maxiter) {
        PhaseTimer pt(c, PH_KRYLOV);
        std::vector<cplx> co(Kn);
        for (int i = 0; i < Kn; ++i) co[i] = U[(size_t)i * Kn + 0];
        K.combine(Kn, co, phi);
        const double ny = std::sqrt(K.norm2(phi));
        scale_f64(K.c, phi, 1.0 / ny, K.nd);
        if (info) *info = inf;
        return theta[0];
      }
      // thick restart: keep (3*kdim + 2*conv)/5 Ritz vectors + residual direction
      // (here Kn == kdim: no step was queued ahead, w is this step's)
      int keep = (3 * kdim + 2 * conv) / 5;
      keep = std::max(1, std::min(keep, Kn - 1));
      {
        PhaseTimer pt(c, PH_KRYLOV);
        for (int j = 0; j < keep; ++j) {
          std::vector<cplx> co(Kn);
          for (int i = 0; i < Kn; ++i) co[i] = U[(size_t)i * Kn + j];
          K.combine(Kn, co, K.V(kdim + 1 + j));
        }
        for (int j = 0; j < keep; ++j) copy_f64(K.c, K.V(j), K.V(kdim + 1 + j), K.nd);
        scaled_copy(K.c, K.V(keep), K.w, 1.0 / beta, K.nd);
      }
      std::fill(Hm.begin(), Hm.end(), 0.0);
      for (int j = 0; j < keep; ++j) {
        H(j, j) = theta[j];
        H(keep, j) = H(j, keep) = beta * U[(size_t)(Kn - 1) * Kn + j];
      }
      Kn = keep + 1;
      inf.numiter++;
      apply(K.V(Kn - 1), K.w);
      PhaseTimer pt(c, PH_KRYLOV);
      K.orthogonalize_async(Kn, sl);
    } else {
      // the step queued ahead made V(Kn) = w / beta and its product
      H(Kn, Kn - 1) = H(Kn - 1, Kn) = beta;
      ++Kn;
      sl ^= 1;
    }
  }
}

namespace {

double krylov_eig(tn_env* e, double* phi, int kdim, int maxiter, double tol, int prec, tn_krylov_info* info) {
  const Dom D = domain(e);
  return krylov_eig_op(
      e->ctx, e->cplx(), D.n, [&](const double* x, double* y) { apply_op(e, x, y, prec); }, phi, kdim, maxiter,
      tol, info);
}

// exp(t A) x0 by a Lanczos exponential integrator (src/base/solver.jl:80)
void krylov_expm(tn_env* e, double* phi, cplx t, int kdim, int maxiter, double tol, int prec,
                 tn_krylov_info* info) {
  TN_REQUIRE(kdim >= 1 && kdim <= 60, TN_E_INVALID, "exp_solver: bad krylovdim");
  TN_REQUIRE(e->cplx() || t.imag() == 0.0, TN_E_INVALID,
             "exp_solver: complex time step needs complex (TN_C128) tensors");
  const Dom D = domain(e);
  Krylov K(e, D.n, kdim + 1);
  tn_krylov_info inf{0, 0, 0, 0.0};
  double remaining = 1.0;  // fraction of t still to go
  double last = 1.0;       // last accepted sub-step
  std::vector<cplx> h;
  double r2;
  while (remaining > 0.0 && inf.numiter < maxiter) {
    inf.numiter++;
    double nx;
    {
      PhaseTimer pt(e->ctx, PH_KRYLOV);
      nx = std::sqrt(K.norm2(phi));
    }
    if (nx == 0.0) break;
    {
      PhaseTimer pt(e->ctx, PH_KRYLOV);
      scaled_copy(K.c, K.V(0), phi, 1.0 / nx, K.nd);
    }
    std::vector<double> Hm((size_t)(kdim + 1) * (kdim + 1), 0.0);
    auto H = [&](int i, int j) -> double& { return Hm[(size_t)i * (kdim + 1) + j]; };
    // Kn basis vectors done; the step of V(Kn) is queued in slot sl, and the
    // one after it is queued ahead while this one is read (discarded if this
    // step converges)
    int Kn = 0, sl = 0;
    apply_op(e, K.V(0), K.w, prec);
    {
      PhaseTimer pt(e->ctx, PH_KRYLOV);
      K.orthogonalize_async(1, sl);
    }
    while (true) {
      if (Kn + 1 < kdim) {
        {
          PhaseTimer pt(e->ctx, PH_KRYLOV);
          K.next_from_w(Kn + 1);
        }
        apply_op(e, K.V(Kn + 1), K.w, prec);
        PhaseTimer pt(e->ctx, PH_KRYLOV);
        K.orthogonalize_async(Kn + 2, sl ^ 1);
      }
      {
        PhaseTimer pt(e->ctx, PH_KRYLOV);
        K.orthogonalize_wait(Kn + 1, sl, h, r2);
      }
      inf.numops++;
      H(Kn, Kn) = h[Kn].real();
      TN_REQUIRE(std::isfinite(r2) && std::isfinite(h[Kn].real()), TN_E_NUMERIC,
                 "exp_solver: non-finite product or residual");
      ++Kn;
      const double beta = std::sqrt(std::max(0.0, r2));
      std::vector<double> Hs((size_t)Kn * Kn), th, U;
      double hmax = 1.0;
      for (int i = 0; i < Kn; ++i)
        for (int j = 0; j < Kn; ++j) {
          Hs[(size_t)i * Kn + j] = H(i, j);
          hmax = std::max(hmax, std::fabs(H(i, j)));
        }
      sym_eig(Kn, Hs, th, U);
      auto expv = [&](cplx tau) {
        std::vector<cplx> y(Kn, 0.0);
        for (int j = 0; j < Kn; ++j) {
          const cplx f = std::exp(tau * th[j]) * U[(size_t)0 * Kn + j];
          for (int i = 0; i < Kn; ++i) y[i] += U[(size_t)i * Kn + j] * f;
        }
        return y;
      };
      const cplx tau = remaining * t;
      std::vector<cplx> y = expv(tau);
      const double err = std::abs(beta * y[Kn - 1]) * nx;
      inf.residual = err;
      if (err <= tol * remaining || beta <= 1e-14 * hmax) {
        PhaseTimer pt(e->ctx, PH_KRYLOV);
        for (auto& v : y) v *= nx;
        K.combine(Kn, y, phi);  // waits for the step queued ahead, then drops it
        remaining = 0.0;
        break;
      }
      if (Kn == kdim) {
        // largest sub-step (halving from twice the last accepted one) whose
        // error estimate meets the per-unit-time tolerance
        double frac = std::min(remaining, 2.0 * last);
        std::vector<cplx> yy;
        while (true) {
          yy = expv(frac * t);
          if (std::abs(beta * yy[Kn - 1]) * nx <= tol * frac || frac <= 1e-12) break;
          frac *= 0.5;
        }
        PhaseTimer pt(e->ctx, PH_KRYLOV);
        for (auto& v : yy) v *= nx;
        K.combine(Kn, yy, phi);
        remaining -= frac;
        last = frac;
        break;
      }
      // the step queued ahead made V(Kn) = w / beta and its product
      H(Kn, Kn - 1) = H(Kn - 1, Kn) = beta;
      sl ^= 1;
    }
  }
  inf.converged = remaining == 0.0 ? 1 : 0;
  if (info) *info = inf;
}

// <phi|H|phi> with one FP64 matvec
cplx rayleigh(tn_env* e, const double* phi) {
  const Dom D = domain(e);
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const long nd = cx ? 2 * D.n : D.n;
  double* y = (double*)c->workspace(WS_KRY_TMP, (size_t)nd * sizeof(double));
  matvec(e, phi, y, TN_PREC_FP64);
  double* co = (double*)c->workspace(WS_MISC, 64 * sizeof(double));
  multi_dot(c, phi, nd, 1, y, D.n, cx, co);
  multi_dot(c, phi, nd, 1, phi, D.n, cx, co + 4);
  double* h = c->pinned;
  TN_HIP(hipMemcpyAsync(h, co, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
  return cplx(h[0], cx ? h[1] : 0.0) / h[4];
}

// Energy of a reduced-precision eig_solver result phi (Ritz value th):
// refine_fp64 = 1 takes the Rayleigh quotient of one FP64 product; k >= 2
// continues with an FP64 Lanczos of dimension k started from phi (one
// restart cycle, KrylovKit's Ritz-value semantics, solver.jl:36-42), which
// also moves phi to that FP64 Ritz vector.
double refine_energy(tn_env* e, double* phi, const tn_update_params* p, tn_krylov_info* info, double th) {
  if (!p->refine_fp64 || p->precision == TN_PREC_FP64) return th;
  if (p->refine_fp64 == 1) return rayleigh(e, phi).real();
  tn_krylov_info ri{};
  const int rp = p->refine_precision == TN_PREC_FP32 ? TN_PREC_FP32 : TN_PREC_FP64;
  const double th64 = krylov_eig(e, phi, p->refine_fp64, 1, p->tol, rp, &ri);
  if (info) {
    info->numops += ri.numops;
    info->residual = ri.residual;
  }
  return th64;
}

// phi = A_pos * A_{pos+1}
void contract_phi(tn_env* e, int pos, void* phi) {
  TN_REQUIRE(pos >= 1 && pos < e->N, TN_E_POSITION, "contract_phi: bond out of range");
  const SiteT& a = e->A[pos - 1];
  const SiteT& b = e->A[pos];
  TN_REQUIRE(a.chir == b.chil, TN_E_SHAPE, "contract_phi: bond dimensions disagree");
  GemmDesc g;
  g.M = a.chil * e->d[pos - 1]; g.N = e->d[pos] * b.chir; g.K = a.chir;
  g.A = a.buf.p; g.lda = a.chir;
  g.B = b.buf.p; g.ldb = e->d[pos] * b.chir;
  g.C = phi; g.ldc = e->d[pos] * b.chir;
  gemm(e->ctx, gemm_prec_full(e), g);
}

// order of the p eigenvalues, largest first (stable), and the weights P of
// the first r = rank-bound of them (negative rounding residue clipped)
void sorted_weights(const std::vector<double>& w, long r, std::vector<int>& order, std::vector<double>& P) {
  const long p = (long)w.size();
  order.resize(p);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x] > w[y]; });
  P.resize(r);
  for (long i = 0; i < r; ++i) P[i] = std::max(0.0, w[order[i]]);
}

// Factor X of the density-matrix perturbation noiseterm(PH, phi, ortho) of a
// two-site update at bond b (ITensorMPS 0.3 noiseterm(::AbstractProjMPO),
// called at src/mps/update_site.jl:60-62; the in-tree mirror for coupling
// models is src/mps/projcouplingmodel.jl:391-492): the operator applied up to
// the middle MPO bond w,
//   left:  X[(a' s'), (w t b)] = sum L[w0][a'][a] W_b[w0, s', s, w] phi[a, s, t, b]
//   right: X[(w a s), (t' b')] = sum phi[a, s, t, b] W_{b+1}[w, t', t, w2] R[w2][b][b']
// and the perturbation is X X^H (left) or X^H X (right) in the convention of
// gram_eig.  Needs the projected operator positioned at b with nsite = 2.
const void* noise_factor(tn_env* e, int b, const double* phi, bool left, long* kx, long* ldx) {
  tn_ctx* c = e->ctx;
  TN_REQUIRE(e->nsite == 2 && e->lpos == b - 1 && e->rpos == b + 2, TN_E_POSITION,
             "noise term only defined for 2-site ProjMPO positioned at the bond");
  const bool cx = e->cplx();
  const long m2 = cx ? 2 : 1;
  const Prec P = gemm_prec_full(e);
  const long chil = e->A[b - 1].chil, chir = e->A[b].chir;
  const long d1 = e->d[b - 1], d2 = e->d[b];
  const long w0 = e->w[b - 1], wm = e->w[b], w2 = e->w[b + 1];
  const size_t es = e->esz;
  if (left) {
    const EnvT& L = e->LR[b - 1];
    TN_REQUIRE(L.chi == chil && L.w == w0, TN_E_SHAPE, "noise term: left environment does not match phi");
    void* X1 = c->workspace(WS_ENV_X, (size_t)(w0 * chil * d1 * d2 * chir) * es);
    void* X2 = c->workspace(WS_ENV_Y, (size_t)(chil * d1 * wm * d2 * chir) * es);
    GemmDesc g;  // X1[(w0 a'), (s t b)] = L[(w0 a'), a] phi[a, (s t b)]
    g.M = w0 * chil; g.N = d1 * d2 * chir; g.K = chil;
    g.A = L.buf.p; g.lda = chil;
    g.B = phi; g.ldb = g.N;
    g.C = X1; g.ldc = g.N;
    gemm(c, P, g);
    WStepDesc ws;  // (w0, s) -> (w, s'), output rows (a' s'), columns (w t b)
    ws.G = (int)w0; ws.Gp = (int)wm; ws.D = (int)d1; ws.Dp = (int)d1;
    ws.nA = chil; ws.nB = d2 * chir * m2;
    ws.in = X1; ws.in_sg = chil * d1 * d2 * chir * m2; ws.in_sA = d1 * d2 * chir * m2; ws.in_sS = d2 * chir * m2;
    ws.out = X2; ws.out_sg = d2 * chir * m2; ws.out_sA = d1 * wm * d2 * chir * m2; ws.out_sS = wm * d2 * chir * m2;
    ws.op = &op_left(e, b);
    wstep(c, false, ws);
    // columns (w t b) with w slowest: the kept channels are a column range
    const long sk = e->noise_skip_l.empty() ? 0 : e->noise_skip_l[b];
    *ldx = wm * d2 * chir;
    *kx = (wm - sk) * d2 * chir;
    return (const char*)X2 + (size_t)(sk * d2 * chir) * es;
  }
  const EnvT& R = e->LR[b + 2];
  TN_REQUIRE(R.chi == chir && R.w == w2, TN_E_SHAPE, "noise term: right environment does not match phi");
  void* X1 = c->workspace(WS_ENV_X, (size_t)(w2 * chil * d1 * d2 * chir) * es);
  void* X2 = c->workspace(WS_ENV_Y, (size_t)(wm * chil * d1 * d2 * chir) * es);
  GemmDesc g;  // X1[w2][(a s t), b'] = phi[(a s t), b] R[w2][b][b']
  g.M = chil * d1 * d2; g.N = chir; g.K = chir;
  g.A = phi; g.lda = chir; g.sA = 0;
  g.B = R.buf.p; g.ldb = chir; g.sB = chir * chir;
  g.C = X1; g.ldc = chir; g.sC = chil * d1 * d2 * chir;
  g.batch = (int)w2;
  gemm(c, P, g);
  WStepDesc ws;  // (w2, t) -> (w, t'), output rows (w a s), columns (t' b')
  ws.G = (int)w2; ws.Gp = (int)wm; ws.D = (int)d2; ws.Dp = (int)d2;
  ws.nA = chil * d1; ws.nB = chir * m2;
  ws.in = X1; ws.in_sg = chil * d1 * d2 * chir * m2; ws.in_sA = d2 * chir * m2; ws.in_sS = chir * m2;
  ws.out = X2; ws.out_sg = chil * d1 * d2 * chir * m2; ws.out_sA = d2 * chir * m2; ws.out_sS = chir * m2;
  ws.op = &op_right(e, b + 1);
  wstep(c, false, ws);
  // rows (w a s) with w slowest: the kept channels are the leading rows
  const long sk = e->noise_skip_r.empty() ? 0 : e->noise_skip_r[b];
  *kx = (wm - sk) * chil * d1;
  *ldx = 0;
  return X2;
}

// replacebond!(psi, b, phi; maxdim, mindim, cutoff, ortho, normalize, eigen_perturbation)
// ortho = "left":  A_b = top-k eigenvectors of rho = phi phi^T, A_{b+1} = A_b^T phi / |.|
// ortho = "right": A_{b+1} = top-k eigenvectors of rho = phi^T phi (rows), A_b = phi A_{b+1}^T / |.|
// Same factors as U, S V^T (left) or U S, V^T (right) of the truncated SVD.
// noise != 0 adds noise * noiseterm(PH, phi, ortho) to rho (ITensors 0.9
// factorize_eigen with eigen_perturbation); the spectrum is then that of the
// perturbed rho and the second factor is normalised by its own norm.
void replacebond_qn(tn_env* e, int b, const double* phi, bool ortho_left, long maxdim, long mindim, double cutoff,
                    bool normalize, double* truncerr_out, std::vector<double>* eigs_out, long* nkept);

void replacebond(tn_env* e, int b, const double* phi, bool ortho_left, long maxdim, long mindim, double cutoff,
                 bool normalize, double* truncerr_out, std::vector<double>* eigs_out, long* nkept,
                 double noise = 0.0) {
  PhaseTimer pt(e->ctx, PH_SVD);
  TN_REQUIRE(b >= 1 && b < e->N, TN_E_POSITION, "replacebond!: bond out of range");
  if (e->nq > 0) {
    TN_REQUIRE(std::fabs(noise) <= 1e-15, TN_E_UNSUPPORTED, "QN-conserving factorisation: noise is not built");
    replacebond_qn(e, b, phi, ortho_left, maxdim, mindim, cutoff, normalize, truncerr_out, eigs_out, nkept);
    return;
  }
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const int w2 = cx ? 2 : 1;  // doubles per element
  const long chil = e->A[b - 1].chil, chir = e->A[b].chir;
  const int d1 = e->d[b - 1], d2 = e->d[b];
  const long m = chil * d1, n = (long)d2 * chir;
  GramEig ge;
  const bool perturbed = std::fabs(noise) > 1e-15;
  long kx = 0, ldx = 0;
  const void* X = perturbed ? noise_factor(e, b, phi, ortho_left, &kx, &ldx) : nullptr;
  gram_eig(c, phi, m, n, ortho_left, cx, ge, X, kx, perturbed ? noise : 0.0, ldx);
  std::vector<int> order;
  std::vector<double> P;
  // rank(phi phi^H) <= min(m, n); the perturbed density matrix has full rank p
  sorted_weights(ge.w, perturbed ? ge.p : std::min(m, n), order, P);
  const TruncRule tr = truncate_weights(P, maxdim, mindim, cutoff);
  const long k = tr.n;
  double nrm = 0.0;
  for (long i = 0; i < k; ++i) nrm += P[i];
  nrm = std::sqrt(nrm);
  const double inv_norm = (normalize && nrm > 0.0 && !perturbed) ? 1.0 / nrm : 1.0;
  int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)k * sizeof(int));
  double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)k * sizeof(double));
  std::vector<double> ones(k, 1.0);
  TN_HIP(hipMemcpyAsync(dperm, order.data(), k * sizeof(int), hipMemcpyHostToDevice, c->stream));
  TN_HIP(hipMemcpyAsync(dsc, ones.data(), k * sizeof(double), hipMemcpyHostToDevice, c->stream));
  SiteT& Sa = e->A[b - 1];
  SiteT& Sb = e->A[b];
  Sa.buf.ensure((size_t)m * k * e->esz);
  Sb.buf.ensure((size_t)k * n * e->esz);
  const double* Vt = (const double*)ge.Vt;  // V^T (real) or V^H (complex) rows
  const Prec P64 = gemm_prec_full(e);
  if (ortho_left) {
    // A_b = V_k (m x k), A_{b+1} = V_k^H phi / |.|
    if (cx) gather_rows_T_z(c, Vt, ge.ppad, dperm, dsc, k, m, Sa.buf.as<double>(), true);
    else gather_rows_T(c, Vt, ge.ppad, dperm, dsc, k, m, Sa.buf.as<double>());
    double* Vk = (double*)c->workspace(WS_SVD_GRAM, (size_t)k * m * e->esz);
    gather_rows(c, Vt, ge.ppad * w2, dperm, dsc, k, m * w2, Vk);
    GemmDesc g;
    g.M = k; g.N = n; g.K = m;
    g.A = Vk; g.lda = m;
    g.B = phi; g.ldb = n;
    g.C = Sb.buf.p; g.ldc = n;
    g.alpha = inv_norm;
    gemm(c, P64, g);
  } else {
    // A_{b+1} = V_k^H (k x n), A_b = phi V_k / |.|
    gather_rows(c, Vt, ge.ppad * w2, dperm, dsc, k, n * w2, Sb.buf.as<double>());
    GemmDesc g;
    g.M = m; g.N = k; g.K = n;
    g.A = phi; g.lda = n;
    g.B = Sb.buf.p; g.ldb = n; g.tB = 1; g.cB = cx;
    g.C = Sa.buf.p; g.ldc = k;
    g.alpha = inv_norm;
    gemm(c, P64, g);
  }
  if (perturbed && normalize) {  // the second factor is no longer sqrt(sum P) in norm
    double* F = ortho_left ? Sb.buf.as<double>() : Sa.buf.as<double>();
    const long nel = ortho_left ? k * n : m * k;
    double* co = (double*)c->workspace(WS_PROJ_C, 64 * sizeof(double));
    multi_dot(c, F, nel * w2, 1, F, nel, cx, co);
    TN_HIP(hipMemcpyAsync(c->pinned, co, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    const double f2 = c->pinned[0];
    if (f2 > 0.0) scale_f64(c, F, 1.0 / std::sqrt(f2), nel * w2);
  }
  TN_HIP(hipStreamSynchronize(c->stream));  // host vectors above go out of scope
  Sa.chir = k;
  Sb.chil = k;
  invalidate_site(e, b);
  invalidate_site(e, b + 1);
  e->ortho = ortho_left ? b + 1 : b;
  if (truncerr_out) *truncerr_out = tr.truncerr;
  if (eigs_out) eigs_out->assign(P.begin(), P.begin() + k);
  if (nkept) *nkept = k;
}

// replacebond! of a QN-conserving two-site tensor (NDTensors BlockSparse
// semantics, upstream ITensors/NDTensors; called at update_site.jl:64-76):
// the density matrix of the kept side is block diagonal in the QN of its
// combined (link, site) index.  Rows of M (ortho left; columns, through the
// transpose, for ortho right) are gathered sector by sector; blocks up to 32
// rows are formed and diagonalised in one launch (one wavefront each), larger
// ones one by one (gram_eig: divide and conquer).  All weights go through
// truncate!!; each block keeps its weights above docut (whole multiplets split
// across blocks are dropped together, exact zeros go).  The new link lists
// the kept sectors in ascending QN order, weights descending within each.
void replacebond_qn(tn_env* e, int b, const double* phi, bool ortho_left, long maxdim, long mindim, double cutoff,
                    bool normalize, double* truncerr_out, std::vector<double>* eigs_out, long* nkept) {
  tn_ctx* c = e->ctx;
  TN_REQUIRE(!e->cplx(), TN_E_UNSUPPORTED, "QN-conserving factorisation: real tensors only");
  const long chil = e->A[b - 1].chil, chir = e->A[b].chir;
  const int d1 = e->d[b - 1], d2 = e->d[b];
  const long m = chil * d1, n = (long)d2 * chir;
  using Q = std::array<int, 2>;
  const auto& ql = e->link_qn[b - 1];
  const auto& qr = e->link_qn[b + 1];
  const auto& q1 = e->site_qn[b - 1];
  const auto& q2 = e->site_qn[b];
  TN_REQUIRE((long)ql.size() == chil && (long)qr.size() == chir, TN_E_SHAPE,
             "QN-conserving factorisation: link QNs do not match the bond dimensions");
  // QN of every row of G (G = M for ortho left, M^T for ortho right)
  const long p = ortho_left ? m : n, L = ortho_left ? n : m;
  std::vector<Q> qrow(p);
  if (ortho_left) {
    for (long a = 0; a < chil; ++a)
      for (int s = 0; s < d1; ++s) qrow[a * d1 + s] = Q{ql[a][0] + q1[s][0], ql[a][1] + q1[s][1]};
  } else {
    for (int t = 0; t < d2; ++t)
      for (long k = 0; k < chir; ++k) qrow[t * chir + k] = Q{qr[k][0] - q2[t][0], qr[k][1] - q2[t][1]};
  }
  std::map<Q, std::vector<int>> rows;
  for (long r = 0; r < p; ++r) rows[qrow[r]].push_back((int)r);
  const int ns = (int)rows.size();
  std::vector<Q> sq;
  std::vector<int> perm, roff, psz;
  perm.reserve(p);
  for (auto& kv : rows) {
    sq.push_back(kv.first);
    roff.push_back((int)perm.size());
    psz.push_back((int)kv.second.size());
    perm.insert(perm.end(), kv.second.begin(), kv.second.end());
  }
  // eigenvalue / eigenvector slots of every block
  std::vector<long> woff(ns), voff(ns);
  long wtot = 0, vtot = 0;
  for (int i = 0; i < ns; ++i) {
    woff[i] = wtot;
    voff[i] = vtot;
    wtot += psz[i];
    vtot += (long)psz[i] * psz[i];
  }
  TN_REQUIRE(vtot < (1L << 31), TN_E_UNSUPPORTED, "QN-conserving factorisation: blocks too large");
  // G gathered by sector: Gp[i] = G[perm[i]] (p x L)
  double* Gp = (double*)c->workspace(WS_QN_M, (size_t)p * L * sizeof(double));
  int* ibuf = (int*)c->workspace(WS_QN_IDX, (size_t)(p + 4 * std::max<long>(ns, p) + 16) * sizeof(int));
  int* dperm = ibuf;
  double* dones = (double*)c->workspace(WS_SVD_SCALE, (size_t)p * sizeof(double));
  fill_f64(c, dones, 1.0, p);
  TN_HIP(hipMemcpyAsync(dperm, perm.data(), p * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const double* Gsrc = phi;
  if (!ortho_left) {
    double* Mt = (double*)c->workspace(WS_QN_T, (size_t)m * n * sizeof(double));
    transpose_f64(c, phi, m, n, Mt);
    Gsrc = Mt;
  }
  gather_rows(c, Gsrc, L, dperm, dones, p, L, Gp);
  double* Wd = (double*)c->workspace(WS_QN_W, (size_t)wtot * sizeof(double));
  double* Vd = (double*)c->workspace(WS_QN_V, (size_t)std::max<long>(vtot, 1) * sizeof(double));
  // Blocks of a size class (up to 32, 64, 96 rows) go to one batched Jacobi
  // launch when the block-size histogram says it is cheaper than the dense
  // eigensolver block by block: measured on MI355X, a launch of class c
  // takes about qn_batch_ms[c] (set by its largest block, all blocks run
  // side by side) and one block of p rows through the dense eigensolver
  // about 0.35 + 0.005 p ms (launch-latency bound below 136 rows).  The
  // batched blocks' rho_q = G_q G_q^T come from one block-sparse GEMM
  // launch into a stack (ld QN_RHO_LD); the rest use the dense eigensolver.
  static constexpr double qn_batch_ms[3] = {0.25, 1.2, 4.0};
  double serial_ms[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < ns; ++i)
    if (psz[i] <= QN_JACOBI_MAX) serial_ms[(psz[i] + 31) / 32 - 1] += (psz[i] < 3 ? 0.05 : 0.35 + 0.005 * psz[i]);
  bool batch_cls[3];
  for (int cl = 0; cl < 3; ++cl) batch_cls[cl] = serial_ms[cl] > qn_batch_ms[cl];
  std::vector<long> soff(ns, -1);
  long srows = 0;
  for (int i = 0; i < ns; ++i)
    if (psz[i] <= QN_JACOBI_MAX && batch_cls[(psz[i] + 31) / 32 - 1]) {
      soff[i] = srows;
      srows += psz[i];
    }
  std::vector<BsTask> gt;
  std::vector<BsSeg> gs;
  std::vector<int4> gtl;
  std::vector<int> meta;
  std::vector<char> host;  // staging of the block plan, alive until the stream is synchronised
  int cls_n[3] = {0, 0, 0};
  if (srows > 0) {
    double* Rst = (double*)c->workspace(WS_QN_RHO, (size_t)srows * QN_RHO_LD * sizeof(double));
    for (int i = 0; i < ns; ++i) {
      if (soff[i] < 0) continue;
      BsTask t;
      t.c_off = soff[i] * QN_RHO_LD;
      t.M = t.N = psz[i];
      t.seg0 = (int)gs.size();
      t.nseg = 1;
      gs.push_back(BsSeg{(long)roff[i] * L, (long)roff[i] * L, L});
      push_tiles(gtl, (int)gt.size(), t.M, t.N);
      gt.push_back(t);
    }
    pad_tiles(gtl);
    // size classes 32, 64, 96: metadata of class 0, then 1, then 2
    for (int cl = 0; cl < 3; ++cl)
      for (int i = 0; i < ns; ++i) {
        if (soff[i] < 0 || (psz[i] + 31) / 32 - 1 != cl) continue;
        meta.insert(meta.end(), {(int)soff[i], psz[i], (int)woff[i], (int)voff[i]});
        ++cls_n[cl];
      }
    auto al = [](size_t x) { return (x + 15) / 16 * 16; };
    const size_t b_t = al(gt.size() * sizeof(BsTask)), b_s = al(gs.size() * sizeof(BsSeg));
    const size_t b_l = al(gtl.size() * sizeof(int4)), b_m = al(meta.size() * sizeof(int));
    host.assign(b_t + b_s + b_l + b_m, 0);
    std::memcpy(host.data(), gt.data(), gt.size() * sizeof(BsTask));
    std::memcpy(host.data() + b_t, gs.data(), gs.size() * sizeof(BsSeg));
    std::memcpy(host.data() + b_t + b_s, gtl.data(), gtl.size() * sizeof(int4));
    std::memcpy(host.data() + b_t + b_s + b_l, meta.data(), meta.size() * sizeof(int));
    char* dplan = (char*)c->workspace(WS_QN_PLAN, host.size());
    TN_HIP(hipMemcpyAsync(dplan, host.data(), host.size(), hipMemcpyHostToDevice, c->stream));
    BsDesc g;
    g.A = Gp; g.lda = L;
    g.B = Gp; g.ldb = L; g.tB = 1;
    g.C = Rst; g.ldc = QN_RHO_LD;
    g.tasks = (const BsTask*)dplan;
    g.segs = (const BsSeg*)(dplan + b_t);
    g.tiles = (const int4*)(dplan + b_t + b_s);
    g.ntiles = (long)gtl.size();
    gemm_blocksparse(c, Prec::F64, g);
    const int* dm = (const int*)(dplan + b_t + b_s + b_l);
    for (int cl = 0, o = 0; cl < 3; o += cls_n[cl], ++cl)
      qn_block_eig(c, Rst, dm + 4 * o, cls_n[cl], 32 * (cl + 1), Wd, Vd);
  }
  std::vector<std::vector<double>> w(ns);
  // large blocks: the dense eigensolver, one block at a time
  for (int i = 0; i < ns; ++i) {
    if (soff[i] >= 0) continue;
    GramEig ge;
    gram_eig(c, Gp + (long)roff[i] * L, psz[i], L, true, false, ge);
    w[i] = ge.w;
    TN_HIP(hipMemcpy2DAsync(Vd + voff[i], (size_t)psz[i] * sizeof(double), ge.Vt, (size_t)ge.ppad * sizeof(double),
                            (size_t)psz[i] * sizeof(double), (size_t)psz[i], hipMemcpyDeviceToDevice, c->stream));
  }
  if (srows > 0) {
    std::vector<double> hw(wtot);
    TN_HIP(hipMemcpyAsync(hw.data(), Wd, wtot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < ns; ++i)
      if (soff[i] >= 0) w[i].assign(hw.begin() + woff[i], hw.begin() + woff[i] + psz[i]);
  }
  // weights of every block in descending order, then NDTensors' block rule
  std::vector<std::vector<int>> ord(ns);
  std::vector<double> allP;
  allP.reserve(p);
  for (int i = 0; i < ns; ++i) {
    ord[i].resize(psz[i]);
    std::iota(ord[i].begin(), ord[i].end(), 0);
    std::stable_sort(ord[i].begin(), ord[i].end(), [&](int x, int y) { return w[i][x] > w[i][y]; });
    for (int x : ord[i]) allP.push_back(std::max(w[i][x], 0.0));
  }
  std::sort(allP.begin(), allP.end(), std::greater<double>());
  const TruncRule tr = truncate_weights(allP, maxdim, mindim, cutoff);
  std::vector<int> col;      // kept columns: {V offset, row offset, eigenvector, block size}
  std::vector<double> kept;
  std::vector<Q> newq;
  for (int i = 0; i < ns; ++i) {
    for (int j = 0; j < psz[i]; ++j) {
      const double x = w[i][ord[i][j]];
      if (!(x > tr.docut)) break;
      col.insert(col.end(), {(int)voff[i], roff[i], ord[i][j], psz[i]});
      kept.push_back(x);
      newq.push_back(sq[i]);
    }
  }
  const long k = (long)kept.size();
  TN_REQUIRE(k >= 1, TN_E_NUMERIC, "QN-conserving factorisation: no weight above the truncation threshold");
  double nrm = 0.0;
  for (double x : kept) nrm += x;
  nrm = std::sqrt(nrm);
  const double inv_norm = (normalize && nrm > 0.0) ? 1.0 / nrm : 1.0;
  int* dcol = (int*)c->workspace(WS_SVD_PERM, (size_t)4 * k * sizeof(int));
  TN_HIP(hipMemcpyAsync(dcol, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  SiteT& Sa = e->A[b - 1];
  SiteT& Sb = e->A[b];
  Sa.buf.ensure((size_t)m * k * e->esz);
  Sb.buf.ensure((size_t)k * n * e->esz);
  if (ortho_left) {
    // A_b = U (m x k, zero outside each column's sector), A_{b+1} = U^T phi / |.|
    fill_f64(c, Sa.buf.as<double>(), 0.0, m * k);
    qn_scatter(c, Vd, dcol, dperm, (int)k, k, false, Sa.buf.as<double>());
    GemmDesc g;
    g.M = k; g.N = n; g.K = m;
    g.A = Sa.buf.p; g.lda = k; g.tA = 1;
    g.B = phi; g.ldb = n;
    g.C = Sb.buf.p; g.ldc = n;
    g.alpha = inv_norm;
    gemm(c, Prec::F64, g);
  } else {
    // A_{b+1} = U^T (k x n), A_b = phi U / |.|
    fill_f64(c, Sb.buf.as<double>(), 0.0, k * n);
    qn_scatter(c, Vd, dcol, dperm, (int)k, n, true, Sb.buf.as<double>());
    GemmDesc g;
    g.M = m; g.N = k; g.K = n;
    g.A = phi; g.lda = n;
    g.B = Sb.buf.p; g.ldb = n; g.tB = 1;
    g.C = Sa.buf.p; g.ldc = k;
    g.alpha = inv_norm;
    gemm(c, Prec::F64, g);
  }
  TN_HIP(hipStreamSynchronize(c->stream));  // host index arrays above go out of scope
  Sa.chir = k;
  Sb.chil = k;
  e->link_qn[b] = newq;
  ++e->link_ver[b];
  invalidate_site(e, b);
  invalidate_site(e, b + 1);
  e->ortho = ortho_left ? b + 1 : b;
  if (truncerr_out) *truncerr_out = tr.truncerr;
  if (eigs_out) {
    std::sort(kept.begin(), kept.end(), std::greater<double>());
    eigs_out->swap(kept);
  }
  if (nkept) *nkept = k;
}

// orthogonalize!(psi, site): exact gauge move.  Each step keeps the full row
// (column) space of the site matrix: the top min(m, n) eigenvectors of its
// Gram matrix become the orthonormal factor, the exact remainder moves on.
void orthogonalize(tn_env* e, int site) {
  TN_REQUIRE(site >= 1 && site <= e->N, TN_E_POSITION, "orthogonalize!: site out of range");
  TN_REQUIRE(e->nq == 0 || e->ortho == site, TN_E_UNSUPPORTED,
             "QN-conserving environment: the orthocenter moves only through two-site updates "
             "(start from a right-orthonormal state with orthocenter 1)");
  if (e->nq > 0) return;
  require_sites(e);
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const int w2 = cx ? 2 : 1;
  const size_t es = e->esz;
  const Prec P64 = gemm_prec_full(e);
  auto upload_perm = [&](const std::vector<int>& order, long r, int*& dperm, double*& dsc) {
    dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)r * sizeof(int));
    dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)r * sizeof(double));
    std::vector<double> ones(r, 1.0);
    TN_HIP(hipMemcpyAsync(dperm, order.data(), r * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipMemcpyAsync(dsc, ones.data(), r * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
  };
  // right part: sites N..site+1 become right-orthonormal
  for (int k = e->N; k > site; --k) {
    SiteT& S = e->A[k - 1];
    SiteT& Pv = e->A[k - 2];
    const long m = S.chil, n = (long)e->d[k - 1] * S.chir;
    GramEig ge;
    gram_eig(c, S.buf.p, m, n, false, cx, ge);
    std::vector<int> order;
    std::vector<double> P;
    const long r = std::min(m, n);
    sorted_weights(ge.w, r, order, P);
    int* dperm;
    double* dsc;
    upload_perm(order, r, dperm, dsc);
    DevBuf newA;
    newA.ensure((size_t)r * n * es);
    gather_rows(c, (const double*)ge.Vt, ge.ppad * w2, dperm, dsc, r, n * w2, newA.as<double>());  // V_r^H
    double* SV = (double*)c->workspace(WS_PHI, (size_t)m * r * es);
    GemmDesc g1;
    g1.M = m; g1.N = r; g1.K = n;
    g1.A = S.buf.p; g1.lda = n;
    g1.B = newA.p; g1.ldb = n; g1.tB = 1; g1.cB = cx;
    g1.C = SV; g1.ldc = r;
    gemm(c, P64, g1);                                                    // S V_r (m x r)
    const long pm = Pv.chil * e->d[k - 2];
    DevBuf newP;
    newP.ensure((size_t)pm * r * es);
    GemmDesc g2;
    g2.M = pm; g2.N = r; g2.K = m;
    g2.A = Pv.buf.p; g2.lda = m;
    g2.B = SV; g2.ldb = r;
    g2.C = newP.p; g2.ldc = r;
    gemm(c, P64, g2);                                                    // A_{k-1} S V_r
    TN_HIP(hipStreamSynchronize(c->stream));
    S.buf = std::move(newA);
    S.chil = r;
    Pv.buf = std::move(newP);
    Pv.chir = r;
    invalidate_site(e, k);
    invalidate_site(e, k - 1);
  }
  // left part: sites 1..site-1 become left-orthonormal
  for (int k = 1; k < site; ++k) {
    SiteT& S = e->A[k - 1];
    SiteT& Nx = e->A[k];
    const long m = S.chil * e->d[k - 1], n = S.chir;
    GramEig ge;
    gram_eig(c, S.buf.p, m, n, true, cx, ge);
    std::vector<int> order;
    std::vector<double> P;
    const long r = std::min(m, n);
    sorted_weights(ge.w, r, order, P);
    int* dperm;
    double* dsc;
    upload_perm(order, r, dperm, dsc);
    DevBuf newA;
    newA.ensure((size_t)m * r * es);
    if (cx) gather_rows_T_z(c, (const double*)ge.Vt, ge.ppad, dperm, dsc, r, m, newA.as<double>(), true);
    else gather_rows_T(c, (const double*)ge.Vt, ge.ppad, dperm, dsc, r, m, newA.as<double>());  // U_r
    double* US = (double*)c->workspace(WS_PHI, (size_t)r * n * es);
    GemmDesc g1;
    g1.M = r; g1.N = n; g1.K = m;
    g1.A = newA.p; g1.lda = r; g1.tA = 1; g1.cA = cx;
    g1.B = S.buf.p; g1.ldb = n;
    g1.C = US; g1.ldc = n;
    gemm(c, P64, g1);                                                    // U_r^H S (r x n)
    const long nn = (long)e->d[k] * Nx.chir;
    DevBuf newN;
    newN.ensure((size_t)r * nn * es);
    GemmDesc g2;
    g2.M = r; g2.N = nn; g2.K = n;
    g2.A = US; g2.lda = n;
    g2.B = Nx.buf.p; g2.ldb = nn;
    g2.C = newN.p; g2.ldc = nn;
    gemm(c, P64, g2);
    TN_HIP(hipStreamSynchronize(c->stream));
    S.buf = std::move(newA);
    S.chir = r;
    Nx.buf = std::move(newN);
    Nx.chil = r;
    invalidate_site(e, k);
    invalidate_site(e, k + 1);
  }
  e->ortho = site;
}


// _update_two_site! (src/mps/update_site.jl:27-90)
void update_two_site(tn_env* e, int pos, bool ortho_left, const tn_update_params* p, tn_update_result* r,
                     std::vector<double>* eigs_out = nullptr) {
  TN_REQUIRE(p != nullptr, TN_E_INVALID, "null params");
  TN_REQUIRE(pos > 0 && pos < e->N, TN_E_POSITION, "update_position!: bond out of range");
  TN_REQUIRE(!(p->solver == 0 && p->reverse_step), TN_E_INVALID,
             "update_position!: reverse_step needs a time step (exp_solver)");
  TN_REQUIRE(!(std::fabs(p->noise) > 1e-15 && p->reverse_step), TN_E_INVALID,
             "`updatePosition()` :: `noise` cannot be greater than zero for `reverse_step=true` !!");
  TN_REQUIRE((ortho_left && e->ortho == pos) || (!ortho_left && e->ortho == pos + 1), TN_E_POSITION,
             "update_position!: orthocenter is " + std::to_string(e->ortho) + ", expected " +
                 std::to_string(ortho_left ? pos : pos + 1));
  tn_ctx* c = e->ctx;
  e->nsite = 2;
  const long chil = e->A[pos - 1].chil, chir = e->A[pos].chir;
  const long n = chil * e->d[pos - 1] * e->d[pos] * chir;
  const long nd = e->cplx() ? 2 * n : n;
  double* phi = (double*)c->workspace(WS_PHI, (size_t)nd * sizeof(double));
  contract_phi(e, pos, phi);
  position(e, pos);
  tn_update_result res{};
  if (r) { res.eigs = r->eigs; res.eigs_cap = r->eigs_cap; }
  cplx energy(NAN, 0.0);
  const cplx t(p->time_step_re, p->time_step_im);
  if (p->solver == 0) {
    const double th = krylov_eig(e, phi, p->krylovdim, p->maxiter, p->tol, p->precision, &res.info);
    if (p->check_convergence && res.info.converged < 1)
      throw Error(TN_E_NOTCONV, "`eig_solver()` not converged !!");
    energy = refine_energy(e, phi, p, &res.info, th);
  } else {
    krylov_expm(e, phi, t, p->krylovdim, p->maxiter, p->tol, p->precision, &res.info);
    if (p->check_convergence && res.info.converged < 1)
      throw Error(TN_E_NOTCONV, "`eig_solver()` not converged !!");
  }
  if (p->normalize) {
    Krylov K(e, n, 1);
    const double nr = std::sqrt(K.norm2(phi));
    if (nr > 0) scale_f64(c, phi, 1.0 / nr, nd);
  }
  if (std::isnan(energy.real())) energy = rayleigh(e, phi);
  double terr = 0.0;
  long nk = 0;
  std::vector<double> eigs;
  replacebond(e, pos, phi, ortho_left, p->maxdim, p->mindim, p->cutoff, p->normalize != 0, &terr, &eigs, &nk,
              p->noise);
  const bool done = (pos == 1 && !ortho_left) || (pos == e->N - 1 && ortho_left);
  if (p->reverse_step && !done) {
    const int pos1 = ortho_left ? pos + 1 : pos;
    SiteT& S = e->A[pos1 - 1];
    const long n1 = S.chil * e->d[pos1 - 1] * S.chir;
    const long nd1 = e->cplx() ? 2 * n1 : n1;
    double* phi0 = (double*)c->workspace(WS_PHI, (size_t)nd1 * sizeof(double));
    copy_f64(c, phi0, S.buf.as<double>(), nd1);
    e->nsite = 1;
    position(e, pos1);
    krylov_expm(e, phi0, -t, p->krylovdim, p->maxiter, p->tol, p->precision, nullptr);
    if (p->normalize) {
      Krylov K(e, n1, 1);
      const double nr = std::sqrt(K.norm2(phi0));
      if (nr > 0) scale_f64(c, phi0, 1.0 / nr, nd1);
    }
    energy = rayleigh(e, phi0);
    copy_f64(c, S.buf.as<double>(), phi0, nd1);
    invalidate_site(e, pos1);
  }
  res.energy = energy.real();
  res.energy_im = energy.imag();
  res.truncerr = terr;
  res.nkept = nk;
  if (res.eigs)
    for (long i = 0; i < std::min<long>(res.eigs_cap, (long)eigs.size()); ++i) res.eigs[i] = eigs[i];
  if (r) *r = res;
  if (eigs_out) eigs_out->swap(eigs);
}


// SVD split of the one-site tensor phi of site pos (update_site.jl:154-173).
// ortho left:  A_pos = U_k (chil d x k), C = S_k V_k^H (k x chir);
// ortho right: A_pos = V_k^H (k x d chir), C = U_k S_k (chil x k).
// S is normalised when `normalize`.  C is left in WS_BOND; returns k.
long split_site(tn_env* e, int pos, const double* phi, bool ortho_left, long maxdim, long mindim, double cutoff,
                bool normalize, double* truncerr_out, std::vector<double>* eigs_out, double** C_out) {
  PhaseTimer pt(e->ctx, PH_SVD);
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const int w2 = cx ? 2 : 1;
  SiteT& S = e->A[pos - 1];
  const long chil = S.chil, chir = S.chir;
  const int d = e->d[pos - 1];
  const long m = ortho_left ? chil * d : chil;
  const long n = ortho_left ? chir : (long)d * chir;
  GramEig ge;
  gram_eig(c, phi, m, n, ortho_left, cx, ge);
  std::vector<int> order;
  std::vector<double> P;
  sorted_weights(ge.w, std::min(m, n), order, P);
  const TruncRule tr = truncate_weights(P, maxdim, mindim, cutoff);
  const long k = tr.n;
  double nrm = 0.0;
  for (long i = 0; i < k; ++i) nrm += P[i];
  nrm = std::sqrt(nrm);
  const double inv_norm = (normalize && nrm > 0.0) ? 1.0 / nrm : 1.0;
  int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)k * sizeof(int));
  double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)k * sizeof(double));
  std::vector<double> ones(k, 1.0);
  TN_HIP(hipMemcpyAsync(dperm, order.data(), k * sizeof(int), hipMemcpyHostToDevice, c->stream));
  TN_HIP(hipMemcpyAsync(dsc, ones.data(), k * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const double* Vt = (const double*)ge.Vt;
  const Prec P64 = gemm_prec_full(e);
  double* C = (double*)c->workspace(WS_BOND, (size_t)(ortho_left ? k * n : m * k) * e->esz);
  S.buf.ensure((size_t)m * k * e->esz > (size_t)k * n * e->esz ? (size_t)m * k * e->esz : (size_t)k * n * e->esz);
  if (ortho_left) {
    if (cx) gather_rows_T_z(c, Vt, ge.ppad, dperm, dsc, k, m, S.buf.as<double>(), true);
    else gather_rows_T(c, Vt, ge.ppad, dperm, dsc, k, m, S.buf.as<double>());
    double* Vk = (double*)c->workspace(WS_SVD_GRAM, (size_t)k * m * e->esz);
    gather_rows(c, Vt, ge.ppad * w2, dperm, dsc, k, m * w2, Vk);
    GemmDesc g;  // C = V_k^H phi / |.|
    g.M = k; g.N = n; g.K = m;
    g.A = Vk; g.lda = m;
    g.B = phi; g.ldb = n;
    g.C = C; g.ldc = n;
    g.alpha = inv_norm;
    gemm(c, P64, g);
    S.chir = k;
  } else {
    gather_rows(c, Vt, ge.ppad * w2, dperm, dsc, k, n * w2, S.buf.as<double>());
    GemmDesc g;  // C = phi V_k / |.|
    g.M = m; g.N = k; g.K = n;
    g.A = phi; g.lda = n;
    g.B = S.buf.p; g.ldb = n; g.tB = 1; g.cB = cx;
    g.C = C; g.ldc = k;
    g.alpha = inv_norm;
    gemm(c, P64, g);
    S.chil = k;
  }
  TN_HIP(hipStreamSynchronize(c->stream));  // host vectors above go out of scope
  invalidate_site(e, pos);
  if (truncerr_out) *truncerr_out = tr.truncerr;
  if (eigs_out) eigs_out->assign(P.begin(), P.begin() + k);
  *C_out = C;
  return k;
}

// _update_one_site! (src/mps/update_site.jl:94-197)
void update_one_site(tn_env* e, int pos, bool ortho_left, const tn_update_params* p, tn_update_result* r,
                     std::vector<double>* eigs_out = nullptr) {
  TN_REQUIRE(e->nq == 0, TN_E_UNSUPPORTED, "QN-conserving environment: one-site updates are not built");
  TN_REQUIRE(p != nullptr, TN_E_INVALID, "null params");
  TN_REQUIRE(pos >= 1 && pos <= e->N, TN_E_POSITION, "update_position!: site out of range");
  TN_REQUIRE(!(p->solver == 0 && p->reverse_step), TN_E_INVALID,
             "update_position!: reverse_step needs a time step (exp_solver)");
  TN_REQUIRE(!(std::fabs(p->noise) > 1e-15 && p->reverse_step), TN_E_INVALID,
             "`updatePosition()` :: `noise` cannot be greater than zero for `reverse_step=true` !!");
  TN_REQUIRE(e->ortho == pos, TN_E_POSITION,
             "update_position!: orthocenter is " + std::to_string(e->ortho) + ", expected " + std::to_string(pos));
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  e->nsite = 1;
  SiteT& S = e->A[pos - 1];
  const long n = S.chil * e->d[pos - 1] * S.chir;
  const long nd = cx ? 2 * n : n;
  double* phi = (double*)c->workspace(WS_PHI, (size_t)nd * sizeof(double));
  copy_f64(c, phi, S.buf.as<double>(), nd);
  position(e, pos);
  tn_update_result res{};
  if (r) { res.eigs = r->eigs; res.eigs_cap = r->eigs_cap; }
  cplx energy(NAN, 0.0);
  const cplx t(p->time_step_re, p->time_step_im);
  if (p->solver == 0) {
    const double th = krylov_eig(e, phi, p->krylovdim, p->maxiter, p->tol, p->precision, &res.info);
    if (p->check_convergence && res.info.converged < 1)
      throw Error(TN_E_NOTCONV, "`eig_solver()` not converged !!");
    energy = refine_energy(e, phi, p, &res.info, th);
  } else {
    krylov_expm(e, phi, t, p->krylovdim, p->maxiter, p->tol, p->precision, &res.info);
    if (p->check_convergence && res.info.converged < 1)
      throw Error(TN_E_NOTCONV, "`eig_solver()` not converged !!");
  }
  if (p->normalize) {
    Krylov K(e, n, 1);
    const double nr = std::sqrt(K.norm2(phi));
    if (nr > 0) scale_f64(c, phi, 1.0 / nr, nd);
  }
  if (std::isnan(energy.real())) energy = rayleigh(e, phi);
  double terr = 0.0;
  long nk = 0;
  std::vector<double> eigs;
  const bool done = (pos == 1 && !ortho_left) || (pos == e->N && ortho_left);
  if (done) {
    copy_f64(c, S.buf.as<double>(), phi, nd);
    invalidate_site(e, pos);
  } else if (std::fabs(p->noise) > 1e-15) {
    // two-site density-matrix factorisation of phi and its neighbour
    const int pos0 = ortho_left ? pos : pos - 1;
    const SiteT& Sa = e->A[pos0 - 1];
    const SiteT& Sb = e->A[pos0];
    const long n2 = Sa.chil * e->d[pos0 - 1] * e->d[pos0] * Sb.chir;
    double* phi2 = (double*)c->workspace(WS_PHI2, (size_t)n2 * e->esz);
    GemmDesc g;
    g.M = Sa.chil * e->d[pos0 - 1]; g.N = (long)e->d[pos0] * Sb.chir; g.K = Sa.chir;
    g.A = ortho_left ? (const void*)phi : Sa.buf.p; g.lda = Sa.chir;
    g.B = ortho_left ? Sb.buf.p : (const void*)phi; g.ldb = (long)e->d[pos0] * Sb.chir;
    g.C = phi2; g.ldc = (long)e->d[pos0] * Sb.chir;
    gemm(c, gemm_prec_full(e), g);
    e->nsite = 2;
    position(e, pos0);
    replacebond(e, pos0, phi2, ortho_left, p->maxdim, p->mindim, p->cutoff, p->normalize != 0, &terr, &eigs, &nk,
                p->noise);
    e->nsite = 1;
  } else {
    double* C = nullptr;
    nk = split_site(e, pos, phi, ortho_left, p->maxdim, p->mindim, p->cutoff, p->normalize != 0, &terr, &eigs, &C);
    const int nxt = ortho_left ? pos + 1 : pos - 1;
    SiteT& Nx = e->A[nxt - 1];
    const long crow = ortho_left ? nk : Nx.chir;     // C is crow x ccol
    const long ccol = ortho_left ? Nx.chil : nk;
    if (p->reverse_step) {
      const int pos1 = ortho_left ? pos + 1 : pos;
      const long nb = crow * ccol;
      const long ndb = cx ? 2 * nb : nb;
      e->nsite = 0;
      position(e, pos1);
      krylov_expm(e, C, -t, p->krylovdim, p->maxiter, p->tol, TN_PREC_FP64, nullptr);
      if (p->normalize) {
        Krylov K(e, nb, 1);
        const double nr = std::sqrt(K.norm2(C));
        if (nr > 0) scale_f64(c, C, 1.0 / nr, ndb);
      }
      energy = rayleigh(e, C);
    }
    DevBuf newN;
    GemmDesc g;
    if (ortho_left) {  // A_{pos+1} <- C A_{pos+1}
      const long nn = (long)e->d[nxt - 1] * Nx.chir;
      newN.ensure((size_t)nk * nn * e->esz);
      g.M = nk; g.N = nn; g.K = Nx.chil;
      g.A = C; g.lda = Nx.chil;
      g.B = Nx.buf.p; g.ldb = nn;
      g.C = newN.p; g.ldc = nn;
    } else {           // A_{pos-1} <- A_{pos-1} C
      const long mm = Nx.chil * e->d[nxt - 1];
      newN.ensure((size_t)mm * nk * e->esz);
      g.M = mm; g.N = nk; g.K = Nx.chir;
      g.A = Nx.buf.p; g.lda = Nx.chir;
      g.B = C; g.ldb = nk;
      g.C = newN.p; g.ldc = nk;
    }
    gemm(c, gemm_prec_full(e), g);
    TN_HIP(hipStreamSynchronize(c->stream));
    Nx.buf = std::move(newN);
    if (ortho_left) Nx.chil = nk;
    else Nx.chir = nk;
    invalidate_site(e, nxt);
    e->ortho = nxt;
  }
  res.energy = energy.real();
  res.energy_im = energy.imag();
  res.truncerr = terr;
  res.nkept = nk;
  if (res.eigs)
    for (long i = 0; i < std::min<long>(res.eigs_cap, (long)eigs.size()); ++i) res.eigs[i] = eigs[i];
  if (r) *r = res;
  if (eigs_out) eigs_out->swap(eigs);
}

// ---------------------------------------------------------------------------
// global subspace expansion (krylov_extend!, src/mps/sweep.jl:378-534)
// ---------------------------------------------------------------------------
using Sites = std::vector<SiteT>;

struct EnvOwner {
  tn_env* e = nullptr;
  ~EnvOwner() {
    if (e) tn_env_destroy(e);
  }
};

void copy_sites(tn_env* e, const Sites& from, Sites& to) {
  to.resize(from.size());
  for (size_t k = 0; k < from.size(); ++k) {
    const long n = from[k].chil * e->d[k] * from[k].chir;
    to[k].buf.ensure((size_t)n * e->esz);
    copy_f64(e->ctx, to[k].buf.as<double>(), from[k].buf.as<double>(), e->cplx() ? 2 * n : n);
    to[k].chil = from[k].chil;
    to[k].chir = from[k].chir;
  }
}

// H psi by ITensorMPS's "densitymatrix" contraction: e2 holds the MPO
// (H^dag H)[(w w'), s', s, (v v')] = sum_t W[w, t, s, v] conj(W[w', t, s', v'])
// whose left environments are those of <H psi|H psi>; from the right, the
// reduced density matrix of H psi on (site j, kept right basis) gives the
// right-orthonormal site j (its dominant eigenvectors, NDTensors truncation
// with maxdim / cutoff), the first site keeps the rest.
void apply_mpo(tn_env* e, tn_env* e2, const Sites& in, long maxdim, double cutoff, Sites& out) {
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const int m2 = cx ? 2 : 1;
  const Prec P64 = gemm_prec_full(e);
  const int N = e->N;
  copy_sites(e, in, e2->A);
  for (int k = 1; k <= N; ++k) invalidate_site(e2, k);
  makeL(e2, N - 1);
  out.resize(N);
  DevBuf Rb, Pb, Eb, Zb, rb, Qb;
  // R_N[w_{N-1}][a][t][1] = sum_s W_N[w, t, s, 1] psi_N[a, s]
  long K = 1;
  {
    const long chi = in[N - 1].chil;
    const int d = e->d[N - 1];
    Rb.ensure((size_t)e->w[N - 1] * chi * d * K * e->esz);
    WStepDesc ws;
    ws.G = e->w[N]; ws.Gp = e->w[N - 1]; ws.D = d; ws.Dp = d;
    ws.nA = chi; ws.nB = K * m2;
    ws.in = in[N - 1].buf.p; ws.in_sg = chi * d * K * m2; ws.in_sA = d * K * m2; ws.in_sS = K * m2;
    ws.out = Rb.p; ws.out_sg = chi * d * K * m2; ws.out_sA = d * K * m2; ws.out_sS = K * m2;
    ws.op = &op_right(e, N);
    wstep(c, false, ws);
  }
  for (int j = N; j >= 2; --j) {
    const long chi = in[j - 1].chil;
    const int W = e->w[j - 1];
    const int d = e->d[j - 1];
    const long p = (long)d * K;
    const long WC = (long)W * chi;
    const EnvT& E = e2->LR[j - 1];
    TN_REQUIRE(E.chi == chi && E.w == W * W, TN_E_SHAPE, "apply: environment shapes disagree");
    Eb.ensure((size_t)WC * WC * e->esz);
    swap01_f64(c, E.buf.as<double>(), W, WC, chi * m2, Eb.as<double>());  // E'[(w' a')][w][a]
    Zb.ensure((size_t)WC * p * e->esz);
    GemmDesc g1;  // Z[(w' a'), q] = E'[(w' a'), (w a)] R[(w a), q]
    g1.M = WC; g1.N = p; g1.K = WC;
    g1.A = Eb.p; g1.lda = WC;
    g1.B = Rb.p; g1.ldb = p;
    g1.C = Zb.p; g1.ldc = p;
    gemm(c, P64, g1);
    rb.ensure((size_t)p * p * e->esz);
    GemmDesc g2;  // rho[q', q] = sum conj(R[(w' a'), q']) Z[(w' a'), q]
    g2.M = p; g2.N = p; g2.K = WC;
    g2.A = Rb.p; g2.lda = p; g2.tA = 1; g2.cA = cx;
    g2.B = Zb.p; g2.ldb = p;
    g2.C = rb.p; g2.ldc = p;
    gemm(c, P64, g2);
    GramEig ge;
    herm_eig(c, rb.p, p, p, cx, ge);
    std::vector<int> order;
    std::vector<double> P;
    sorted_weights(ge.w, p, order, P);
    const TruncRule tr = truncate_weights(P, maxdim, 1, cutoff);
    const long kn = tr.n;
    int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)kn * sizeof(int));
    double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)kn * sizeof(double));
    std::vector<double> ones(kn, 1.0);
    TN_HIP(hipMemcpyAsync(dperm, order.data(), kn * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipMemcpyAsync(dsc, ones.data(), kn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SiteT& O = out[j - 1];
    O.buf.ensure((size_t)kn * p * e->esz);
    gather_rows(c, (const double*)ge.Vt, ge.ppad * m2, dperm, dsc, kn, p * m2, O.buf.as<double>());
    O.chil = kn;
    O.chir = K;
    Qb.ensure((size_t)WC * kn * e->esz);
    GemmDesc g3;  // Q[(w a), n] = R[(w a), q] conj(Ut[n, q])
    g3.M = WC; g3.N = kn; g3.K = p;
    g3.A = Rb.p; g3.lda = p;
    g3.B = O.buf.p; g3.ldb = p; g3.tB = 1; g3.cB = cx;
    g3.C = Qb.p; g3.ldc = kn;
    gemm(c, P64, g3);
    const long chi0 = in[j - 2].chil;
    const int d0 = e->d[j - 2];
    Pb.ensure((size_t)W * chi0 * d0 * kn * e->esz);
    GemmDesc g4;  // P[w][(a0 s)][n] = psi_{j-1}[(a0 s), a] Q[w][a][n]
    g4.M = chi0 * d0; g4.N = kn; g4.K = chi;
    g4.A = in[j - 2].buf.p; g4.lda = chi; g4.sA = 0;
    g4.B = Qb.p; g4.ldb = kn; g4.sB = chi * kn;
    g4.C = Pb.p; g4.ldc = kn; g4.sC = chi0 * d0 * kn;
    g4.batch = W;
    gemm(c, P64, g4);
    Rb.ensure((size_t)e->w[j - 2] * chi0 * d0 * kn * e->esz);
    WStepDesc ws;  // R_{j-1}[w0][a0][t][n] = sum W_{j-1}[w0, t, s, w] P[w][a0][s][n]
    ws.G = W; ws.Gp = e->w[j - 2]; ws.D = d0; ws.Dp = d0;
    ws.nA = chi0; ws.nB = kn * m2;
    ws.in = Pb.p; ws.in_sg = chi0 * d0 * kn * m2; ws.in_sA = d0 * kn * m2; ws.in_sS = kn * m2;
    ws.out = Rb.p; ws.out_sg = chi0 * d0 * kn * m2; ws.out_sA = d0 * kn * m2; ws.out_sS = kn * m2;
    ws.op = &op_right(e, j - 1);
    wstep(c, false, ws);
    TN_HIP(hipStreamSynchronize(c->stream));  // host vectors above go out of scope
    K = kn;
  }
  SiteT& O1 = out[0];
  const long n1 = (long)e->d[0] * K;
  O1.buf.ensure((size_t)n1 * e->esz);
  copy_f64(c, O1.buf.as<double>(), Rb.as<double>(), cx ? 2 * n1 : n1);
  O1.chil = 1;
  O1.chir = K;
}

// _krylov_addbasis!(psi, phis, cutoff) (src/mps/sweep.jl:458-534) on the
// environment's psi; phis are brought to orthocenter N on e2.
void krylov_addbasis(tn_env* e, tn_env* e2, std::vector<Sites>& phis, double ext_cutoff) {
  tn_ctx* c = e->ctx;
  const bool cx = e->cplx();
  const int m2 = cx ? 2 : 1;
  const Prec P64 = gemm_prec_full(e);
  const int N = e->N;
  orthogonalize(e, N);
  for (Sites& ph : phis) {
    for (int k = 0; k < N; ++k) e2->A[k] = std::move(ph[k]);
    orthogonalize(e2, N);
    for (int k = 0; k < N; ++k) ph[k] = std::move(e2->A[k]);
  }
  DevBuf Bx, G, GP, T1, T2, Tb;
  for (int j = N; j >= 2; --j) {
    SiteT& S = e->A[j - 1];
    const long m = S.chil, n = (long)e->d[j - 1] * S.chir;
    // B: right singular vectors of psi_j (no truncation)
    GramEig ge;
    gram_eig(c, S.buf.p, m, n, false, cx, ge);
    std::vector<int> order;
    std::vector<double> P;
    const long r = std::min(m, n);
    sorted_weights(ge.w, r, order, P);
    int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)n * sizeof(int));
    double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)n * sizeof(double));
    std::vector<double> ones(n, 1.0);
    TN_HIP(hipMemcpyAsync(dperm, order.data(), r * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipMemcpyAsync(dsc, ones.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    Bx.ensure((size_t)2 * n * n * e->esz);
    gather_rows(c, (const double*)ge.Vt, ge.ppad * m2, dperm, dsc, r, n * m2, Bx.as<double>());
    TN_HIP(hipStreamSynchronize(c->stream));
    // G = sum_phi phi_j^H phi_j / tr  (the conjugate of rho = sum phi^T conj(phi))
    G.ensure((size_t)n * n * e->esz);
    double tr = 0.0;
    for (size_t i = 0; i < phis.size(); ++i) {
      SiteT& F = phis[i][j - 1];
      TN_REQUIRE(F.chir == S.chir, TN_E_SHAPE, "krylov_extend!: Krylov vector bonds disagree");
      GemmDesc g;
      g.M = n; g.N = n; g.K = F.chil;
      g.A = F.buf.p; g.lda = n; g.tA = 1; g.cA = cx;
      g.B = F.buf.p; g.ldb = n;
      g.C = G.p; g.ldc = n;
      g.beta = i == 0 ? 0.0 : 1.0;
      gemm(c, P64, g);
      Krylov K(c, cx, F.chil * n, 1);
      tr += K.norm2(F.buf.as<double>());
    }
    TN_REQUIRE(tr > 0.0, TN_E_NUMERIC, "krylov_extend!: zero Krylov vector");
    scale_f64(c, G.as<double>(), 1.0 / tr, n * n * m2);
    // X = (1 - B^H B) G (1 - B^H B)
    T1.ensure((size_t)r * n * e->esz);
    GemmDesc a1;  // T1 = B G
    a1.M = r; a1.N = n; a1.K = n;
    a1.A = Bx.p; a1.lda = n;
    a1.B = G.p; a1.ldb = n;
    a1.C = T1.p; a1.ldc = n;
    gemm(c, P64, a1);
    GP.ensure((size_t)n * n * e->esz);
    copy_f64(c, GP.as<double>(), G.as<double>(), n * n * m2);
    GemmDesc a2;  // GP = G - B^H T1
    a2.M = n; a2.N = n; a2.K = r;
    a2.A = Bx.p; a2.lda = n; a2.tA = 1; a2.cA = cx;
    a2.B = T1.p; a2.ldb = n;
    a2.C = GP.p; a2.ldc = n;
    a2.alpha = -1.0; a2.beta = 1.0;
    gemm(c, P64, a2);
    T2.ensure((size_t)n * r * e->esz);
    GemmDesc a3;  // T2 = GP B^H
    a3.M = n; a3.N = r; a3.K = n;
    a3.A = GP.p; a3.lda = n;
    a3.B = Bx.p; a3.ldb = n; a3.tB = 1; a3.cB = cx;
    a3.C = T2.p; a3.ldc = r;
    gemm(c, P64, a3);
    copy_f64(c, G.as<double>(), GP.as<double>(), n * n * m2);
    GemmDesc a4;  // X = GP - T2 B
    a4.M = n; a4.N = n; a4.K = r;
    a4.A = T2.p; a4.lda = r;
    a4.B = Bx.p; a4.ldb = n;
    a4.C = G.p; a4.ldc = n;
    a4.alpha = -1.0; a4.beta = 1.0;
    gemm(c, P64, a4);
    long kx = r;
    {
      Krylov K(c, cx, n * n, 1);
      const double xn = std::sqrt(K.norm2(G.as<double>()));
      if (xn > 1e-14) {
        GramEig gx;
        herm_eig(c, G.p, n, n, cx, gx);
        std::vector<int> ox;
        std::vector<double> Px;
        sorted_weights(gx.w, n, ox, Px);
        const TruncRule t = truncate_weights(Px, std::numeric_limits<long>::max(), 1, ext_cutoff);
        TN_HIP(hipMemcpyAsync(dperm, ox.data(), t.n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        gather_rows(c, (const double*)gx.Vt, gx.ppad * m2, dperm, dsc, t.n, n * m2,
                    Bx.as<double>() + (size_t)r * n * m2);
        TN_HIP(hipStreamSynchronize(c->stream));
        kx = r + t.n;
      }
    }
    // every state: site j-1 absorbs site_j Bx^H, site j becomes Bx
    auto move_state = [&](SiteT& Aj, SiteT& Ap) {
      Tb.ensure((size_t)Aj.chil * kx * e->esz);
      GemmDesc g;
      g.M = Aj.chil; g.N = kx; g.K = n;
      g.A = Aj.buf.p; g.lda = n;
      g.B = Bx.p; g.ldb = n; g.tB = 1; g.cB = cx;
      g.C = Tb.p; g.ldc = kx;
      gemm(c, P64, g);
      const long mm = Ap.chil * e->d[j - 2];
      DevBuf np;
      np.ensure((size_t)mm * kx * e->esz);
      GemmDesc h;
      h.M = mm; h.N = kx; h.K = Aj.chil;
      h.A = Ap.buf.p; h.lda = Aj.chil;
      h.B = Tb.p; h.ldb = kx;
      h.C = np.p; h.ldc = kx;
      gemm(c, P64, h);
      Aj.buf.ensure((size_t)kx * n * e->esz);
      copy_f64(c, Aj.buf.as<double>(), Bx.as<double>(), kx * n * m2);
      TN_HIP(hipStreamSynchronize(c->stream));
      Ap.buf = std::move(np);
      Ap.chir = kx;
      Aj.chil = kx;
    };
    move_state(e->A[j - 1], e->A[j - 2]);
    for (Sites& ph : phis) move_state(ph[j - 1], ph[j - 2]);
    invalidate_site(e, j);
    invalidate_site(e, j - 1);
  }
  orthogonalize(e, 1);
}

void krylov_extend(tn_env* e, int krylovdim, long applyH_maxdim, double applyH_cutoff, double ext_cutoff) {
  TN_REQUIRE(e->nq == 0, TN_E_UNSUPPORTED, "QN-conserving environment: subspace expansion is not built");
  TN_REQUIRE(krylovdim >= 1, TN_E_INVALID, "krylov_extend!: extension_krylovdim must be >= 1");
  TN_REQUIRE(e->proj.empty(), TN_E_UNSUPPORTED, "krylov_extend!: the StateEnvs must be created by a single MPO");
  require_mpo(e);
  require_sites(e);
  tn_ctx* c = e->ctx;
  const int N = e->N;
  // H^dag H on the host (real MPO)
  std::vector<int> w2(N + 1);
  for (int k = 0; k <= N; ++k) w2[k] = e->w[k] * e->w[k];
  EnvOwner own;
  int rc = tn_env_create(c, N, e->d.data(), w2.data(), e->dtype, &own.e);
  if (rc != 0) throw Error(rc, "krylov_extend!: cannot create the H^dag H environment");
  tn_env* e2 = own.e;
  for (int k = 1; k <= N; ++k) {
    const int wl = e->w[k - 1], wr = e->w[k], d = e->d[k - 1];
    const std::vector<double>& W = e->Wh[k - 1];
    std::vector<double> HH((size_t)wl * wl * d * d * wr * wr, 0.0);
    auto Wi = [&](int a, int t, int s, int b) { return W[(((size_t)a * d + t) * d + s) * wr + b]; };
    for (int a = 0; a < wl; ++a)
      for (int a2 = 0; a2 < wl; ++a2)
        for (int sp = 0; sp < d; ++sp)
          for (int s0 = 0; s0 < d; ++s0)
            for (int b = 0; b < wr; ++b)
              for (int b2 = 0; b2 < wr; ++b2) {
                double acc = 0.0;
                for (int t = 0; t < d; ++t) acc += Wi(a, t, s0, b) * Wi(a2, t, sp, b2);
                HH[((((size_t)(a * wl + a2) * d + sp) * d + s0) * wr + b) * wr + b2] = acc;
              }
    rc = tn_env_set_mpo(e2, k, HH.data());
    if (rc != 0) throw Error(rc, "krylov_extend!: cannot set the H^dag H MPO");
  }
  long maxlink = 1;
  for (int k = 0; k < N; ++k) maxlink = std::max(maxlink, e->A[k].chir);
  if (applyH_maxdim <= 0) applyH_maxdim = maxlink + 2;
  std::vector<Sites> phis(krylovdim);
  for (int k = 0; k < krylovdim; ++k) {
    apply_mpo(e, e2, k == 0 ? e->A : phis[k - 1], applyH_maxdim, applyH_cutoff, phis[k]);
    SiteT& F = phis[k][0];
    const long n1 = F.chil * e->d[0] * F.chir;
    Krylov K(c, e->cplx(), n1, 1);
    const double nr = std::sqrt(K.norm2(F.buf.as<double>()));
    TN_REQUIRE(nr > 0.0, TN_E_NUMERIC, "krylov_extend!: H psi vanishes");
    scale_f64(c, F.buf.as<double>(), 1.0 / nr, e->cplx() ? 2 * n1 : n1);
  }
  krylov_addbasis(e, e2, phis, ext_cutoff);
  for (int k = 1; k <= N; ++k) invalidate_site(e, k);
  e->lpos = 0;
  e->rpos = N + 1;
  e->nsite = 2;
  e->ortho = 1;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int tn_ctx_create(int device, tn_ctx** out) {
  return guarded(nullptr, [&] {
    TN_REQUIRE(out != nullptr, TN_E_INVALID, "null output pointer");
    int ndev = 0;
    TN_HIP(hipGetDeviceCount(&ndev));
    TN_REQUIRE(device >= 0 && device < ndev, TN_E_INVALID, "no such device");
    TN_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TN_HIP(hipGetDeviceProperties(&prop, device));
    TN_REQUIRE(std::string(prop.gcnArchName).rfind("gfx950", 0) == 0, TN_E_UNSUPPORTED,
               std::string("libtnhip is built for gfx950 only, device is ") + prop.gcnArchName);
    auto* c = new tn_ctx();
    c->device = device;
    TN_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->ws.resize(WS_SLOTS);
    c->ones_buf.ensure(16);
    double one[2] = {1.0, 0.0};
    TN_HIP(hipMemcpy(c->ones_buf.p, one, 16, hipMemcpyHostToDevice));
    c->ones = c->ones_buf.as<double>();
    TN_HIP(hipHostMalloc((void**)&c->pinned, tn_ctx::PINNED_N * sizeof(double), hipHostMallocMapped));
    TN_HIP(hipHostGetDevicePointer((void**)&c->pinned_dev, c->pinned, 0));
    c->scratch_dev = scratch_address(device);
    const char* tv = std::getenv("TNHIP_TIMING");
    c->timing = tv && tv[0] == '1';
    *out = c;
  });
}

int tn_ctx_destroy(tn_ctx* c) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  if (!c) return TN_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& pr : c->ev_pool) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (hipEvent_t& ev : c->kry_ev)
    if (ev) (void)hipEventDestroy(ev);
  c->ws.clear();
  c->ones_buf.release();
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->hstage) (void)hipHostFree(c->hstage);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return TN_OK;
}

// a copy per calling thread, taken under the library lock: the context's
// message may be replaced by a call on another thread right after
const char* tn_last_error(const tn_ctx* c) {
  thread_local std::string copy;
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  copy = c ? c->last_error : g_thread_error;
  return copy.c_str();
}

int tn_ctx_synchronize(tn_ctx* c) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    TN_HIP(hipStreamSynchronize(c->stream));
  });
}

int tn_alloc(tn_ctx* c, size_t bytes, void** dptr) {
  return guarded(c, [&] {
    TN_REQUIRE(c && dptr, TN_E_INVALID, "null argument");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
    if (e != hipSuccess) throw Error(TN_E_NOMEM, std::string("tn_alloc: ") + hipGetErrorString(e));
  });
}
int tn_free(tn_ctx* c, void* dptr) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    if (dptr) {
      TN_HIP(hipStreamSynchronize(c->stream));
      TN_HIP(hipFree(dptr));
    }
  });
}
int tn_memcpy_h2d(tn_ctx* c, void* dst, const void* src, size_t bytes) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    if (!bytes) return;
    TN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
  });
}
int tn_memcpy_d2h(tn_ctx* c, void* dst, const void* src, size_t bytes) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    if (!bytes) return;
    TN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
  });
}
int tn_ctx_timers(tn_ctx* c, double* ms4, int reset) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    flush_timers(c);
    if (ms4)
      for (int i = 0; i < 4; ++i) ms4[i] = c->phase_ms[i];
    if (reset)
      for (int i = 0; i < PH_COUNT; ++i) { c->phase_ms[i] = 0.0; c->phase_n[i] = 0; c->phase_bytes[i] = 0.0; }
  });
}
int tn_ctx_enable_timing(tn_ctx* c, int on) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    flush_timers(c);
    c->timing = on != 0;
    c->timing_mask = on != 0 ? ~0u : 0u;
  });
}
int tn_ctx_enable_timing_mask(tn_ctx* c, unsigned mask) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    flush_timers(c);
    c->timing = mask != 0;
    c->timing_mask = mask;
  });
}

int tn_ctx_phase_stats(tn_ctx* c, int phase, double* ms, long* count) {
  return guarded(c, [&] {
    TN_REQUIRE(c, TN_E_INVALID, "null context");
    TN_REQUIRE(phase >= 0 && phase < PH_COUNT, TN_E_INVALID, "phase_stats: unknown phase");
    flush_timers(c);
    if (ms) *ms = c->phase_ms[phase];
    if (count) *count = c->phase_n[phase];
  });
}

int tn_ctx_phase_bytes(tn_ctx* c, int phase, double* bytes) {
  return guarded(c, [&] {
    TN_REQUIRE(c && bytes, TN_E_INVALID, "null argument");
    TN_REQUIRE(phase >= 0 && phase < PH_COUNT, TN_E_INVALID, "phase_bytes: unknown phase");
    *bytes = c->phase_bytes[phase];
  });
}

void tn_update_params_default(tn_update_params* p, int solver) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->solver = solver;
  p->normalize = 1;
  p->maxdim = 9223372036854775807L;
  p->mindim = 1;
  p->cutoff = 1e-15;
  p->reverse_step = solver == 1 ? 1 : 0;
  p->krylovdim = solver == 1 ? 30 : 5;
  p->maxiter = solver == 1 ? 100 : 2;
  p->tol = solver == 1 ? 1e-12 : 1e-14;
  p->check_convergence = 0;
  p->precision = TN_PREC_FP64;
  p->refine_fp64 = 1;
  p->refine_precision = TN_PREC_FP64;
}

int tn_env_create(tn_ctx* c, int N, const int* d, const int* w, int dtype, tn_env** out) {
  return guarded(c, [&] {
    TN_REQUIRE(c && d && w && out, TN_E_INVALID, "null argument");
    TN_REQUIRE(N >= 2, TN_E_INVALID, "StateEnvs: need at least 2 sites");
    TN_REQUIRE(dtype == TN_F64 || dtype == TN_C128, TN_E_INVALID, "StateEnvs: unknown dtype");
    TN_REQUIRE(w[0] == 1 && w[N] == 1, TN_E_SHAPE, "StateEnvs: MPO edge bond dimensions must be 1");
    for (int k = 0; k < N; ++k) TN_REQUIRE(d[k] >= 1 && w[k] >= 1, TN_E_INVALID, "StateEnvs: bad dims");
    auto* e = new tn_env();
    e->ctx = c;
    e->N = N;
    e->d.assign(d, d + N);
    e->w.assign(w, w + N + 1);
    e->dtype = dtype;
    e->esz = dtype == TN_C128 ? 16 : 8;
    e->Wh.resize(N);
    e->W_set.assign(N, 0);
    e->A.resize(N);
    e->LR.resize(N + 2);
    e->LR_version.assign(N + 2, 0);
    e->lpos = 0;
    e->rpos = N + 1;
    e->opL.resize(N);
    e->opR.resize(N);
    e->op2.resize(N);
    e->opL_ok.assign(N, 0);
    e->opR_ok.assign(N, 0);
    e->op2_ok.assign(N, 0);
    set_edge(e, 0);
    set_edge(e, N + 1);
    *out = e;
  });
}

int tn_env_destroy(tn_env* e) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  if (!e) return TN_OK;
  (void)hipStreamSynchronize(e->ctx->stream);
  delete e;
  return TN_OK;
}

int tn_env_set_mpo(tn_env* e, int k, const double* W) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k >= 1 && k <= e->N && W, TN_E_INVALID, "set_mpo: bad site");
    const size_t n = (size_t)e->w[k - 1] * e->d[k - 1] * e->d[k - 1] * e->w[k];
    e->Wh[k - 1].assign(W, W + n);
    e->W_set[k - 1] = 1;
    ++e->mpo_ver;
    e->opL_ok[k - 1] = e->opR_ok[k - 1] = 0;
    if (k >= 2) e->op2_ok[k - 2] = 0;
    e->op2_ok[k - 1] = 0;
    e->lpos = std::min(e->lpos, k - 1);
    e->rpos = std::max(e->rpos, k + 1);
  });
}

int tn_env_set_site(tn_env* e, int k, const void* A, long chil, long chir) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k >= 1 && k <= e->N && A, TN_E_INVALID, "set_site: bad site");
    TN_REQUIRE(chil >= 1 && chir >= 1, TN_E_SHAPE, "set_site: bad bond dimensions");
    TN_REQUIRE((k > 1 || chil == 1) && (k < e->N || chir == 1), TN_E_SHAPE,
               "set_site: edge bond dimensions must be 1");
    SiteT& S = e->A[k - 1];
    const size_t bytes = (size_t)chil * e->d[k - 1] * chir * e->esz;
    S.buf.ensure(bytes);
    TN_HIP(hipMemcpyAsync(S.buf.p, A, bytes, hipMemcpyHostToDevice, e->ctx->stream));
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
    S.chil = chil;
    S.chir = chir;
    invalidate_site(e, k);
    e->ortho = 0;
  });
}

int tn_env_get_site_dims(tn_env* e, int k, long* chil, long* chir) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k >= 1 && k <= e->N, TN_E_INVALID, "get_site_dims: bad site");
    if (chil) *chil = e->A[k - 1].chil;
    if (chir) *chir = e->A[k - 1].chir;
  });
}

int tn_env_get_site(tn_env* e, int k, void* A) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k >= 1 && k <= e->N && A, TN_E_INVALID, "get_site: bad site");
    const SiteT& S = e->A[k - 1];
    TN_REQUIRE(S.buf.p, TN_E_INVALID, "get_site: site not set");
    const size_t bytes = (size_t)S.chil * e->d[k - 1] * S.chir * e->esz;
    TN_HIP(hipMemcpyAsync(A, S.buf.p, bytes, hipMemcpyDeviceToHost, e->ctx->stream));
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}

int tn_env_set_orthocenter(tn_env* e, int k) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k >= 0 && k <= e->N, TN_E_INVALID, "set_orthocenter: bad site");
    e->ortho = k;
  });
}
int tn_env_orthocenter(tn_env* e, int* k) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(k, TN_E_INVALID, "null output");
    *k = e->ortho;
  });
}
int tn_env_set_nsite(tn_env* e, int nsite) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(nsite >= 0 && nsite <= 2, TN_E_UNSUPPORTED, "set_nsite!: only nsite 0, 1 and 2 are built");
    e->nsite = nsite;
  });
}
int tn_env_nsite(tn_env* e, int* nsite) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(nsite, TN_E_INVALID, "null output");
    *nsite = e->nsite;
  });
}
int tn_env_position(tn_env* e, int pos) {
  return guarded(C(e), [&] {
    check_env(e);
    position(e, pos);
  });
}
int tn_env_set_projectors(tn_env* e, int nproj, double weight) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(nproj >= 0, TN_E_INVALID, "set_projectors: negative count");
    TN_REQUIRE(nproj == 0 || weight > 0.0, TN_E_INVALID,
               "`weight` parameter should be > 0.0 (value passed was `weight=" + std::to_string(weight) + "`)");
    e->proj.clear();
    e->proj.resize(nproj);
    for (ProjT& P : e->proj) {
      P.M.resize(e->N);
      P.LR.resize(e->N + 2);
      P.chiM.assign(e->N + 2, 0);
      P.lpos = 0;
      P.rpos = e->N + 1;
    }
    e->proj_weight = weight;
  });
}

int tn_env_set_noise_channels(tn_env* e, const int* skip_left, const int* skip_right) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(skip_left && skip_right, TN_E_INVALID, "set_noise_channels: null count array");
    for (int k = 1; k < e->N; ++k)
      TN_REQUIRE(skip_left[k] >= 0 && skip_right[k] >= 0 && skip_left[k] <= e->w[k] && skip_right[k] <= e->w[k],
                 TN_E_INVALID, "set_noise_channels: count out of range at bond " + std::to_string(k));
    e->noise_skip_l.assign(e->N + 1, 0);
    e->noise_skip_r.assign(e->N + 1, 0);
    for (int k = 1; k < e->N; ++k) {
      e->noise_skip_l[k] = skip_left[k];
      e->noise_skip_r[k] = skip_right[k];
    }
  });
}

int tn_env_set_projector_site(tn_env* e, int j, int k, const void* A, long chil, long chir) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(j >= 0 && j < (int)e->proj.size(), TN_E_INVALID, "set_projector_site: bad projector");
    TN_REQUIRE(k >= 1 && k <= e->N && A, TN_E_INVALID, "set_projector_site: bad site");
    TN_REQUIRE(chil >= 1 && chir >= 1 && (k > 1 || chil == 1) && (k < e->N || chir == 1), TN_E_SHAPE,
               "set_projector_site: bad bond dimensions");
    ProjT& P = e->proj[j];
    SiteT& S = P.M[k - 1];
    const size_t bytes = (size_t)chil * e->d[k - 1] * chir * e->esz;
    S.buf.ensure(bytes);
    TN_HIP(hipMemcpyAsync(S.buf.p, A, bytes, hipMemcpyHostToDevice, e->ctx->stream));
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
    S.chil = chil;
    S.chir = chir;
    if (k <= P.lpos) P.lpos = k - 1;
    if (k >= P.rpos) P.rpos = k + 1;
    P.m_ok = false;
  });
}

int tn_env_reset(tn_env* e) {
  return guarded(C(e), [&] {
    check_env(e);
    for (ProjT& P : e->proj) {
      P.lpos = 0;
      P.rpos = e->N + 1;
      P.m_ok = false;
    }
    e->lpos = 0;
    e->rpos = e->N + 1;
  });
}
int tn_env_phi_dims(tn_env* e, long* dims4, long* n) {
  return guarded(C(e), [&] {
    check_env(e);
    const Dom D = domain(e);
    if (dims4) { dims4[0] = D.chil; dims4[1] = D.d1; dims4[2] = D.d2; dims4[3] = D.chir; }
    if (n) *n = D.n;
  });
}
int tn_env_matvec(tn_env* e, const void* v, void* out, int precision) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(v && out, TN_E_INVALID, "matvec: null vector");
    matvec(e, v, out, precision);
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}
int tn_env_set_qn(tn_env* e, int nq, const int* site_qn, const int* link_qn) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(nq >= 0 && nq <= 2, TN_E_INVALID, "set_qn: nq must be 0, 1 or 2");
    if (nq == 0) {
      e->nq = 0;
      e->site_qn.clear();
      e->link_qn.clear();
      return;
    }
    TN_REQUIRE(site_qn && link_qn, TN_E_INVALID, "set_qn: null QN array");
    require_sites(e);
    TN_REQUIRE(!e->cplx(), TN_E_UNSUPPORTED, "set_qn: QN conservation is built for real tensors");
    TN_REQUIRE(e->proj.empty(), TN_E_UNSUPPORTED, "set_qn: QN conservation with penalty projectors is not built");
    std::vector<std::vector<std::array<int, 2>>> sq(e->N), lq(e->N + 1);
    long o = 0;
    for (int k = 0; k < e->N; ++k)
      for (int s2 = 0; s2 < e->d[k]; ++s2, ++o) sq[k].push_back({site_qn[o * nq], nq > 1 ? site_qn[o * nq + 1] : 0});
    o = 0;
    for (int k = 0; k <= e->N; ++k) {
      const long chi = (k == 0) ? e->A[0].chil : e->A[k - 1].chir;
      for (long a = 0; a < chi; ++a, ++o) lq[k].push_back({link_qn[o * nq], nq > 1 ? link_qn[o * nq + 1] : 0});
    }
    for (const auto& q : lq[0]) TN_REQUIRE(q[0] == 0 && q[1] == 0, TN_E_INVALID, "set_qn: link 0 must carry QN 0");
    e->nq = nq;
    e->site_qn.swap(sq);
    e->link_qn.swap(lq);
    e->link_ver.resize(e->N + 1, 0);
    for (auto& v : e->link_ver) ++v;
    e->flux_ver = -1;
    bool all = true;
    for (int k = 0; k < e->N; ++k) all = all && e->W_set[k];
    if (all) infer_mpo_flux(e);  // rejects an MPO that does not conserve the QNs
  });
}

int tn_env_get_link_qn(tn_env* e, int k, int* out, long cap, long* n) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(e->nq > 0, TN_E_INVALID, "get_link_qn: no QNs set");
    TN_REQUIRE(k >= 0 && k <= e->N, TN_E_INVALID, "get_link_qn: link out of range");
    const auto& q = e->link_qn[k];
    if (n) *n = (long)q.size();
    if (out)
      for (long a = 0; a < (long)q.size() && (a + 1) * e->nq <= cap; ++a)
        for (int j = 0; j < e->nq; ++j) out[a * e->nq + j] = q[a][j];
  });
}

int tn_env_matvec_noalloc(tn_env* e, const void* v, void* out) {
  return guarded(C(e), [&] {
    check_env(e);
    matvec_noalloc(e, v, out);
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}

int tn_env_contract_phi(tn_env* e, int pos, void* phi) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(phi, TN_E_INVALID, "contract_phi: null output");
    contract_phi(e, pos, phi);
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}
int tn_env_orthogonalize(tn_env* e, int site) {
  return guarded(C(e), [&] {
    check_env(e);
    orthogonalize(e, site);
  });
}

int tn_krylov_eig(tn_env* e, void* phi, int kdim, int maxiter, double tol, int precision, double* eig,
                  tn_krylov_info* info) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(phi, TN_E_INVALID, "eig_solver: null vector");
    const double th = krylov_eig(e, (double*)phi, kdim, maxiter, tol, precision, info);
    if (eig) *eig = th;
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}
int tn_krylov_expm(tn_env* e, void* phi, double t_re, double t_im, int kdim, int maxiter, double tol,
                   int precision, tn_krylov_info* info) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(phi, TN_E_INVALID, "exp_solver: null vector");
    krylov_expm(e, (double*)phi, cplx(t_re, t_im), kdim, maxiter, tol, precision, info);
    TN_HIP(hipStreamSynchronize(e->ctx->stream));
  });
}
int tn_rayleigh_fp64(tn_env* e, const void* phi, double* e2) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(phi && e2, TN_E_INVALID, "rayleigh: null argument");
    const cplx r = rayleigh(e, (const double*)phi);
    e2[0] = r.real();
    e2[1] = r.imag();
  });
}

int tn_replacebond(tn_env* e, int b, const void* phi, int ortho_left, long maxdim, long mindim, double cutoff,
                   double noise, int normalize, double* truncerr, double* eigs_out, long eigs_cap, long* nkept) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(phi, TN_E_INVALID, "replacebond!: null phi");
    std::vector<double> eigs;
    replacebond(e, b, (const double*)phi, ortho_left != 0, maxdim, mindim, cutoff, normalize != 0, truncerr,
                &eigs, nkept, noise);
    if (eigs_out)
      for (long i = 0; i < std::min<long>(eigs_cap, (long)eigs.size()); ++i) eigs_out[i] = eigs[i];
  });
}

int tn_svd_trunc(tn_ctx* c, const double* M, long m, long n, long maxdim, long mindim, double cutoff, double* U,
                 double* S, double* Vt, long* nkept, double* truncerr) {
  return guarded(c, [&] {
    TN_REQUIRE(c && M, TN_E_INVALID, "svd: null argument");
    PhaseTimer pt(c, PH_SVD);
    // eigen-decomposition of the smaller Gram matrix; that side's singular
    // vectors are its (orthonormal) eigenvectors, the other side's are
    // M v / s (or u^T M / s)
    const bool left = m <= n;
    GramEig ge;
    gram_eig(c, M, m, n, left, false, ge);
    std::vector<int> order;
    std::vector<double> P;
    sorted_weights(ge.w, std::min(m, n), order, P);
    const TruncRule tr = truncate_weights(P, maxdim, mindim, cutoff);
    const long k = tr.n;
    std::vector<double> one(k, 1.0);
    int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)k * sizeof(int));
    double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)k * sizeof(double));
    TN_HIP(hipMemcpyAsync(dperm, order.data(), k * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipMemcpyAsync(dsc, one.data(), k * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const long L = left ? n : m;  // length of the non-orthonormal singular vectors
    double* W = (double*)c->workspace(WS_SVD_GRAM, (size_t)k * std::max(m, n) * sizeof(double));
    double* F = (double*)c->workspace(WS_MISC, (size_t)k * L * sizeof(double));
    // W = top-k eigenvectors as rows; F (k x L) = W M (left) or W M^T (right)
    // holds s_i times the other side's singular vectors
    gather_rows(c, (const double*)ge.Vt, ge.ppad, dperm, dsc, k, left ? m : n, W);
    if (left && U) gather_rows_T(c, (const double*)ge.Vt, ge.ppad, dperm, dsc, k, m, U);
    if (!left && Vt) copy_f64(c, Vt, W, k * n);
    {
      GemmDesc g;
      g.M = k; g.N = L;
      if (left) {
        g.K = m; g.A = W; g.lda = m; g.B = M; g.ldb = n;
      } else {
        g.K = n; g.A = W; g.lda = n; g.B = M; g.ldb = n; g.tB = 1;
      }
      g.C = F; g.ldc = L;
      gemm(c, Prec::F64, g);
    }
    // s_i = |F_i|: U S V^T = M V V^T (or U U^T M) holds exactly, whatever the
    // rounding of weights near eps * sum(P)
    double* dn2 = (double*)c->workspace(WS_SVD_S, (size_t)k * sizeof(double));
    row_norms2(c, F, k, L, L, dn2);
    std::vector<double> n2(k), inv(k);
    TN_HIP(hipMemcpyAsync(n2.data(), dn2, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TN_HIP(hipStreamSynchronize(c->stream));
    for (long i = 0; i < k; ++i) {
      const double sv = std::sqrt(n2[i]);
      inv[i] = sv > 0 ? 1.0 / sv : 0.0;
      if (S) S[i] = sv;
    }
    double* dinv = (double*)c->workspace(WS_SVD_SCALE, (size_t)k * sizeof(double));
    TN_HIP(hipMemcpyAsync(dinv, inv.data(), k * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<int> ident(k);
    std::iota(ident.begin(), ident.end(), 0);
    TN_HIP(hipMemcpyAsync(dperm, ident.data(), k * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (left && Vt) gather_rows(c, F, n, dperm, dinv, k, n, Vt);     // Vt[i][j] = F[i][j] / s_i
    if (!left && U) gather_rows_T(c, F, m, dperm, dinv, k, m, U);   // U[j][i] = F[i][j] / s_i
    TN_HIP(hipStreamSynchronize(c->stream));
    if (nkept) *nkept = k;
    if (truncerr) *truncerr = tr.truncerr;
  });
}


int tn_update_two_site(tn_env* e, int pos, int ortho_left, const tn_update_params* p, tn_update_result* r) {
  return guarded(C(e), [&] {
    check_env(e);
    update_two_site(e, pos, ortho_left != 0, p, r);
  });
}

int tn_env_krylov_extend(tn_env* e, int krylovdim, long applyH_maxdim, double applyH_cutoff, double ext_cutoff) {
  return guarded(C(e), [&] {
    check_env(e);
    krylov_extend(e, krylovdim, applyH_maxdim, applyH_cutoff, ext_cutoff);
  });
}

int tn_update_one_site(tn_env* e, int pos, int ortho_left, const tn_update_params* p, tn_update_result* r) {
  return guarded(C(e), [&] {
    check_env(e);
    update_one_site(e, pos, ortho_left != 0, p, r);
  });
}

int tn_fullsweep(tn_env* e, const tn_update_params* p, double* energy, double* maxtruncerr, long* maxchi,
                 double* mid_eigs, long cap, long* nmid) {
  return guarded(C(e), [&] {
    check_env(e);
    TN_REQUIRE(p, TN_E_INVALID, "null params");
    if (e->ortho != 1) orthogonalize(e, 1);
    const int N = e->N;
    double en = NAN, mte = 0.0;
    std::vector<double> mid;
    tn_update_result r{};
    // the sweep reports the energy of its last update; a Rayleigh-quotient
    // refinement (refine_fp64 = 1) changes only that number, so it runs there
    // (a Lanczos refinement in FP32 moves the state at every update; the last
    // update refines in FP64, so the reported energy is an FP64 Ritz value)
    tn_update_params pm = *p, pl = *p;
    if (p->refine_fp64 == 1 && p->precision != TN_PREC_FP64) pm.refine_fp64 = 0;
    if (p->refine_fp64 >= 2) pl.refine_precision = TN_PREC_FP64;
    for (int b = 1; b <= N - 1; ++b) {
      update_two_site(e, b, true, &pm, &r);
      en = r.energy;
      mte = std::max(mte, r.truncerr);
    }
    for (int b = N - 1; b >= 1; --b) {
      std::vector<double> eigs;
      update_two_site(e, b, false, b == 1 ? &pl : &pm, &r, &eigs);
      en = r.energy;
      mte = std::max(mte, r.truncerr);
      if (b == N / 2) mid.swap(eigs);  // lasteigs[N/2] (sweep.jl:141,158)
    }
    long mc = 0;
    for (int k = 0; k < N; ++k) mc = std::max(mc, e->A[k].chir);
    if (energy) *energy = en;
    if (maxtruncerr) *maxtruncerr = mte;
    if (maxchi) *maxchi = mc;
    if (nmid) *nmid = (long)mid.size();
    if (mid_eigs)
      for (long i = 0; i < std::min<long>(cap, (long)mid.size()); ++i) mid_eigs[i] = mid[i];
  });
}

}  // extern "C"
