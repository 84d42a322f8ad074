// C ABI for tree tensor networks (StateEnvsTTN / LinkTensorsTTN,
// src/ttn/state_envs_ttn.jl, src/ttn/linktensors.jl) on one MI355X.
//
// Every node tensor has three legs (slots 0, 1, 2; row-major A[a0][a1][a2]).
// The Hamiltonian enters as a tree operator: each link carries channels c
// and each node a sparse coupling W[c0][c1][c2] between the channels of its
// three legs, so that the projected operator at a node is
//   (H v)[a0',a1',a2'] = sum_W w E0[c0] E1[c1] E2[c2] v
// with E_s the environment of slot s (the far side of that leg).  A site leg
// is a slot whose environment holds the channel operators of that site.
//
// Environment layouts (FP64, per slot):
//   slots 0, 1   E[c][bra][ket]     (the first and the batched middle GEMM)
//   slot 2       E[c][ket][bra]     (the closing GEMM, K = (c, ket))
// Node product at node n (D_s the leg dimensions, n_s the channel counts):
//   A[(c0 a0'), (a1 a2)]  = E0[(c0 a0'), a0] v[a0, (a1 a2)]           GEMM
//   S[a0'][a1'][c2][a2]  += w E1[c1] A[c0][a0'][a1][a2]  per W entry    GEMM batched over a0'
//   out[(a0' a1'), a2']   = S[(a0' a1'), (c2 a2)] E2[(c2 a2), a2']       GEMM
// The environment of slot s toward the neighbour (move_linktensors_to_next!)
// closes the same chains with the bra copy of v.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include <cstring>

#include "tn_internal.h"

using namespace tn;

struct TTNSlot {
  int nbr = -1;       // neighbouring node, -1 for a physical site
  int nbr_slot = -1;  // slot of this link at the neighbour
  int nchan = 0;
  DevBuf env;         // environment of the far side, layout by slot
  long env_dim = 0;   // leg dimension the environment was built for
  bool env_ok = false;
};

// the node's coupling terms sorted by one output channel for the grouped
// GEMMs: q = 0 by c2 (A index c1, B index c0), q = 1 by c0 (c1, c2),
// q = 2 by c1 (c0, c2); device arrays off (ng + 1), ia, ib, w
struct TermGroups {
  DevBuf buf;
  int ng = 0;
  const int* off = nullptr;
  const int* ia = nullptr;
  const int* ib = nullptr;
  const double* w = nullptr;
};

struct TTNNode {
  long dims[3] = {0, 0, 0};
  DevBuf A;
  DevBuf spare;  // grow-only second buffer: new tensors are written here, then swapped with A
  TTNSlot slot[3];
  std::vector<int> c[3];
  std::vector<double> w;
  bool has_tensor = false;
  TermGroups tg[3];
  bool tg_ok = false;
  long size() const { return dims[0] * dims[1] * dims[2]; }
};

// penalty projector TTN M (LinkProjTTN, src/ttn/linkproj.jl): its node
// tensors and, per node and slot, the overlap environment of the far side
// E[x][y] (x: psi's leg, y: M's leg); physical slots are the identity
struct TTNProjNode {
  long dims[3] = {0, 0, 0};
  DevBuf M;
  DevBuf env[3];
  long env_x[3] = {0, 0, 0}, env_y[3] = {0, 0, 0};
  bool env_ok[3] = {false, false, false};
  DevBuf m;            // local projection at this node (psi's node shape)
  long m_epoch = -1;
};

struct tn_ttn {
  tn_ctx* ctx = nullptr;
  std::vector<TTNNode> nodes;
  int center = -1;                       // orthogonality center (-1: none)
  uint64_t rng = 0x853c49e6748fea9bULL;  // SplitMix64 state of the expansion padding
  std::vector<std::vector<TTNProjNode>> proj;  // [projector][node]
  double proj_weight = 0.0;
  long epoch = 0;  // bumped whenever an overlap environment changes
};

namespace {
thread_local std::string g_err;

template <class F>
int guarded_t(tn_ctx* c, F&& f) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  try {
    f();
    return TN_OK;
  } catch (const Error& e) {
    g_err = e.what();
    if (c) c->last_error = g_err;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    if (c) c->last_error = g_err;
    return TN_E_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    if (c) c->last_error = g_err;
    return TN_E_INVALID;
  }
}

tn_ctx* CT(tn_ttn* t) { return t ? t->ctx : nullptr; }

TTNNode& node_of(tn_ttn* t, int n) {
  TN_REQUIRE(t && n >= 0 && n < (int)t->nodes.size(), TN_E_INVALID, "TTN: node index out of range");
  return t->nodes[n];
}

void require_env(const TTNNode& N, int s) {
  const TTNSlot& S = N.slot[s];
  TN_REQUIRE(S.env_ok && S.env_dim == N.dims[s], TN_E_POSITION,
             "TTN: environment of slot " + std::to_string(s) + " is not built for the current bond");
}

void gemm64(tn_ctx* c, GemmDesc g) { gemm(c, Prec::F64, g); }

const double* proj_local(tn_ttn* t, int n, size_t j);

void build_groups(tn_ttn* t, TTNNode& N) {
  if (N.tg_ok) return;
  const int nt = (int)N.w.size();
  const int gkey[3] = {2, 0, 1}, akey[3] = {1, 1, 0}, bkey[3] = {0, 2, 2};
  for (int q = 0; q < 3; ++q) {
    const int ng = N.slot[gkey[q]].nchan;
    std::vector<int> off(ng + 1, 0), ia(nt), ib(nt);
    std::vector<double> w(nt);
    for (int k = 0; k < nt; ++k) {
      for (int x = 0; x < 3; ++x)
        TN_REQUIRE(N.c[x][k] >= 0 && N.c[x][k] < N.slot[x].nchan, TN_E_INVALID, "TTN: coupling channel out of range");
      ++off[N.c[gkey[q]][k] + 1];
    }
    for (int g = 0; g < ng; ++g) off[g + 1] += off[g];
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int k = 0; k < nt; ++k) {
      const int at = pos[N.c[gkey[q]][k]]++;
      ia[at] = N.c[akey[q]][k];
      ib[at] = N.c[bkey[q]][k];
      w[at] = N.w[k];
    }
    const size_t ni = (size_t)(ng + 1 + 2 * nt);
    const size_t wofs = ((ni * sizeof(int) + 7) / 8) * 8;
    std::vector<char> host(wofs + (size_t)nt * sizeof(double) + 8);
    std::memcpy(host.data(), off.data(), (ng + 1) * sizeof(int));
    std::memcpy(host.data() + (ng + 1) * sizeof(int), ia.data(), nt * sizeof(int));
    std::memcpy(host.data() + (ng + 1 + nt) * sizeof(int), ib.data(), nt * sizeof(int));
    std::memcpy(host.data() + wofs, w.data(), nt * sizeof(double));
    TermGroups& G = N.tg[q];
    G.buf.ensure(host.size());
    TN_HIP(hipMemcpyAsync(G.buf.p, host.data(), host.size(), hipMemcpyHostToDevice, t->ctx->stream));
    TN_HIP(hipStreamSynchronize(t->ctx->stream));
    const char* base = (const char*)G.buf.p;
    G.ng = ng;
    G.off = (const int*)base;
    G.ia = G.off + ng + 1;
    G.ib = G.ia + nt;
    G.w = (const double*)(base + wofs);
  }
  N.tg_ok = true;
}

// S[a0'][a1'][c2][a2] = sum_W w E0[c0] E1[c1] v   (legs 0 and 1 applied, leg 2 open)
void apply_legs01(tn_ttn* t, const TTNNode& N, const double* v, double* S) {
  tn_ctx* c = t->ctx;
  const long D0 = N.dims[0], D1 = N.dims[1], D2 = N.dims[2];
  const long n0 = N.slot[0].nchan, n2 = N.slot[2].nchan;
  double* A = (double*)c->workspace(WS_TTN_A, (size_t)(n0 * D0 * D1 * D2) * sizeof(double));
  {
    GemmDesc g;
    g.M = n0 * D0; g.N = D1 * D2; g.K = D0;
    g.A = N.slot[0].env.p; g.lda = D0;
    g.B = v; g.ldb = D1 * D2;
    g.C = A; g.ldc = D1 * D2;
    gemm64(c, g);
  }
  // S[a0'][a1'][c2][a2] = sum_{terms of c2} w E1[c1] A[c0][a0']   (one grouped launch)
  build_groups(t, const_cast<TTNNode&>(N));
  const TermGroups& G = N.tg[0];
  GroupedDesc d;
  d.M = D1; d.N = D2; d.K = D1;
  d.A = (const double*)N.slot[1].env.p; d.lda = D1; d.sA = 0; d.tsA = D1 * D1;
  d.B = A; d.ldb = D2; d.sB = D1 * D2; d.tsB = D0 * D1 * D2;
  d.C = S; d.ldc = n2 * D2; d.sC = D1 * n2 * D2; d.gsC = D2;
  d.off = G.off; d.ia = G.ia; d.ib = G.ib; d.w = G.w;
  d.ngroups = G.ng; d.batch = (int)D0;
  TN_REQUIRE(G.ng == n2, TN_E_SHAPE, "TTN: channel count of slot 2 changed");
  gemm_grouped_f64(c, d);
}

void matvec_node(tn_ttn* t, int n, const double* v, double* out) {
  PhaseTimer pt(t->ctx, PH_MATVEC);
  const TTNNode& N = t->nodes[n];
  for (int s = 0; s < 3; ++s) require_env(N, s);
  const long D0 = N.dims[0], D1 = N.dims[1], D2 = N.dims[2], n2 = N.slot[2].nchan;
  double* S = (double*)t->ctx->workspace(WS_TTN_S, (size_t)(D0 * D1 * n2 * D2) * sizeof(double));
  apply_legs01(t, N, v, S);
  GemmDesc g;
  g.M = D0 * D1; g.N = D2; g.K = n2 * D2;
  g.A = S; g.lda = n2 * D2;
  g.B = N.slot[2].env.p; g.ldb = D2;
  g.C = out; g.ldc = D2;
  gemm64(t->ctx, g);
  // product(EnvCouplingModelProjTTN) (environment.jl:104-110): + weight m <m, v>
  if (!t->proj.empty()) {
    double* co = (double*)t->ctx->workspace(WS_PROJ_C, 64 * sizeof(double));
    for (size_t j = 0; j < t->proj.size(); ++j) {
      const double* m = proj_local(t, n, j);
      multi_dot(t->ctx, m, N.size(), 1, v, N.size(), false, co);
      add_scaled_dev(t->ctx, out, m, co, t->proj_weight, N.size(), false);
    }
  }
}

// G[c2][(a0 a1)][a2'] = v[(a0 a1)][a2] E2[c2][a2][a2']
double* close_leg2(tn_ttn* t, const TTNNode& N, const double* v) {
  tn_ctx* c = t->ctx;
  const long D0 = N.dims[0], D1 = N.dims[1], D2 = N.dims[2], n2 = N.slot[2].nchan;
  double* G = (double*)c->workspace(WS_TTN_G, (size_t)(n2 * D0 * D1 * D2) * sizeof(double));
  GemmDesc g;
  g.M = D0 * D1; g.N = D2; g.K = D2;
  g.A = v; g.lda = D2; g.sA = 0;
  g.B = N.slot[2].env.p; g.ldb = D2; g.sB = D2 * D2;
  g.C = G; g.ldc = D2; g.sC = D0 * D1 * D2;
  g.batch = (int)n2;
  gemm64(c, g);
  return G;
}

// T with leg `leg` mapped through E (ex x ey): to_y contracts T's leg (dim
// ex) with E's first index, else T's leg (dim ey) with E's second index.
void apply_leg(tn_ctx* c, const double* T, const long dims[3], int leg, const double* E, long ex, long ey, bool to_y,
               double* out) {
  const long kin = to_y ? ex : ey, kout = to_y ? ey : ex;
  TN_REQUIRE(dims[leg] == kin, TN_E_SHAPE, "TTN projector: leg and overlap environment disagree");
  GemmDesc g;
  if (leg == 0) {  // out[o][x1 x2] = sum_k E'[k][o] T[k][x1 x2]
    g.M = kout; g.N = dims[1] * dims[2]; g.K = kin;
    g.A = E; g.lda = ey; g.tA = to_y ? 1 : 0;
    g.B = T; g.ldb = dims[1] * dims[2];
    g.C = out; g.ldc = dims[1] * dims[2];
  } else if (leg == 1) {  // batched over x0
    g.M = kout; g.N = dims[2]; g.K = kin;
    g.A = E; g.lda = ey; g.tA = to_y ? 1 : 0; g.sA = 0;
    g.B = T; g.ldb = dims[2]; g.sB = dims[1] * dims[2];
    g.C = out; g.ldc = dims[2]; g.sC = kout * dims[2];
    g.batch = (int)dims[0];
  } else {  // out[x0 x1][o] = T[x0 x1][k] E'[k][o]
    g.M = dims[0] * dims[1]; g.N = kout; g.K = kin;
    g.A = T; g.lda = dims[2];
    g.B = E; g.ldb = ey; g.tB = to_y ? 0 : 1;
    g.C = out; g.ldc = kout;
  }
  gemm64(c, g);
}

// move_linkproj_to_next! (src/ttn/linkproj.jl:48-66) for every projector:
// E[x_s][y_s] = sum psi_n[x] M_n[y] prod_{o != s} E_o[x_o][y_o], written into
// the neighbour across slot s.
void proj_update_env(tn_ttn* t, int n, int s) {
  tn_ctx* c = t->ctx;
  const TTNNode& N = t->nodes[n];
  const int nb = N.slot[s].nbr, ts = N.slot[s].nbr_slot;
  for (auto& P : t->proj) {
    TTNProjNode& Q = P[n];
    TN_REQUIRE(Q.M.p != nullptr, TN_E_INVALID, "TTN projector: node tensor not set");
    long dims[3] = {N.dims[0], N.dims[1], N.dims[2]};
    const double* T = N.A.as<double>();
    int cur = 0;
    for (int o = 0; o < 3; ++o) {
      if (o == s) continue;
      if (N.slot[o].nbr < 0) {
        TN_REQUIRE(Q.dims[o] == N.dims[o], TN_E_SHAPE, "TTN projector: physical dimensions disagree");
        continue;
      }
      TN_REQUIRE(Q.env_ok[o] && Q.env_x[o] == N.dims[o] && Q.env_y[o] == Q.dims[o], TN_E_POSITION,
                 "TTN projector: overlap environment of slot " + std::to_string(o) + " is not built");
      long nd[3] = {dims[0], dims[1], dims[2]};
      nd[o] = Q.dims[o];
      const size_t need = (size_t)(nd[0] * nd[1] * nd[2]) * sizeof(double);
      double* out = (double*)c->workspace(cur == 0 ? WS_TTNP_0 : WS_TTNP_1, need);
      apply_leg(c, T, dims, o, Q.env[o].as<double>(), N.dims[o], Q.dims[o], true, out);
      T = out;
      dims[o] = nd[o];
      cur ^= 1;
    }
    TTNProjNode& R = P[nb];
    const long X = N.dims[s], Y = Q.dims[s];
    R.env[ts].ensure((size_t)(X * Y) * sizeof(double));
    double* E = R.env[ts].as<double>();
    const double* Mt = Q.M.as<double>();
    GemmDesc g;
    if (s == 0) {
      g.M = X; g.N = Y; g.K = Q.dims[1] * Q.dims[2];
      g.A = T; g.lda = g.K;
      g.B = Mt; g.ldb = g.K; g.tB = 1;
    } else if (s == 2) {
      g.M = X; g.N = Y; g.K = Q.dims[0] * Q.dims[1];
      g.A = T; g.lda = X; g.tA = 1;
      g.B = Mt; g.ldb = Y;
    } else {
      double* Ts = (double*)c->workspace(WS_TTNP_2, (size_t)(Q.dims[0] * X * Q.dims[2]) * sizeof(double));
      double* Ms = (double*)c->workspace(WS_TTNP_3, (size_t)(Q.dims[0] * Y * Q.dims[2]) * sizeof(double));
      swap01_f64(c, T, Q.dims[0], X, Q.dims[2], Ts);
      swap01_f64(c, Mt, Q.dims[0], Y, Q.dims[2], Ms);
      g.M = X; g.N = Y; g.K = Q.dims[0] * Q.dims[2];
      g.A = Ts; g.lda = g.K;
      g.B = Ms; g.ldb = g.K; g.tB = 1;
    }
    g.C = E; g.ldc = Y;
    gemm64(c, g);
    R.env_x[ts] = X;
    R.env_y[ts] = Y;
    R.env_ok[ts] = true;
  }
  ++t->epoch;
}

// local projection m[x] = sum_y M[y] prod_s E_s[x_s][y_s] at node n (contract /
// product of LinkProjTTN, linkproj.jl:126-165), cached until an overlap
// environment changes
const double* proj_local(tn_ttn* t, int n, size_t j) {
  tn_ctx* c = t->ctx;
  const TTNNode& N = t->nodes[n];
  TTNProjNode& Q = t->proj[j][n];
  if (Q.m_epoch == t->epoch) return Q.m.as<double>();
  long dims[3] = {Q.dims[0], Q.dims[1], Q.dims[2]};
  const double* T = Q.M.as<double>();
  int cur = 0;
  for (int s = 0; s < 3; ++s) {
    if (N.slot[s].nbr < 0) {
      TN_REQUIRE(Q.dims[s] == N.dims[s], TN_E_SHAPE, "TTN projector: physical dimensions disagree");
      continue;
    }
    TN_REQUIRE(Q.env_ok[s] && Q.env_x[s] == N.dims[s] && Q.env_y[s] == Q.dims[s], TN_E_POSITION,
               "TTN projector: overlap environment of slot " + std::to_string(s) + " is not built");
    long nd[3] = {dims[0], dims[1], dims[2]};
    nd[s] = N.dims[s];
    double* out = (double*)c->workspace(cur == 0 ? WS_TTNP_0 : WS_TTNP_1, (size_t)(nd[0] * nd[1] * nd[2]) * sizeof(double));
    apply_leg(c, T, dims, s, Q.env[s].as<double>(), N.dims[s], Q.dims[s], false, out);
    T = out;
    dims[s] = nd[s];
    cur ^= 1;
  }
  Q.m.ensure((size_t)N.size() * sizeof(double));
  copy_f64(c, Q.m.as<double>(), T, N.size());
  Q.m_epoch = t->epoch;
  return Q.m.as<double>();
}

// move_linktensors_to_next! (src/ttn/linktensors.jl:63-118): the environment
// of the link at slot s of node n, seen from n, written into the neighbour.
void update_env(tn_ttn* t, int n, int s) {
  PhaseTimer pt(t->ctx, PH_ENV);
  tn_ctx* c = t->ctx;
  TTNNode& N = t->nodes[n];
  TN_REQUIRE(s >= 0 && s < 3 && N.slot[s].nbr >= 0, TN_E_INVALID, "TTN: slot has no neighbouring node");
  TN_REQUIRE(N.has_tensor, TN_E_INVALID, "TTN: node tensor not set");
  for (int o = 0; o < 3; ++o)
    if (o != s) require_env(N, o);
  TTNNode& M = t->nodes[N.slot[s].nbr];
  const int ts = N.slot[s].nbr_slot;
  TTNSlot& out = M.slot[ts];
  const long D0 = N.dims[0], D1 = N.dims[1], D2 = N.dims[2];
  const long Ds = N.dims[s], ns = N.slot[s].nchan;
  TN_REQUIRE(out.nchan == ns, TN_E_SHAPE, "TTN: channel counts of a link disagree");
  out.env.ensure((size_t)(ns * Ds * Ds) * sizeof(double));
  double* F = out.env.as<double>();
  const double* v = N.A.as<double>();
  if (s == 2) {
    const long n2 = ns;
    double* S = (double*)c->workspace(WS_TTN_S, (size_t)(D0 * D1 * n2 * D2) * sizeof(double));
    apply_legs01(t, N, v, S);
    if (ts == 2) {  // F[c][ket][bra] = S^T v
      GemmDesc g;
      g.M = n2 * D2; g.N = D2; g.K = D0 * D1;
      g.A = S; g.lda = n2 * D2; g.tA = 1;
      g.B = v; g.ldb = D2;
      g.C = F; g.ldc = D2;
      gemm64(c, g);
    } else {  // F[c][bra][ket] = v^T S[c]
      GemmDesc g;
      g.M = D2; g.N = D2; g.K = D0 * D1;
      g.A = v; g.lda = D2; g.tA = 1; g.sA = 0;
      g.B = S; g.ldb = n2 * D2; g.sB = D2;
      g.C = F; g.ldc = D2; g.sC = D2 * D2;
      g.batch = (int)n2;
      gemm64(c, g);
    }
  } else {
    TN_REQUIRE(ts == 2, TN_E_INVALID, "TTN: a child link must enter the child at its parent slot");
    double* G = close_leg2(t, N, v);
    const long n2 = N.slot[2].nchan;
    (void)n2;
    double* K = (double*)c->workspace(WS_TTN_A, (size_t)(ns * D0 * D1 * D2) * sizeof(double));
    build_groups(t, N);
    if (s == 0) {
      // K[c0][a0][a1'][a2'] = sum_{terms of c0} w E1[c1][a1'][a1] G[c2][a0][a1][a2']   (batched over a0)
      const TermGroups& TG = N.tg[1];
      TN_REQUIRE(TG.ng == ns, TN_E_SHAPE, "TTN: channel count of slot 0 changed");
      GroupedDesc d;
      d.M = D1; d.N = D2; d.K = D1;
      d.A = (const double*)N.slot[1].env.p; d.lda = D1; d.sA = 0; d.tsA = D1 * D1;
      d.B = G; d.ldb = D2; d.sB = D1 * D2; d.tsB = D0 * D1 * D2;
      d.C = K; d.ldc = D2; d.sC = D1 * D2; d.gsC = D0 * D1 * D2;
      d.off = TG.off; d.ia = TG.ia; d.ib = TG.ib; d.w = TG.w;
      d.ngroups = TG.ng; d.batch = (int)D0;
      gemm_grouped_f64(c, d);
      // F[c0][a0][a0'] = K[(c0 a0)][(a1' a2')] v[a0'][(a1' a2')]^T
      GemmDesc g;
      g.M = ns * D0; g.N = D0; g.K = D1 * D2;
      g.A = K; g.lda = D1 * D2;
      g.B = v; g.ldb = D1 * D2; g.tB = 1;
      g.C = F; g.ldc = D0;
      gemm64(c, g);
    } else {
      // K[c1][a1][a0'][a2'] = sum_{terms of c1} w E0[c0][a0'][a0] G[c2][a0][a1][a2']   (batched over a1)
      const TermGroups& TG = N.tg[2];
      TN_REQUIRE(TG.ng == ns, TN_E_SHAPE, "TTN: channel count of slot 1 changed");
      GroupedDesc d;
      d.M = D0; d.N = D2; d.K = D0;
      d.A = (const double*)N.slot[0].env.p; d.lda = D0; d.sA = 0; d.tsA = D0 * D0;
      d.B = G; d.ldb = D1 * D2; d.sB = D2; d.tsB = D0 * D1 * D2;
      d.C = K; d.ldc = D2; d.sC = D0 * D2; d.gsC = D1 * D0 * D2;
      d.off = TG.off; d.ia = TG.ia; d.ib = TG.ib; d.w = TG.w;
      d.ngroups = TG.ng; d.batch = (int)D1;
      gemm_grouped_f64(c, d);
      double* vp = (double*)c->workspace(WS_TTN_P, (size_t)(D0 * D1 * D2) * sizeof(double));
      swap01_f64(c, v, D0, D1, D2, vp);  // vp[a1'][a0'][a2']
      GemmDesc g;
      g.M = ns * D1; g.N = D1; g.K = D0 * D2;
      g.A = K; g.lda = D0 * D2;
      g.B = vp; g.ldb = D0 * D2; g.tB = 1;
      g.C = F; g.ldc = D1;
      gemm64(c, g);
    }
  }
  out.env_dim = Ds;
  out.env_ok = true;
  if (!t->proj.empty()) proj_update_env(t, n, s);
}

// moveisometry_to_next! (src/ttn/ttn.jl:266-310): node n becomes an isometry
// toward its other legs (QR; truncating SVD when cutoff > 0 or the bond
// exceeds maxdim), the remainder is absorbed into the neighbour at slot s.
void move_isometry(tn_ttn* t, int n, int s, long maxdim, long mindim, double cutoff, double* truncerr, long* newdim) {
  PhaseTimer pt(t->ctx, PH_SVD);
  tn_ctx* c = t->ctx;
  TTNNode& N = t->nodes[n];
  TN_REQUIRE(s >= 0 && s < 3 && N.slot[s].nbr >= 0, TN_E_INVALID, "moveisometry_to_next!: slot has no neighbour");
  TTNNode& M = t->nodes[N.slot[s].nbr];
  const int ts = N.slot[s].nbr_slot;
  const long D0 = N.dims[0], D1 = N.dims[1], D2 = N.dims[2];
  const long Ds = N.dims[s];
  const long m = N.size() / Ds;
  TN_REQUIRE(M.dims[ts] == Ds, TN_E_SHAPE, "moveisometry_to_next!: bond dimensions of the link disagree");
  double* X = (double*)c->workspace(WS_QR_X, (size_t)(Ds * m) * sizeof(double));
  const double* A = N.A.as<double>();
  if (s == 0) copy_f64(c, X, A, Ds * m);
  else if (s == 1) swap01_f64(c, A, D0, D1, D2, X);
  else transpose_f64(c, A, D0 * D1, D2, X);
  const long k0 = std::min(m, Ds);
  double* Qt = (double*)c->workspace(WS_QR_Q, (size_t)(k0 * m) * sizeof(double));
  double* R = (double*)c->workspace(WS_QR_R, (size_t)(k0 * Ds) * sizeof(double));
  householder_qr(c, X, m, Ds, Qt, R);
  long k = k0;
  double terr = 0.0;
  const double* Qn = Qt;
  const double* Rn = R;
  if (cutoff > 0.0 || Ds > maxdim) {
    // SVD of R through its Gram matrix R R^T: the left factor rotates Q
    GramEig ge;
    gram_eig(c, R, k0, Ds, true, false, ge);
    std::vector<int> order(k0);
    for (long i = 0; i < k0; ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ge.w[x] > ge.w[y]; });
    std::vector<double> P(k0);
    for (long i = 0; i < k0; ++i) P[i] = std::max(0.0, ge.w[order[i]]);
    const TruncRule tr = truncate_weights(P, maxdim, mindim, cutoff);
    k = tr.n;
    terr = tr.truncerr;
    int* dperm = (int*)c->workspace(WS_SVD_PERM, (size_t)k * sizeof(int));
    double* dsc = (double*)c->workspace(WS_SVD_SCALE, (size_t)k * sizeof(double));
    std::vector<double> ones(k, 1.0);
    TN_HIP(hipMemcpyAsync(dperm, order.data(), k * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TN_HIP(hipMemcpyAsync(dsc, ones.data(), k * sizeof(double), hipMemcpyHostToDevice, c->stream));
    double* Ut = (double*)c->workspace(WS_SVD_GRAM, (size_t)(k * k0) * sizeof(double));
    gather_rows(c, (const double*)ge.Vt, ge.ppad, dperm, dsc, k, k0, Ut);  // U_k^T (k x k0)
    double* Q2 = (double*)c->workspace(WS_TTN_F, (size_t)(k * m + k * Ds) * sizeof(double));
    double* R2 = Q2 + k * m;
    GemmDesc g;
    g.M = k; g.N = m; g.K = k0;
    g.A = Ut; g.lda = k0;
    g.B = Qt; g.ldb = m;
    g.C = Q2; g.ldc = m;
    gemm64(c, g);
    GemmDesc h;
    h.M = k; h.N = Ds; h.K = k0;
    h.A = Ut; h.lda = k0;
    h.B = R; h.ldb = Ds;
    h.C = R2; h.ldc = Ds;
    gemm64(c, h);
    TN_HIP(hipStreamSynchronize(c->stream));  // host vectors above go out of scope
    Qn = Q2;
    Rn = R2;
  }
  // node n <- Q (leg s of dimension k)
  long nd[3] = {D0, D1, D2};
  nd[s] = k;
  DevBuf& An = N.spare;
  An.ensure((size_t)(nd[0] * nd[1] * nd[2]) * sizeof(double));
  if (s == 0) copy_f64(c, An.as<double>(), Qn, k * m);
  else if (s == 1) swap01_f64(c, Qn, k, D0, D2, An.as<double>());
  else transpose_f64(c, Qn, k, D0 * D1, An.as<double>());
  // neighbour <- R x_ts neighbour
  const long B0 = M.dims[0], B1 = M.dims[1], B2 = M.dims[2];
  long md[3] = {B0, B1, B2};
  md[ts] = k;
  DevBuf& Bn = M.spare;
  Bn.ensure((size_t)(md[0] * md[1] * md[2]) * sizeof(double));
  const double* B = M.A.as<double>();
  GemmDesc g;
  if (ts == 0) {
    g.M = k; g.N = B1 * B2; g.K = Ds;
    g.A = Rn; g.lda = Ds;
    g.B = B; g.ldb = B1 * B2;
    g.C = Bn.p; g.ldc = B1 * B2;
  } else if (ts == 2) {
    g.M = B0 * B1; g.N = k; g.K = Ds;
    g.A = B; g.lda = Ds;
    g.B = Rn; g.ldb = Ds; g.tB = 1;
    g.C = Bn.p; g.ldc = k;
  } else {
    g.M = k; g.N = B2; g.K = Ds;
    g.A = Rn; g.lda = Ds; g.sA = 0;
    g.B = B; g.ldb = B2; g.sB = Ds * B2;
    g.C = Bn.p; g.ldc = B2; g.sC = k * B2;
    g.batch = (int)B0;
  }
  gemm64(c, g);
  std::swap(N.A, N.spare);
  std::swap(M.A, M.spare);
  for (int i = 0; i < 3; ++i) {
    N.dims[i] = nd[i];
    M.dims[i] = md[i];
  }
  N.slot[s].env_ok = false;
  M.slot[ts].env_ok = false;
  if (truncerr) *truncerr = terr;
  if (newdim) *newdim = k;
}

double node_norm2(tn_ttn* t, const TTNNode& N) {
  tn_ctx* c = t->ctx;
  double* co = (double*)c->workspace(WS_PROJ_C, 64 * sizeof(double));
  multi_dot(c, N.A.as<double>(), N.size(), 1, N.A.as<double>(), N.size(), false, co);
  TN_HIP(hipMemcpyAsync(c->pinned, co, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TN_HIP(hipStreamSynchronize(c->stream));
  return c->pinned[0];
}

// ---- sweep driver (src/ttn/state_envs_ttn.jl:120-140, update_site_ttn.jl,
// sweep_ttn.jl:117-274) --------------------------------------------------

// shortest path a -> b in the tree (unique), as node indices
std::vector<int> tree_path(tn_ttn* t, int a, int b) {
  const int n = (int)t->nodes.size();
  std::vector<int> par(n, -2);
  std::vector<int> q{a};
  par[a] = -1;
  for (size_t h = 0; h < q.size() && par[b] == -2; ++h) {
    const int x = q[h];
    for (int s = 0; s < 3; ++s) {
      const int y = t->nodes[x].slot[s].nbr;
      if (y >= 0 && par[y] == -2) {
        par[y] = x;
        q.push_back(y);
      }
    }
  }
  TN_REQUIRE(par[b] != -2, TN_E_INVALID, "`shortest_path()`: destination is not reachable");
  std::vector<int> out;
  for (int x = b; x != -1; x = par[x]) out.push_back(x);
  std::reverse(out.begin(), out.end());
  return out;
}

int slot_to(const TTNNode& N, int m) {
  for (int s = 0; s < 3; ++s)
    if (N.slot[s].nbr == m) return s;
  throw Error(TN_E_INVALID, "TTN: nodes are not neighbours");
}

// position!(sysenv, node; ...): isometrize! then move_environment! along the path
void position(tn_ttn* t, int node, long maxdim, long mindim, double cutoff, bool normalize, int skip) {
  TN_REQUIRE(t->center >= 0, TN_E_POSITION, "`isometrize!()`: TTN does not have a proper orthogonality center !!");
  const std::vector<int> path = tree_path(t, t->center, node);
  const double cut = cutoff > 0.0 ? cutoff : -1.0;
  for (size_t k = 0; k + 1 < path.size(); ++k)
    move_isometry(t, path[k], slot_to(t->nodes[path[k]], path[k + 1]), maxdim, mindim, cut, nullptr, nullptr);
  t->center = node;
  if (normalize) {
    TTNNode& N = t->nodes[node];
    const double nr = std::sqrt(node_norm2(t, N));
    if (nr > 0.0) scale_f64(t->ctx, N.A.as<double>(), 1.0 / nr, N.size());
  }
  for (size_t k = 0; k + 1 < path.size(); ++k) {
    if (path[k] == skip) continue;
    update_env(t, path[k], slot_to(t->nodes[path[k]], path[k + 1]));
  }
}

double eig_at(tn_ttn* t, int node, int kdim, int maxiter, double tol, bool check) {
  TTNNode& N = t->nodes[node];
  for (int s = 0; s < 3; ++s) require_env(N, s);
  tn_krylov_info info{};
  const double e = krylov_eig_op(
      t->ctx, false, N.size(), [&](const double* x, double* y) { matvec_node(t, node, x, y); }, N.A.as<double>(), kdim,
      maxiter, tol, &info);
  if (check && info.converged < 1) throw Error(TN_E_NOTCONV, "`eig_solver()` not converged !!");
  return e;
}

// standard normals: SplitMix64 -> two 53-bit uniforms -> Box-Muller
double next_u01(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ULL;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
void normals(uint64_t& st, double* v, size_t n) {
  for (size_t i = 0; i < n; i += 2) {
    const double u1 = next_u01(st), u2 = next_u01(st);
    const double r = std::sqrt(-2.0 * std::log(u1));
    v[i] = r * std::cos(2.0 * M_PI * u2);
    if (i + 1 < n) v[i + 1] = r * std::sin(2.0 * M_PI * u2);
  }
}

// subspace_expand!(psi, node, nextnode, max_expand_dim, noise) (update_site_ttn.jl:65-109)
void expand_random(tn_ttn* t, int node, int slot, long max_expand_dim, double noise) {
  TTNNode& A = t->nodes[node];
  TN_REQUIRE(slot >= 0 && slot < 3 && A.slot[slot].nbr >= 0, TN_E_INVALID, "subspace_expand!: bad slot");
  TTNNode& B = t->nodes[A.slot[slot].nbr];
  const int sb = A.slot[slot].nbr_slot;
  long sa_sum = 0, sb_sum = 0;
  for (int k = 0; k < 3; ++k) {
    if (k != slot) sa_sum += A.dims[k];
    if (k != sb) sb_sum += B.dims[k];
  }
  const long p = std::min(std::min(sa_sum, sb_sum), max_expand_dim);
  if (p <= 0) return;
  tn_ctx* c = t->ctx;
  auto grow = [&](TTNNode& X, int s) {
    long pd[3] = {X.dims[0], X.dims[1], X.dims[2]};
    pd[s] = p;
    const size_t np = (size_t)(pd[0] * pd[1] * pd[2]);
    const double xn = std::sqrt(node_norm2(t, X));  // synchronises the stream
    // the padding is drawn into the context's pinned staging buffer (free:
    // the stream is idle) and uploaded from there, not from pageable memory
    double* P = (double*)c->host_stage(np * sizeof(double));
    normals(t->rng, P, np);
    double pn = 0.0;
    for (size_t i = 0; i < np; ++i) pn += P[i] * P[i];
    const double sc = pn > 0.0 ? noise * xn / std::sqrt(pn) : 0.0;
    for (size_t i = 0; i < np; ++i) P[i] *= sc;
    const size_t pb = np * sizeof(double);
    double* Pd = (double*)c->workspace(WS_TTN_P, pb);
    TN_HIP(hipMemcpyAsync(Pd, P, pb, hipMemcpyHostToDevice, c->stream));
    long nd[3] = {X.dims[0], X.dims[1], X.dims[2]};
    nd[s] += p;
    X.spare.ensure((size_t)(nd[0] * nd[1] * nd[2]) * sizeof(double));
    concat3_f64(c, X.A.as<double>(), X.dims, s, Pd, p, X.spare.as<double>());
    TN_HIP(hipStreamSynchronize(c->stream));  // the staging buffer is reused
    std::swap(X.A, X.spare);
    for (int i = 0; i < 3; ++i) X.dims[i] = nd[i];
    X.slot[s].env_ok = false;
  };
  grow(B, sb);  // nextnode's padding first, as the reference
  grow(A, slot);
}

long sat_add(long a, long b) { return a > std::numeric_limits<long>::max() - b ? std::numeric_limits<long>::max() : a + b; }

long max_link_dim(tn_ttn* t) {
  long md = 1;
  for (const TTNNode& N : t->nodes)
    for (int s = 0; s < 3; ++s)
      if (N.slot[s].nbr >= 0) md = std::max(md, N.dims[s]);
  return md;
}

}  // namespace

extern "C" {

int tn_ttn_create(tn_ctx* ctx, int nnodes, tn_ttn** out) {
  return guarded_t(ctx, [&] {
    TN_REQUIRE(ctx && out && nnodes >= 1, TN_E_INVALID, "tn_ttn_create: bad arguments");
    tn_ttn* t = new tn_ttn;
    t->ctx = ctx;
    t->nodes.resize(nnodes);
    *out = t;
  });
}

int tn_ttn_destroy(tn_ttn* t) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  if (!t) return TN_OK;
  if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
  delete t;
  return TN_OK;
}

int tn_ttn_set_link(tn_ttn* t, int a, int sa, int b, int sb, int nchan) {
  return guarded_t(CT(t), [&] {
    TTNNode& A = node_of(t, a);
    TTNNode& B = node_of(t, b);
    TN_REQUIRE(sa >= 0 && sa < 3 && sb >= 0 && sb < 3 && a != b && nchan >= 1, TN_E_INVALID,
               "tn_ttn_set_link: bad slots");
    A.slot[sa].nbr = b;
    A.slot[sa].nbr_slot = sb;
    A.slot[sa].nchan = nchan;
    A.tg_ok = false;
    A.slot[sa].env_ok = false;
    B.slot[sb].nbr = a;
    B.slot[sb].nbr_slot = sa;
    B.slot[sb].nchan = nchan;
    B.tg_ok = false;
    B.slot[sb].env_ok = false;
  });
}

int tn_ttn_set_site(tn_ttn* t, int node, int slot, int d, int nchan, const double* ops) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(slot >= 0 && slot < 2 && d >= 1 && nchan >= 1 && ops, TN_E_INVALID,
               "tn_ttn_set_site: sites occupy slot 0 or 1");
    TTNSlot& S = N.slot[slot];
    S.nbr = -1;
    S.nbr_slot = -1;
    S.nchan = nchan;
    N.tg_ok = false;
    const size_t bytes = (size_t)nchan * d * d * sizeof(double);
    S.env.ensure(bytes);
    TN_HIP(hipMemcpyAsync(S.env.p, ops, bytes, hipMemcpyHostToDevice, t->ctx->stream));
    TN_HIP(hipStreamSynchronize(t->ctx->stream));
    S.env_dim = d;
    S.env_ok = true;
  });
}

int tn_ttn_set_coupling(tn_ttn* t, int node, int nnz, const int* c0, const int* c1, const int* c2, const double* w) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(nnz >= 1 && c0 && c1 && c2 && w, TN_E_INVALID, "tn_ttn_set_coupling: empty coupling");
    for (int s = 0; s < 3; ++s) N.c[s].assign(s == 0 ? c0 : s == 1 ? c1 : c2, (s == 0 ? c0 : s == 1 ? c1 : c2) + nnz);
    N.w.assign(w, w + nnz);
    N.tg_ok = false;
    for (int k = 0; k < nnz; ++k)
      for (int s = 0; s < 3; ++s)
        TN_REQUIRE(N.c[s][k] >= 0 && N.c[s][k] < N.slot[s].nchan, TN_E_SHAPE,
                   "tn_ttn_set_coupling: channel index out of range (set links and sites first)");
  });
}

int tn_ttn_set_node(tn_ttn* t, int node, const long* dims, const double* A) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(dims && A && dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, TN_E_INVALID, "tn_ttn_set_node: bad dims");
    for (int s = 0; s < 3; ++s) {
      if (N.dims[s] != dims[s]) N.slot[s].env_ok = N.slot[s].env_ok && N.slot[s].env_dim == dims[s];
      N.dims[s] = dims[s];
    }
    const size_t bytes = (size_t)N.size() * sizeof(double);
    N.A.ensure(bytes);
    TN_HIP(hipMemcpyAsync(N.A.p, A, bytes, hipMemcpyHostToDevice, t->ctx->stream));
    TN_HIP(hipStreamSynchronize(t->ctx->stream));
    N.has_tensor = true;
  });
}

int tn_ttn_get_node(tn_ttn* t, int node, long* dims, double* A) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    if (dims)
      for (int s = 0; s < 3; ++s) dims[s] = N.dims[s];
    if (A) {
      TN_REQUIRE(N.has_tensor, TN_E_INVALID, "tn_ttn_get_node: node tensor not set");
      TN_HIP(hipMemcpyAsync(A, N.A.p, (size_t)N.size() * sizeof(double), hipMemcpyDeviceToHost, t->ctx->stream));
      TN_HIP(hipStreamSynchronize(t->ctx->stream));
    }
  });
}

int tn_ttn_node_ptr(tn_ttn* t, int node, double** dev) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(dev && N.has_tensor, TN_E_INVALID, "tn_ttn_node_ptr: node tensor not set");
    *dev = N.A.as<double>();
  });
}

int tn_ttn_set_projectors(tn_ttn* t, int nproj, double weight) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t != nullptr && nproj >= 0, TN_E_INVALID, "set_projectors: bad arguments");
    TN_REQUIRE(nproj == 0 || weight > 0.0, TN_E_INVALID,
               "`weight` parameter should be > 0.0 (value passed was `weight=" + std::to_string(weight) + "`)");
    t->proj.clear();
    t->proj.resize(nproj);
    for (auto& P : t->proj) P.resize(t->nodes.size());
    t->proj_weight = weight;
    ++t->epoch;
  });
}

int tn_ttn_set_projector_node(tn_ttn* t, int j, int node, const long* dims, const double* A) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t != nullptr && j >= 0 && j < (int)t->proj.size(), TN_E_INVALID, "set_projector_node: bad projector");
    node_of(t, node);
    TN_REQUIRE(dims && A && dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, TN_E_INVALID,
               "set_projector_node: bad dimensions");
    TTNProjNode& Q = t->proj[j][node];
    for (int s = 0; s < 3; ++s) {
      Q.dims[s] = dims[s];
      Q.env_ok[s] = false;
    }
    const size_t n = (size_t)(dims[0] * dims[1] * dims[2]);
    Q.M.ensure(n * sizeof(double));
    TN_HIP(hipMemcpyAsync(Q.M.p, A, n * sizeof(double), hipMemcpyHostToDevice, t->ctx->stream));
    TN_HIP(hipStreamSynchronize(t->ctx->stream));
    ++t->epoch;
  });
}

int tn_ttn_update_env(tn_ttn* t, int node, int slot) {
  return guarded_t(CT(t), [&] {
    node_of(t, node);
    update_env(t, node, slot);
  });
}

int tn_ttn_matvec(tn_ttn* t, int node, const double* v, double* out) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(v && out, TN_E_INVALID, "tn_ttn_matvec: null vector");
    (void)N;
    matvec_node(t, node, v, out);
  });
}

int tn_ttn_eig(tn_ttn* t, int node, int krylovdim, int maxiter, double tol, double* energy, tn_krylov_info* info) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(N.has_tensor, TN_E_INVALID, "tn_ttn_eig: node tensor not set");
    for (int s = 0; s < 3; ++s) require_env(N, s);
    const double e = krylov_eig_op(
        t->ctx, false, N.size(), [&](const double* x, double* y) { matvec_node(t, node, x, y); }, N.A.as<double>(),
        krylovdim, maxiter, tol, info);
    if (energy) *energy = e;
  });
}

int tn_ttn_move(tn_ttn* t, int node, int slot, long maxdim, long mindim, double cutoff, double* truncerr,
                long* newdim) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(N.has_tensor, TN_E_INVALID, "moveisometry_to_next!: node tensor not set");
    move_isometry(t, node, slot, maxdim, mindim, cutoff, truncerr, newdim);
  });
}

int tn_ttn_normalize(tn_ttn* t, int node, double* norm) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(N.has_tensor, TN_E_INVALID, "normalize!: node tensor not set");
    const double nr = std::sqrt(node_norm2(t, N));
    if (nr > 0.0) scale_f64(t->ctx, N.A.as<double>(), 1.0 / nr, N.size());
    if (norm) *norm = nr;
  });
}

int tn_ttn_node_norm(tn_ttn* t, int node, double* norm) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(N.has_tensor && norm, TN_E_INVALID, "tn_ttn_node_norm: node tensor not set");
    *norm = std::sqrt(node_norm2(t, N));
  });
}

int tn_ttn_expand(tn_ttn* t, int node, int slot, long pad, const double* padA, const double* padB) {
  return guarded_t(CT(t), [&] {
    TTNNode& N = node_of(t, node);
    TN_REQUIRE(slot >= 0 && slot < 3 && N.slot[slot].nbr >= 0 && pad >= 0 && padA && padB, TN_E_INVALID,
               "subspace_expand!: bad arguments");
    TTNNode& M = t->nodes[N.slot[slot].nbr];
    const int ts = N.slot[slot].nbr_slot;
    if (pad == 0) return;
    tn_ctx* c = t->ctx;
    auto grow = [&](TTNNode& X, int s, const double* P) {
      long pd[3] = {X.dims[0], X.dims[1], X.dims[2]};
      pd[s] = pad;
      const size_t pb = (size_t)(pd[0] * pd[1] * pd[2]) * sizeof(double);
      double* Pd = (double*)c->workspace(WS_TTN_P, pb);
      TN_HIP(hipMemcpyAsync(Pd, P, pb, hipMemcpyHostToDevice, c->stream));
      long nd[3] = {X.dims[0], X.dims[1], X.dims[2]};
      nd[s] += pad;
      X.spare.ensure((size_t)(nd[0] * nd[1] * nd[2]) * sizeof(double));
      concat3_f64(c, X.A.as<double>(), X.dims, s, Pd, pad, X.spare.as<double>());
      TN_HIP(hipStreamSynchronize(c->stream));
      std::swap(X.A, X.spare);
      for (int i = 0; i < 3; ++i) X.dims[i] = nd[i];
      X.slot[s].env_ok = false;
    };
    grow(N, slot, padA);
    grow(M, ts, padB);
  });
}

int tn_ttn_set_center(tn_ttn* t, int node) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t && node >= -1 && node < (int)t->nodes.size(), TN_E_INVALID, "tn_ttn_set_center: bad node");
    t->center = node;
  });
}

int tn_ttn_get_center(tn_ttn* t, int* node) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t && node, TN_E_INVALID, "tn_ttn_get_center: null argument");
    *node = t->center;
  });
}

int tn_ttn_seed(tn_ttn* t, unsigned long long seed) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t, TN_E_INVALID, "tn_ttn_seed: null handle");
    t->rng = (uint64_t)seed;
  });
}

int tn_ttn_position(tn_ttn* t, int node, long maxdim, long mindim, double cutoff, int normalize, int skip) {
  return guarded_t(CT(t), [&] {
    node_of(t, node);
    position(t, node, maxdim, mindim, cutoff, normalize != 0, skip);
  });
}

int tn_ttn_update_position(tn_ttn* t, int node, long maxdim, long mindim, double cutoff, int normalize,
                           int krylovdim, int maxiter, double tol, int check_convergence, double* energy) {
  return guarded_t(CT(t), [&] {
    node_of(t, node);
    position(t, node, maxdim, mindim, cutoff, normalize != 0, -1);
    const double e = eig_at(t, node, krylovdim, maxiter, tol, check_convergence != 0);
    if (energy) *energy = e;
  });
}

int tn_ttn_expand_random(tn_ttn* t, int node, int slot, long max_expand_dim, double noise) {
  return guarded_t(CT(t), [&] {
    node_of(t, node);
    expand_random(t, node, slot, max_expand_dim, noise);
  });
}

int tn_ttn_fullsweep(tn_ttn* t, const int* path, int npath, int central, long maxdim, long mindim, double cutoff,
                     double noise, int expand_dim, int max_expand_dim, int expand_numiter, int krylovdim, int maxiter,
                     double tol, int normalize, int check_convergence, double* energy, long* maxchi) {
  return guarded_t(CT(t), [&] {
    TN_REQUIRE(t && path && npath >= 1, TN_E_INVALID, "fullsweep!: empty sweep path");
    for (int k = 0; k < npath; ++k) node_of(t, path[k]);
    TN_REQUIRE(!(expand_dim != 0 && expand_numiter < 2), TN_E_INVALID,
               "`fullsweep!()`: `expand_numiter=" + std::to_string(expand_numiter) +
                   "` cannot be less than 2 for `expand_dim=" + std::to_string(expand_dim) + "` !!");
    std::vector<int> order(path, path + npath);
    for (int k = npath - 1; k >= 0; --k) order.push_back(path[k]);
    const bool norm = normalize != 0, chk = check_convergence != 0;
    double e = NAN;
    if (std::fabs(noise) < 100 * 1e-15) {
      for (int node : order) {
        position(t, node, maxdim, mindim, -1.0, norm, -1);
        e = eig_at(t, node, krylovdim, maxiter, tol, chk);
      }
    } else {
      node_of(t, central);
      for (int node : order) {
        if (node == central) {
          position(t, node, maxdim, mindim, -1.0, norm, -1);
          e = eig_at(t, node, krylovdim, maxiter, tol, chk);
          continue;
        }
        const int nxt = tree_path(t, node, central)[1];
        position(t, node, maxdim, mindim, -1.0, norm, nxt);
        expand_random(t, node, slot_to(t->nodes[node], nxt), max_expand_dim, noise);
        t->center = nxt;
        for (int dummy = 1; dummy <= expand_numiter; ++dummy) {
          const long md = dummy == expand_numiter ? maxdim : sat_add(maxdim, dummy == 1 ? max_expand_dim : expand_dim);
          const int nn = dummy % 2 == 1 ? node : nxt;
          position(t, nn, md, mindim, cutoff, norm, -1);
          e = eig_at(t, nn, krylovdim, maxiter, tol, chk);
        }
      }
    }
    if (energy) *energy = e;
    if (maxchi) *maxchi = max_link_dim(t);
  });
}

}  // extern "C"
