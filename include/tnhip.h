/*
 * tnhip.h -- C ABI of libtnhip.so, the MI355X (gfx950) effective-Hamiltonian
 * hot path of TenNetLib.jl (two-site DMRG / TDVP on MPS).
 *
 * Every entry point returns an int status: TN_OK (0) or a negative TN_E_*
 * code; tn_last_error() returns the message of the last failure on a context
 * (or of the last failure of the calling thread when ctx == NULL).
 *
 * Ownership: the library owns every device buffer it allocates for
 * environments, MPS tensors and workspaces.  Host pointers are borrowed for
 * the duration of a call.  Device pointers handed in (phi, v, out) are
 * caller-owned buffers from tn_alloc()/tn_free().
 *
 * Threading: one tn_ctx = one HIP stream; a context and the environments
 * created on it are not thread-safe.  Separate contexts may run concurrently.
 *
 * Index conventions (row-major, contiguous):
 *   MPS site      A_k[a, s, b]          (chi_{k-1}, d_k, chi_k)
 *   MPO tensor    W_k[wl, s', s, wr]    (w_{k-1}, d_k, d_k, w_k), s' = bra
 *   two-site phi  phi[a, s, t, b]       (chi_{p-1}, d_p, d_{p+1}, chi_{p+1})
 *   one-site phi  phi[a, s, b]
 * Sites and bonds are 1-based as in the reference.
 *
 * Reference interfaces replaced (TenNetLib.jl 0.2.2 unless marked upstream):
 *   StateEnvs(psi, H::MPO)         src/mps/state_envs.jl:54-60
 *   position!(sysenv, pos)         src/mps/state_envs.jl:364-367
 *   set_nsite!/nsite               src/mps/state_envs.jl:341,352-355
 *   product(sysenv, v)/sysenv(v)   src/mps/state_envs.jl:376-378
 *   ProjMPO makeL!/makeR!/product  upstream ITensorMPS; mirror src/mps/projmps2.jl:77-173
 *   eig_solver / exp_solver        src/base/solver.jl:23-43, 66-88
 *   replacebond!                   upstream ITensorMPS, called src/mps/update_site.jl:64-76
 *   _update_two_site!              src/mps/update_site.jl:27-90
 *   fullsweep!                     src/mps/sweep.jl:94-196
 */
#ifndef TNHIP_H
#define TNHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define TN_OK 0
#define TN_E_INVALID (-1)      /* bad argument / parameter validation       */
#define TN_E_SHAPE (-2)        /* index or order mismatch (projmps2.jl:208-222) */
#define TN_E_POSITION (-3)     /* wrong position / orthocenter (update_site.jl:39-41) */
#define TN_E_NOMEM (-4)        /* device allocation failed                  */
#define TN_E_HIP (-5)          /* HIP runtime error                         */
#define TN_E_NOTCONV (-6)      /* solver not converged when asked (solver.jl:38-41) */
#define TN_E_UNSUPPORTED (-7)  /* feature not built                         */
#define TN_E_NUMERIC (-8)      /* NaN/Inf detected                           */

/* ---- element types and matvec precision ----------------------------- */
#define TN_F64 0   /* real float64 storage                                  */
#define TN_C128 1  /* complex float64 storage (interleaved re, im)          */

#define TN_PREC_FP64 0    /* FP64 MFMA GEMM chain                          */
#define TN_PREC_FP32 1    /* FP32 MFMA GEMM chain, FP64 in/out             */
#define TN_PREC_BF16X3 2  /* bf16 hi/lo split, 3 MFMA products, FP32 accum   */

typedef struct tn_ctx tn_ctx;
typedef struct tn_env tn_env;

typedef struct tn_krylov_info {
  int converged;    /* number of converged Ritz pairs (eig) / 1 if done (exp) */
  int numops;       /* operator applications                                  */
  int numiter;      /* restarts + 1                                           */
  double residual;  /* residual norm of the returned Ritz pair / error est.  */
} tn_krylov_info;

/* Keyword arguments of update_position!/fullsweep! (update_site.jl:235-245,
 * solver.jl:27-35, 72-79).  Fill with tn_update_params_default(). */
typedef struct tn_update_params {
  int solver;          /* 0 = eig_solver, 1 = exp_solver                     */
  double time_step_re; /* exp_solver time step (already halved by tdvpsweep!) */
  double time_step_im;
  int normalize;       /* normalize = true                                    */
  long maxdim;         /* typemax(Int)                                        */
  long mindim;         /* 1                                                   */
  double cutoff;       /* Float64_threshold() = 1e-15                         */
  int reverse_step;    /* TDVP one-site backward step                         */
  int krylovdim;       /* 5 (eig) / 30 (exp)                                  */
  int maxiter;         /* 2 (eig) / 100 (exp)                                 */
  double tol;          /* 1e-14 (eig) / 1e-12 (exp)                           */
  int check_convergence;
  int precision;       /* TN_PREC_*: precision of the Krylov matvecs          */
  int refine_fp64;     /* mixed precision only.  0: none; 1: the reported energy is the
                          Rayleigh quotient of one FP64 product of the converged vector;
                          k >= 2: one FP64 Lanczos run of Krylov dimension k started from
                          the converged vector (moves the state; the BF16x3 default is 2) */
  double noise;        /* 0: density-matrix perturbation noise * noiseterm(PH, phi, ortho)
                          added before the factorisation (update_site.jl:60-62)         */
  int refine_precision; /* products of a refine_fp64 >= 2 Lanczos run: TN_PREC_FP64
                           (default) or TN_PREC_FP32 (enough to take a BF16x3 state to
                           the FP32 form's accuracy at a fraction of an FP64 product)  */
} tn_update_params;

typedef struct tn_update_result {
  double energy;       /* Ritz value, or Rayleigh quotient (update_site.jl:51-57) */
  double energy_im;
  double truncerr;     /* spec.truncerr                                        */
  long nkept;          /* new bond dimension                                   */
  tn_krylov_info info;
  double* eigs;        /* in: caller buffer (or NULL) for spec.eigs = kept P   */
  long eigs_cap;       /* in: capacity of eigs                                 */
} tn_update_result;

/* ---- context and device memory --------------------------------------- */
int tn_ctx_create(int device, tn_ctx** out);
int tn_ctx_destroy(tn_ctx* ctx);
const char* tn_last_error(const tn_ctx* ctx);
int tn_ctx_synchronize(tn_ctx* ctx);
int tn_alloc(tn_ctx* ctx, size_t bytes, void** dptr);
int tn_free(tn_ctx* ctx, void* dptr);
int tn_memcpy_h2d(tn_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int tn_memcpy_d2h(tn_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
/* per-phase device time accumulated on the context's stream (ms):
 * [0] matvec, [1] env update, [2] Krylov vector ops, [3] SVD/replacebond */
int tn_ctx_timers(tn_ctx* ctx, double* ms4, int reset);
/* turn per-phase event timing on (1) or off (0); off by default */
int tn_ctx_enable_timing(tn_ctx* ctx, int on);
/* time only the phases whose bit is set in mask (bit p: phase p of
 * tn_ctx_phase_stats); 0 turns timing off.  Events are recorded on the
 * context's stream without synchronising, so timing one phase (e.g. bit 4,
 * the matvec GEMM launches) leaves a timed region's wall clock as it is */
int tn_ctx_enable_timing_mask(tn_ctx* ctx, unsigned mask);
/* accumulated device ms and event count of one phase: 0 matvec, 1 env update,
 * 2 Krylov vector ops, 3 SVD/replacebond, 4 each GEMM launch of a matvec */
int tn_ctx_phase_stats(tn_ctx* ctx, int phase, double* ms, long* count);
/* phase 5 times each launch of the tridiagonalisation's product kernel
 * (large density matrices); tn_ctx_phase_bytes returns the algorithmic bytes
 * of the timed launches of a phase (8 m^2 per launch for phase 5) */
int tn_ctx_phase_bytes(tn_ctx* ctx, int phase, double* bytes);
void tn_update_params_default(tn_update_params* p, int solver);

/* ---- StateEnvs{ProjMPO}: MPS + cached environments -------------------- */
/* N sites, physical dims d[0..N-1], MPO bond dims w[0..N] (w[0] = w[N] = 1),
 * element type TN_F64 / TN_C128.  StateEnvs(psi, H) state_envs.jl:54-60. */
int tn_env_create(tn_ctx* ctx, int N, const int* d, const int* w, int dtype, tn_env** out);
int tn_env_destroy(tn_env* env);
/* W_host: real (w_{k-1}, d, d, w_k); k is 1-based. */
int tn_env_set_mpo(tn_env* env, int k, const double* W_host);
/* A_host: (chil, d_k, chir) of the env's dtype; invalidates cached envs that
 * contain site k (ProjMPO lpos/rpos semantics). */
int tn_env_set_site(tn_env* env, int k, const void* A_host, long chil, long chir);
int tn_env_get_site_dims(tn_env* env, int k, long* chil, long* chir);
int tn_env_get_site(tn_env* env, int k, void* A_host);
/* orthocenter bookkeeping of the MPS (0 = unknown / not orthogonal) */
int tn_env_set_orthocenter(tn_env* env, int k);
int tn_env_orthocenter(tn_env* env, int* k);
/* set_nsite!/nsite (state_envs.jl:341,352) */
int tn_env_set_nsite(tn_env* env, int nsite);
int tn_env_nsite(tn_env* env, int* nsite);
/* position!(sysenv, pos) -> makeL!(pos-1), makeR!(pos+nsite) */
int tn_env_position(tn_env* env, int pos);
/* drop cached environments (lpos = 0, rpos = N+1) */
int tn_env_reset(tn_env* env);
/* Penalty projectors for excited states: StateEnvs(psi, H, Ms; weight)
 * (state_envs.jl:83-106) builds ProjMPO_MPS2 (projmpo_mps2.jl:37-101), whose
 * product is H v + weight * sum_k m_k (m_k^H v) with m_k the projection of
 * the MPS M_k at the current position (ProjMPS2, projmps2.jl:34-224).
 * set_projectors allocates nproj projectors (weight > 0, else TN_E_INVALID with
 * the reference's message); set_projector_site uploads site k (1-based) of
 * projector j (0-based), shape (chil, d_k, chir) of the env's dtype. */
int tn_env_set_projectors(tn_env* env, int nproj, double weight);
int tn_env_set_projector_site(tn_env* env, int j, int k, const void* A_host, long chil, long chir);
/* Noise term of a projected sum of term strings: StateEnvs(psi, H::CouplingModel)
 * (state_envs.jl:136-175) builds ProjCouplingModel, whose noiseterm
 * (projcouplingmodel.jl:391-492) sums X_id X_id^H over the term ids already
 * acting on the kept side of the bond.  When the MPO set with tn_env_set_mpo
 * is that sum with the ids' channels laid out per bond, skip_left[k] leading
 * channels of bond k (identity on sites <= k) are left out of the left noise
 * factor at bond k and skip_right[k] trailing channels (identity on sites > k)
 * out of the right one; arrays of length N+1, entries 1..N-1 read, each in
 * [0, w[k]].  Default 0 (ITensorMPS noiseterm of a ProjMPO: all channels). */
int tn_env_set_noise_channels(tn_env* env, const int* skip_left, const int* skip_right);
/* Quantum-number (U(1) or U(1)xU(1)) conservation: ConserveQNs of the
 * reference's QN-conserving runs (test/test_MPS_DMRG.jl:10,29; NDTensors
 * BlockSparse factorisation, upstream).  nq charges per QN (1 or 2; 0 turns
 * it off); site_qn: for every site k the QNs of its d_k basis states (nq ints
 * each, sites concatenated); link_qn: for every link 0..N (link k carries the
 * QN of sites 1..k) the QNs of its elements, links concatenated.  The MPS must
 * be block-sparse with respect to them and right-orthonormal with orthocenter
 * 1 (a product state is).  The MPO must conserve the QNs (every nonzero
 * W[wl, s', s, wr] with flux(wr) = flux(wl) + q(s') - q(s), total flux 0;
 * else TN_E_INVALID).  replacebond! then factorises block by block with
 * NDTensors' block truncation and updates the link QNs (kept elements grouped
 * by QN); the two-site FP64 and FP32 products run as block-sparse GEMMs over
 * the link sectors.  One-site updates, noise, projectors and complex tensors
 * are not built in this mode (TN_E_UNSUPPORTED). */
int tn_env_set_qn(tn_env* env, int nq, const int* site_qn, const int* link_qn);
/* QNs of link k (0..N): *n elements, nq ints each into out (capacity cap ints) */
int tn_env_get_link_qn(tn_env* env, int k, int* out, long cap, long* n);
/* shape of the projected-operator domain at the current position:
 * dims[0..3] = (chil, d_p, d_{p+1} or 1, chir); *n = element count */
int tn_env_phi_dims(tn_env* env, long* dims4, long* n);
/* product(sysenv, v) at the current position and nsite; v, out on device */
int tn_env_matvec(tn_env* env, const void* v_dev, void* out_dev, int precision);
/* Allocation-free FP64 product (no counterpart in the reference): the
 * product of tn_env_matvec with TN_PREC_FP64 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * It calls no hipMalloc/hipFree; its intermediates live in one fixed 32 MiB
 * device array per GPU, so (w_l + w_r) * chil * d_p * d_{p+1} * chir elements
 * plus split-K partial sums must fit (else TN_E_INVALID).  Needs position! at
 * the current position (nsite 1 or 2) and no penalty projectors (else
 * TN_E_UNSUPPORTED).  Calls that overlap on one GPU race on that array. */
int tn_env_matvec_noalloc(tn_env* env, const void* v_dev, void* out_dev);
/* phi = A_pos * A_{pos+1} (update_site.jl:46) into phi_dev */
int tn_env_contract_phi(tn_env* env, int pos, void* phi_dev);
/* orthogonalize!(psi, site): exact (untruncated) gauge move on the device */
int tn_env_orthogonalize(tn_env* env, int site);

/* ---- Krylov solvers (solver.jl) --------------------------------------- */
int tn_krylov_eig(tn_env* env, void* phi_dev, int krylovdim, int maxiter, double tol,
                  int precision, double* eig, tn_krylov_info* info);
int tn_krylov_expm(tn_env* env, void* phi_dev, double t_re, double t_im, int krylovdim,
                   int maxiter, double tol, int precision, tn_krylov_info* info);
/* <phi|H|phi> from one FP64 matvec (real part in e[0], imaginary in e[1]) */
int tn_rayleigh_fp64(tn_env* env, const void* phi_dev, double* e2);

/* ---- bond replacement and truncated SVD -------------------------------- */
/* replacebond!(psi, b, phi; maxdim, mindim, cutoff, ortho, normalize,
 * eigen_perturbation = noise * noiseterm(PH, phi, ortho)) (update_site.jl:59-76;
 * the noise term is ITensorMPS noiseterm, in-tree mirror
 * projcouplingmodel.jl:391-492).  noise = 0 is the plain factorisation; a
 * nonzero noise needs the environment positioned at b with nsite = 2.
 * eigs_out receives the kept weights P (eigenvalues of the density matrix,
 * s^2 without noise; capacity eigs_cap). */
int tn_replacebond(tn_env* env, int b, const void* phi_dev, int ortho_left, long maxdim,
                   long mindim, double cutoff, double noise, int normalize, double* truncerr,
                   double* eigs_out, long eigs_cap, long* nkept);
/* Standalone truncated SVD of a real m x n row-major device matrix:
 * M = U S Vt with U (m x k) and Vt (k x n) row-major on the device, S on the
 * host (capacity min(m,n)).  Truncation rule of NDTensors truncate!!. */
int tn_svd_trunc(tn_ctx* ctx, const double* M_dev, long m, long n, long maxdim, long mindim,
                 double cutoff, double* U_dev, double* S_host, double* Vt_dev, long* nkept,
                 double* truncerr);

/* ---- fused updates ---------------------------------------------------- */
/* update_position!(sysenv, solver, pos, 2, ortho; kwargs...) */
int tn_update_two_site(tn_env* env, int pos, int ortho_left, const tn_update_params* p,
                       tn_update_result* r);
/* update_position!(sysenv, solver, pos, 1, ortho; kwargs...) = _update_one_site!
 * (src/mps/update_site.jl:94-197): the solver on site pos (orthocenter pos), then
 * the truncated SVD split of that site with S V (left) or U S (right) moved into
 * the neighbour; with reverse_step the bond matrix is evolved backward by the
 * zero-site operator first (nsite = 0); with noise > 0 the site and its
 * neighbour are factorised as a two-site tensor with the noise term.  At the
 * end of a half sweep (pos = N going left, pos = 1 going right) only the site
 * is replaced.  tn_env_set_nsite accepts 0 for the zero-site operator. */
int tn_update_one_site(tn_env* env, int pos, int ortho_left, const tn_update_params* p,
                       tn_update_result* r);
/* krylov_extend!(sysenv; extension_krylovdim, extension_applyH_cutoff,
 * extension_applyH_maxdim, extension_cutoff) (src/mps/sweep.jl:378-534):
 * global subspace expansion of the environment's MPS with the Krylov vectors
 * H^k psi (k = 1..krylovdim, ITensorMPS density-matrix MPO-MPS products with
 * maxdim applyH_maxdim (<= 0: max bond dimension + 2) and cutoff
 * applyH_cutoff), then _krylov_addbasis! with ext_cutoff; the MPS ends at
 * orthocenter 1 and every environment is dropped.  Needs an environment with
 * one MPO and no projectors (else TN_E_UNSUPPORTED). */
int tn_env_krylov_extend(tn_env* env, int krylovdim, long applyH_maxdim, double applyH_cutoff,
                         double ext_cutoff);
/* fullsweep!(sysenv, solver, 2, swdata; kwargs...): energy of the last
 * update, max truncation error, max bond dim; lasteigs of the mid bond
 * N/2 (right half-sweep) in mid_eigs (capacity cap, count in *nmid). */
int tn_fullsweep(tn_env* env, const tn_update_params* p, double* energy, double* maxtruncerr,
                 long* maxchi, double* mid_eigs, long cap, long* nmid);

/* ---- tree tensor networks (StateEnvsTTN, the src/ttn/ files) ------------ */
/* A tn_ttn holds the node tensors (three legs each, row-major A[a0][a1][a2],
 * FP64) and the environments of a TTN on the device.  The Hamiltonian enters
 * as a tree operator: every link carries `nchan` channels, every site leg the
 * channel operators of its site, and every node a sparse coupling
 * W[c0][c1][c2] between the channels of its three slots.  Slots 0 and 1 hold
 * children (nodes or sites), slot 2 the parent (the root: a third child).
 * Nodes are numbered 0..nnodes-1 by the caller. */
typedef struct tn_ttn tn_ttn;
/* TTN / StateEnvsTTN storage (src/ttn/ttn.jl:1-30, state_envs_ttn.jl:1-30) */
int tn_ttn_create(tn_ctx* ctx, int nnodes, tn_ttn** out);
int tn_ttn_destroy(tn_ttn* ttn);
/* link between slot sa of node a and slot sb of node b (graph edge, src/base/graph.jl:100-125) */
int tn_ttn_set_link(tn_ttn* ttn, int a, int sa, int b, int sb, int nchan);
/* site leg at slot 0/1: ops[c][bra][ket] = <bra|X_c|ket>, d x d per channel
 * (the CouplingModel terms of that site, src/base/couplingmodel.jl:153-241) */
int tn_ttn_set_site(tn_ttn* ttn, int node, int slot, int d, int nchan, const double* ops_host);
/* the node's tree-operator coupling: nnz entries (c0[k], c1[k], c2[k], w[k]) */
int tn_ttn_set_coupling(tn_ttn* ttn, int node, int nnz, const int* c0, const int* c1, const int* c2,
                        const double* w);
/* ttn[node] = A / ttn[node] (src/ttn/ttn.jl:60-90) */
int tn_ttn_set_node(tn_ttn* ttn, int node, const long* dims3, const double* A_host);
int tn_ttn_get_node(tn_ttn* ttn, int node, long* dims3, double* A_host /* nullable */);
int tn_ttn_node_ptr(tn_ttn* ttn, int node, double** A_dev);
/* Penalty projectors of StateEnvsTTN(psi, H, Ms; weight) (state_envs_ttn.jl:36-70,
 * EnvCouplingModelProjTTN, environment.jl:42-110; LinkProjTTN, linkproj.jl):
 * node products add weight * m <m, v> per projector TTN M, with m the
 * projection of M on the node's local space through the overlap environments
 * of its links.  set_projectors allocates nproj projectors (weight > 0, else
 * TN_E_INVALID with the reference's message); set_projector_node uploads node
 * `node` of projector j (0-based; same tree, any bond dimensions, FP64).  Set
 * them before the environments are built: every tn_ttn_update_env also moves
 * the projectors' overlap environments (move_linkproj_to_next!). */
int tn_ttn_set_projectors(tn_ttn* ttn, int nproj, double weight);
int tn_ttn_set_projector_node(tn_ttn* ttn, int j, int node, const long* dims, const double* A_host);
/* move_linktensors_to_next!(env, psi, node, neighbour at `slot`) (src/ttn/linktensors.jl:63-118) */
int tn_ttn_update_env(tn_ttn* ttn, int node, int slot);
/* product(env, psi, v) at the orthogonality center `node` (src/ttn/linktensors.jl:183-221) */
int tn_ttn_matvec(tn_ttn* ttn, int node, const double* v_dev, double* out_dev);
/* eig_solver at `node` (src/ttn/update_site_ttn.jl:51-62, src/base/solver.jl:23-43):
 * Lanczos from the node tensor, which receives the Ritz vector */
int tn_ttn_eig(tn_ttn* ttn, int node, int krylovdim, int maxiter, double tol, double* energy,
               tn_krylov_info* info);
/* moveisometry_to_next!(ttn, node, neighbour at `slot`; maxdim, mindim, cutoff)
 * (src/ttn/ttn.jl:266-310): QR, or truncating SVD when cutoff > 0 or the bond
 * exceeds maxdim; the neighbour absorbs the remainder (not normalised) */
int tn_ttn_move(tn_ttn* ttn, int node, int slot, long maxdim, long mindim, double cutoff,
                double* truncerr, long* newdim);
/* normalize!(ttn) at `node` (src/ttn/ttn.jl:204-215); norm receives the old norm */
int tn_ttn_normalize(tn_ttn* ttn, int node, double* norm);
int tn_ttn_node_norm(tn_ttn* ttn, int node, double* norm);
/* subspace_expand! (src/ttn/update_site_ttn.jl:65-109): grow the bond at
 * `slot` of node and its neighbour by `pad`, appending padA / padB (each the
 * node's tensor shape with that leg replaced by pad, row-major) */
int tn_ttn_expand(tn_ttn* ttn, int node, int slot, long pad, const double* padA_host,
                  const double* padB_host);
/* orthogonality center of the TTN (ttn.orthocenter, src/ttn/ttn.jl:100-110); -1: none */
int tn_ttn_set_center(tn_ttn* ttn, int node);
int tn_ttn_get_center(tn_ttn* ttn, int* node);
/* seed of the subspace-expansion padding (SplitMix64 -> Box-Muller normals;
 * the reference draws it with Julia's RNG, update_site_ttn.jl:91,100) */
int tn_ttn_seed(tn_ttn* ttn, unsigned long long seed);
/* position!(sysenv, node; maxdim, mindim, cutoff, normalize, node_to_skip)
 * (src/ttn/state_envs_ttn.jl:120-140): isometry moves along the path (QR, or
 * SVD when cutoff > 0 or a bond exceeds maxdim), normalisation, then the
 * environments along the same path except from `skip` (-1: none) */
int tn_ttn_position(tn_ttn* ttn, int node, long maxdim, long mindim, double cutoff, int normalize, int skip);
/* update_position!(sysenv, eig_solver, node; ...) (src/ttn/update_site_ttn.jl:42-63) */
int tn_ttn_update_position(tn_ttn* ttn, int node, long maxdim, long mindim, double cutoff, int normalize,
                           int krylovdim, int maxiter, double tol, int check_convergence, double* energy);
/* subspace_expand!(psi, node, neighbour at `slot`, max_expand_dim, noise)
 * (src/ttn/update_site_ttn.jl:65-109) with padding from the seeded generator */
int tn_ttn_expand_random(tn_ttn* ttn, int node, int slot, long max_expand_dim, double noise);
/* fullsweep!(sysenv, sweeppath, eig_solver, swdata; ...) (src/ttn/sweep_ttn.jl:117-274):
 * path and its reverse; with noise >= 1e-13 the expansion branch toward
 * `central` (find_eccentric_central_node of the graph).  energy receives the
 * last update's Ritz value, maxchi the largest bond dimension. */
int tn_ttn_fullsweep(tn_ttn* ttn, const int* path, int npath, int central, long maxdim, long mindim,
                     double cutoff, double noise, int expand_dim, int max_expand_dim, int expand_numiter,
                     int krylovdim, int maxiter, double tol, int normalize, int check_convergence,
                     double* energy, long* maxchi);

#ifdef __cplusplus
}
#endif
#endif /* TNHIP_H */
