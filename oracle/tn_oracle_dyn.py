"""FP64 CPU oracle for one-site updates, one-site and dynamic sweeps, and the
global subspace expansion (numpy).

TEST INFRASTRUCTURE ONLY.  This module is a checker for the HIP path; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
may import it.  The product path never routes through it.

What it restates (reference = TenNetLib.jl 0.2.2):

* ``_update_one_site!`` (``src/mps/update_site.jl:94-197``): solver on the
  one-site tensor; SVD split with truncation, the non-orthonormal factor
  moved to the next site; the backward zero-site step of TDVP; with noise,
  the two-site density-matrix factorisation of the site and its neighbour.
* ``fullsweep!(…, 1, …)`` (``src/mps/sweep.jl:94-196``),
  ``dynamic_fullsweep!`` (``sweep.jl:257-376``: two-site updates until a
  bond's smallest kept weight drops below ``eigthreshold`` or its dimension
  reaches ``maxdim``, one-site updates after), ``tdvpsweep!(…; nsite)``
  (``src/mps/tdvp.jl:247-277``).
* ``krylov_extend!`` (``sweep.jl:378-452``) and ``_krylov_addbasis!``
  (``sweep.jl:458-534``): Krylov vectors H^k psi from MPO-MPS products, then
  the right bases of psi extended site by site with the dominant part of the
  Krylov vectors' reduced density matrices outside psi's own basis.
* The MPO-MPS product of ``apply(H, psi; maxdim, cutoff)``: ITensorMPS 0.3's
  default ``"densitymatrix"`` contraction (the package is absent here; its
  algorithm is restated): left environments of <H psi|H psi>, then from the
  right, at every site the reduced density matrix of H psi with the right
  part already truncated, its dominant eigenvectors (ITensors eigen
  truncation) kept as the right-orthonormal site tensor; the first site keeps
  the rest.  Parity of this contraction against ITensorMPS itself is
  unpinned; the tests pin it by exact contraction (no truncation) and by the
  reduced density matrices of the dense vector.
"""

from __future__ import annotations

import math
from typing import List, Optional

import numpy as np

from oracle import tn_oracle as O

FLOAT64_THRESHOLD = O.FLOAT64_THRESHOLD


# ---------------------------------------------------------------------------
# gauge helpers
# ---------------------------------------------------------------------------

def orthogonalize(A: List[np.ndarray], j: int) -> List[np.ndarray]:
    """orthogonalize!(psi, j) (1-based): QR sweeps so that sites < j are
    left-orthonormal and sites > j right-orthonormal."""
    A = [a.copy() for a in A]
    N = len(A)
    for k in range(j - 1):
        a, s, b = A[k].shape
        Q, R = np.linalg.qr(A[k].reshape(a * s, b))
        A[k] = Q.reshape(a, s, Q.shape[1])
        A[k + 1] = np.tensordot(R, A[k + 1], axes=([1], [0]))
    for k in range(N - 1, j - 1, -1):
        a, s, b = A[k].shape
        Q, R = np.linalg.qr(A[k].reshape(a, s * b).conj().T)
        A[k] = Q.conj().T.reshape(Q.shape[1], s, b)
        A[k - 1] = np.tensordot(A[k - 1], R.conj().T, axes=([2], [0]))
    return A


def _truncated_eigh(rho, maxdim, mindim, cutoff):
    """ITensors eigen(ρ; ishermitian, maxdim, cutoff): eigenpairs sorted by
    decreasing value, truncated by the NDTensors rule."""
    rho = 0.5 * (rho + rho.conj().T)
    w, V = np.linalg.eigh(rho)
    w, V = w[::-1], V[:, ::-1]
    tr = O.truncate_weights(np.maximum(w, 0.0), maxdim, mindim, cutoff)
    return w[:tr.n], V[:, :tr.n]


# ---------------------------------------------------------------------------
# one-site update
# ---------------------------------------------------------------------------

def _solve(env, phi, solver, time_step, krylovdim, maxiter, tol):
    if solver == "eig":
        energy, phi, _ = O.eigsolve_lanczos(env.PH.product, phi, krylovdim, maxiter, tol)
    else:
        phi, _ = O.exponentiate_lanczos(env.PH.product, time_step, phi, krylovdim, maxiter, tol)
        energy = float("nan")
    return energy, phi


def update_one_site(env, pos: int, ortho: str, *, solver: str = "eig", time_step=None,
                    maxdim=2 ** 62, mindim=1, cutoff=FLOAT64_THRESHOLD, normalize=True,
                    krylovdim=5, maxiter=2, tol=1e-14, reverse_step=False, noise=0.0):
    """_update_one_site! (src/mps/update_site.jl:94-197); pos is 1-based.
    Returns (energy, truncerr, eigs)."""
    N = len(env)
    A = env.psi
    env.PH.nsite = 1
    phi = A[pos - 1].copy()
    env.PH.position(A, pos)
    energy, phi = _solve(env, phi, solver, time_step, krylovdim, maxiter, tol)
    if normalize:
        phi = phi / np.linalg.norm(phi)
    if math.isnan(energy):
        energy = float(np.real(np.vdot(phi, env.PH.product(phi))))
    truncerr, eigs = 0.0, np.zeros(0)
    done = (pos == 1 and ortho == "right") or (pos == N and ortho == "left")
    if done:
        A[pos - 1] = phi
        return energy, truncerr, eigs
    nxt = pos + 1 if ortho == "left" else pos - 1
    pos0 = pos if ortho == "left" else pos - 1
    if abs(noise) > FLOAT64_THRESHOLD:
        phi2 = np.tensordot(phi, A[pos], axes=([2], [0])) if ortho == "left" else \
            np.tensordot(A[pos - 2], phi, axes=([2], [0]))
        env.PH.nsite = 2
        env.PH.position(A, pos0)
        drho = noise * O.noiseterm(env.PH, phi2, ortho)
        spec = O.replacebond_perturbed(A, pos0 - 1, phi2, maxdim, mindim, cutoff, ortho, normalize, drho)
        env.PH.nsite = 1
        env.ortho_center = pos0 + 1 if ortho == "left" else pos0
        return energy, spec.truncerr, spec.eigs
    a, s, b = phi.shape
    if ortho == "left":
        U, S, Vh, spec = O.svd_truncate(phi.reshape(a * s, b), maxdim, mindim, cutoff)
        k = S.size
        if normalize:
            S = S / np.linalg.norm(S)
        A[pos - 1] = U.reshape(a, s, k)
        C = S[:, None] * Vh                                     # [k, b]
    else:
        U, S, Vh, spec = O.svd_truncate(phi.reshape(a, s * b).T, maxdim, mindim, cutoff)
        k = S.size
        if normalize:
            S = S / np.linalg.norm(S)
        A[pos - 1] = U.T.reshape(k, s, b)                       # rows (s b) orthonormal
        C = (S[:, None] * Vh).T                                 # [a, k]
    truncerr, eigs = spec.truncerr, spec.eigs
    if reverse_step:
        pos1 = pos + 1 if ortho == "left" else pos
        env.PH.nsite = 0
        env.PH.position(A, pos1)
        C, _ = O.exponentiate_lanczos(env.PH.product, -time_step, C, krylovdim, maxiter, tol)
        if normalize:
            C = C / np.linalg.norm(C)
        energy = float(np.real(np.vdot(C, env.PH.product(C))))
    if ortho == "left":
        A[nxt - 1] = np.tensordot(C, A[nxt - 1], axes=([1], [0]))
    else:
        A[nxt - 1] = np.tensordot(A[nxt - 1], C, axes=([2], [0]))
    env.ortho_center = nxt
    return energy, truncerr, eigs


def update_position(env, pos, nsite, ortho, **kw):
    """update_position!(sysenv, solver, pos, nsite, ortho; kwargs...)
    (src/mps/update_site.jl:231-277)."""
    if kw.get("noise", 0.0) > 0 and kw.get("reverse_step", False):
        raise ValueError("noise cannot be greater than zero for reverse_step")
    if nsite == 2:
        return O.update_two_site(env, pos, ortho, **kw)
    if nsite == 1:
        return update_one_site(env, pos, ortho, **kw)
    raise ValueError(f"`update_position()` with `nsite={nsite}` not implemented !!")


# ---------------------------------------------------------------------------
# sweeps
# ---------------------------------------------------------------------------

def _record(env, swdata, energy, lasteigs, maxtruncerr):
    N = len(env)
    swdata.maxchi.append(max(a.shape[2] for a in env.psi))
    swdata.energy.append(energy)
    mid = lasteigs[N // 2 - 1]
    swdata.entropy.append(O.entropy(mid / mid.sum()))
    swdata.maxtruncerr.append(maxtruncerr)
    swdata.lasteigs = lasteigs
    if swdata.sweepcount > 1:
        return swdata.energy[-1] - swdata.energy[-2], swdata.entropy[-1] - swdata.entropy[-2]
    return float("nan"), float("nan")


def fullsweep(env, swdata, nsite: int, **kw):
    """fullsweep!(sysenv, solver, nsite, swdata; kwargs...) (src/mps/sweep.jl:94-196)."""
    if nsite == 2:
        return O.fullsweep(env, swdata, **kw)
    O._ensure_ortho_first(env)
    N = len(env)
    energy, maxtruncerr = float("nan"), 0.0
    swdata.sweepcount += 1
    lasteigs: List[np.ndarray] = [None] * (N - 1)
    for bond in range(1, N):
        energy, err, _ = update_one_site(env, bond, "left", **kw)
        maxtruncerr = max(err, maxtruncerr)
        if bond == N - 1:
            energy, _, _ = update_one_site(env, bond + 1, "left", **kw)
    for bond in range(N - 1, 0, -1):
        energy, err, eigs = update_one_site(env, bond + 1, "right", **kw)
        lasteigs[bond - 1] = eigs
        maxtruncerr = max(err, maxtruncerr)
        if bond == 1:
            energy, _, _ = update_one_site(env, bond, "right", **kw)
    env.ortho_center = 1
    return _record(env, swdata, energy, lasteigs, maxtruncerr)


def _is_plain_mpo(env) -> bool:
    return isinstance(env.PH, O.Envs) and not env.PH.Ms


def dynamic_fullsweep(env, swdata, *, eigthreshold=1e-12, extendat=None, maxdim=2 ** 62,
                      extension=None, **kw):
    """dynamic_fullsweep! (src/mps/sweep.jl:257-376).  ``extension`` holds the
    keyword arguments of krylov_extend!."""
    extension = extension or {}
    if swdata.sweepcount == 0 or (extendat is not None and (swdata.sweepcount + 1) % extendat == 0):
        if _is_plain_mpo(env):
            krylov_extend(env, **extension)
            return fullsweep(env, swdata, 1, maxdim=maxdim, **kw)
        return fullsweep(env, swdata, 2, maxdim=maxdim, **kw)
    O._ensure_ortho_first(env)
    N = len(env)
    energy, maxtruncerr = float("nan"), 0.0
    swdata.sweepcount += 1
    lasteigs: List[np.ndarray] = [None] * (N - 1)

    def pick(bond):
        small = swdata.lasteigs[bond - 1][-1] < eigthreshold
        return 1 if (small or env.psi[bond - 1].shape[2] >= maxdim) else 2

    for bond in range(1, N):
        nsite = pick(bond)
        energy, err, eigs = update_position(env, bond, nsite, "left", maxdim=maxdim, **kw)
        lasteigs[bond - 1] = eigs
        maxtruncerr = max(err, maxtruncerr)
        if nsite == 1 and bond == N - 1:
            energy, _, _ = update_one_site(env, bond + 1, "left", maxdim=maxdim, **kw)
    swdata.lasteigs = list(lasteigs)
    for bond in range(N - 1, 0, -1):
        nsite = pick(bond)
        site = bond + 1 if nsite == 1 else bond
        energy, err, eigs = update_position(env, site, nsite, "right", maxdim=maxdim, **kw)
        lasteigs[bond - 1] = eigs
        maxtruncerr = max(err, maxtruncerr)
        if nsite == 1 and bond == 1:
            energy, _, _ = update_one_site(env, bond, "right", maxdim=maxdim, **kw)
    env.ortho_center = 1
    return _record(env, swdata, energy, lasteigs, maxtruncerr)


def tdvpsweep(env, swdata, time_step, nsite="dynamic", **kw):
    """tdvpsweep!(engine, time_step; nsite, kwargs...) (src/mps/tdvp.jl:247-277)."""
    kw = dict(kw)
    kw.setdefault("krylovdim", 30)
    kw.setdefault("maxiter", 100)
    kw.setdefault("tol", 1e-12)
    if nsite == "dynamic":
        return dynamic_fullsweep(env, swdata, solver="exp", time_step=0.5 * time_step, reverse_step=True, **kw)
    if kw.get("extendat") is not None:
        raise ValueError("`extendat` must be `nothing` for `nsite == 2` or `nsite == 1`")
    kw.pop("extendat", None)
    kw.pop("extension", None)
    kw.pop("eigthreshold", None)
    return fullsweep(env, swdata, nsite, solver="exp", time_step=0.5 * time_step, reverse_step=True, **kw)


# ---------------------------------------------------------------------------
# MPO-MPS product (ITensorMPS "densitymatrix" contraction)
# ---------------------------------------------------------------------------

def apply_mpo(Ws: List[np.ndarray], psi: List[np.ndarray], maxdim: int, cutoff: float,
              mindim: int = 1) -> List[np.ndarray]:
    """H psi as an MPS whose sites 2..N are right-orthonormal: the
    dominant eigenvectors of the reduced density matrices of H psi, from the
    right, with the right part already truncated."""
    N = len(psi)
    Wc = [W.conj() for W in Ws]
    psic = [a.conj() for a in psi]
    # E[j][a, w, w', a'] over sites 1..j (ket psi, H, conj H, bra psi)
    E = [None] * N
    e = np.ones((1, 1, 1, 1), dtype=np.result_type(psi[0], Ws[0]))
    for j in range(N - 1):
        e = np.einsum("awvx,asb,wtsy,vtrz,xrc->byzc", e, psi[j], Ws[j], Wc[j], psic[j])
        E[j] = e
    out: List[np.ndarray] = [None] * N
    # R[a, w, t, k]: psi and H of sites j..N contracted, outputs t_j and the new link k
    R = np.einsum("asb,wtsv->awtbv", psi[N - 1], Ws[N - 1])[:, :, :, 0, 0][..., None]
    for j in range(N - 1, 0, -1):
        a, w, t, k = R.shape
        Rm = R.reshape(a * w, t * k)
        El = E[j - 1]                                    # [a, w, w', a']
        Z = np.einsum("awvx,awq->vxq", El, R.reshape(a, w, t * k))   # [w', a', (t k)]
        rho = np.einsum("xvp,vxq->pq", R.reshape(a, w, t * k).conj(), Z)   # [(t' k'), (t k)]
        rho = rho.T  # rho[(t k), (t' k')] with ket first
        _, V = _truncated_eigh(rho, maxdim, mindim, cutoff)
        Ut = V.T                                          # rows: kept eigenvectors over (t k)
        kn = Ut.shape[0]
        out[j] = Ut.reshape(kn, t, k)
        Q = np.einsum("awq,nq->awn", R.reshape(a, w, t * k), Ut.conj())     # [a, w, kn]
        R = np.einsum("zsa,yrsw,awn->zyrn", psi[j - 1], Ws[j - 1], Q)       # [a0, w0, t, kn]
        del Rm
    out[0] = R[:, :, :, :].reshape(R.shape[0] * R.shape[1], R.shape[2], R.shape[3])
    return out


# ---------------------------------------------------------------------------
# global subspace expansion
# ---------------------------------------------------------------------------

def krylov_addbasis(psi: List[np.ndarray], phis: List[List[np.ndarray]], extension_cutoff: float):
    """_krylov_addbasis! (src/mps/sweep.jl:458-534); returns the new psi
    (orthocenter 1)."""
    N = len(psi)
    psi = orthogonalize(psi, N)
    phis = [orthogonalize(p, N) for p in phis]
    for j in range(N, 1, -1):
        a, s, b = psi[j - 1].shape
        M = psi[j - 1].reshape(a, s * b)
        _, _, Vh = np.linalg.svd(M, full_matrices=False)
        B = Vh                                              # rows span psi[j]'s row space
        P = np.eye(s * b) - B.T @ B.conj()                  # P[x', x]
        rho = sum(np.einsum("ax,ay->xy", p[j - 1].reshape(p[j - 1].shape[0], -1),
                            p[j - 1].reshape(p[j - 1].shape[0], -1).conj()) for p in phis)
        rho = rho / np.trace(rho)
        PrP = P @ rho @ P
        if np.linalg.norm(PrP) > 1e-14:
            _, V = _truncated_eigh(PrP, 2 ** 62, 1, extension_cutoff)
            Bx = np.concatenate([B, V.T], axis=0)
        else:
            Bx = B
        kx = Bx.shape[0]
        Bx3 = Bx.reshape(kx, s, b)
        for st in [psi] + phis:
            T = np.einsum("asb,ksb->ak", st[j - 1].reshape(-1, s, b), Bx3.conj())
            st[j - 2] = np.tensordot(st[j - 2], T, axes=([2], [0]))
            st[j - 1] = Bx3.copy()
    return orthogonalize(psi, 1)


def krylov_extend(env, krylovdim: int = 3, applyH_cutoff: float = FLOAT64_THRESHOLD,
                  applyH_maxdim: Optional[int] = None, extension_cutoff: float = 1e-7):
    """krylov_extend!(sysenv; kwargs...) (src/mps/sweep.jl:378-452) for a
    StateEnvs over one MPO."""
    if not _is_plain_mpo(env):
        raise ValueError("krylov_extend!: the StateEnvs must be created by a single MPO")
    psi = env.psi
    if applyH_maxdim is None:
        applyH_maxdim = max(a.shape[2] for a in psi) + 2
    phis, prev = [], psi
    for _ in range(krylovdim):
        ph = apply_mpo(env.PH.W, prev, applyH_maxdim, applyH_cutoff)
        ph[0] = ph[0] / np.linalg.norm(ph[0])
        phis.append(ph)
        prev = ph
    env.psi = krylov_addbasis(psi, phis, extension_cutoff)
    env.ortho_center = 1
    env.PH = env.PH.fresh()
