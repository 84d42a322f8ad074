"""Benchmark: DMRG sweep time and matvec MFMA utilisation on one MI355X.

    python bench.py [--steps K] [--warmup W] [--chi 1024] [--N 100] [--prec bf16x3]

Workload (Heisenberg chain, open boundaries, d=2, MPO bond dim w=5):
  * warmup: W two-site DMRG sweeps from the Neel state with maxdim ramped
    64, 128, ... up to --chi (cutoff 1e-12, eig_solver defaults krylovdim=5,
    maxiter=2), mixed precision; W = 7 plus the timed sweep gives the
    8-sweep ramp 64..1024 of the Heisenberg N = 100 workload;
  * timed: K sweeps at maxdim --chi in the mixed form (matvec --prec, FP64
    Lanczos / SVD / final Rayleigh quotient) -> `value` (seconds per sweep);
  * the same K sweeps from the same snapshot in the plain FP64 GPU form,
    reported beside it with |dE/N| between the two;
  * matvec microbenchmark at chi=4096 (dense, d=2, w=5) in FP64, FP32 and
    the mixed form: per-GEMM-launch HIP-event timings on the library's own
    stream give the roofline of the dominant kernel (matvec GEMM).
  * cpu_baseline: the FP64 numpy oracle (OpenBLAS, all host threads) running
    the same timed sweep from the same snapshot.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "tennetlib.jl_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # dense MFMA TFLOP/s, MI355X
METRIC = "DMRG sweep s + matvec %MFMA peak, mixed vs FP64 GPU at |ΔE/N|≤1e-9, χ=4096"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def heis_mpo(N):
    import tennetlib_amd as T
    return T.heisenberg_mpo(N)


def neel(N):
    import tennetlib_amd as T
    return T.neel_state(N)


def matvec_flops(chil, chir, d, wl, wr=None):
    """Two-site matvec: the two GEMMs (L and R sides) plus the W steps."""
    wr = wl if wr is None else wr
    return 2.0 * chil * chir * d * d * (wl * chil + wr * chir) + 4.0 * chil * chir * d ** 3 * wl * wr


def phase(ctx, ph):
    ms, n = C.c_double(), C.c_long()
    ctx.check(ctx.lib.tn_ctx_phase_stats(ctx.h, ph, C.byref(ms), C.byref(n)))
    return ms.value, n.value


def wide_mpo4():
    """Four bulk sites (35..38) of the J1-J2 Ly = 6 x Lx = 12 cylinder MPO
    (inner bonds w = 20), outer bonds closed with fixed random vectors."""
    import tennetlib_amd as T
    Ws = T.j1j2_cylinder_mpo(6, 12, 1.0, 0.5)[35:39]
    rng = np.random.default_rng(11)
    Ws[0] = np.tensordot(rng.uniform(-1, 1, Ws[0].shape[0]), Ws[0], axes=([0], [0]))[None]
    Ws[3] = np.tensordot(Ws[3], rng.uniform(-1, 1, Ws[3].shape[3]), axes=([3], [0]))[..., None]
    return [np.ascontiguousarray(w) for w in Ws]


def matvec_microbench(T, chi, iters, precs, Ws=None):
    """Two-site matvec at (chi, 2, 2, chi) at position 2 of a 4-site MPO
    (default: Heisenberg, w = 5); L/R come from random (seed 7, U[-1,1])
    edge tensors."""
    ctx = T.default_context()
    rng = np.random.default_rng(7)
    N, d = 4, 2
    A = [rng.uniform(-1, 1, (1, d, chi)), rng.uniform(-1, 1, (chi, d, chi)) / chi,
         rng.uniform(-1, 1, (chi, d, chi)) / chi, rng.uniform(-1, 1, (chi, d, 1))]
    Ws = heis_mpo(N) if Ws is None else Ws
    wl, wr = Ws[1].shape[0], Ws[2].shape[3]  # environment bonds at position 2
    env = T.StateEnvs(A, Ws, ctx=ctx)
    T.position_(env, 2)
    psi = T.DeviceArray.from_numpy(ctx, rng.uniform(-1, 1, (chi, d, d, chi)))
    out = T.DeviceArray(ctx, (chi, d, d, chi))
    F = matvec_flops(chi, chi, d, wl, wr)
    Fgemm = 2.0 * 0.5 * (wl + wr) * d * d * chi ** 3  # mean of the two big GEMM launches
    res = {}
    for prec in precs:
        code = T._lib.PRECISIONS[prec]
        ctx.check(ctx.lib.tn_env_matvec(env.h, psi.ptr, out.ptr, code))  # warm (shadows, code objects)
        ctx.enable_timing(True)
        ctx.timers(reset=True)
        t0 = time.perf_counter()
        for _ in range(iters):
            ctx.check(ctx.lib.tn_env_matvec(env.h, psi.ptr, out.ptr, code))
        wall = (time.perf_counter() - t0) / iters
        mv_ms, mv_n = phase(ctx, 0)
        g_ms, g_n = phase(ctx, 4)
        ctx.enable_timing(False)
        g_avg = g_ms / max(g_n, 1)
        mv_avg = mv_ms / max(mv_n, 1)
        res[prec] = dict(matvec_ms=mv_avg, wall_ms=wall * 1e3, matvec_tflops=F / (mv_avg * 1e-3) / 1e12,
                         gemm_launch_ms=g_avg, gemm_tflops=Fgemm / (g_avg * 1e-3) / 1e12)
        log(f"matvec chi={chi} w={wl},{wr} {prec}: {mv_avg:.2f} ms/matvec ({res[prec]['matvec_tflops']:.1f} TF/s), "
            f"GEMM launch {g_avg:.2f} ms ({res[prec]['gemm_tflops']:.1f} TF/s)")
    psi.free()
    out.free()
    env.close()
    return res, F, Fgemm


def sweeps(T, psi, N, chi, nsweeps, prec, cutoff, kd, mi):
    env = T.StateEnvs(psi, heis_mpo(N))
    sw = T.SweepData()
    ctx = env.ctx
    ctx.enable_timing(True)
    ctx.timers(reset=True)
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    t0 = time.perf_counter()
    for _ in range(nsweeps):
        T.fullsweep_(env, T.eig_solver, 2, sw, maxdim=chi, cutoff=cutoff, solver_krylovdim=kd,
                     solver_maxiter=mi, precision=prec, outputlevel=0)
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    dt = (time.perf_counter() - t0) / nsweeps
    tm = ctx.timers(reset=True)
    ctx.enable_timing(False)
    return env, sw, dt, tm


def tdvp_run(T, N, steps, dt, maxdim, cutoff):
    """Two-site TDVP of the XXZ chain (Delta = 0.5) from the Neel state,
    time_step = -i dt per tdvpsweep!; returns s/step, energy drift, max bond."""
    env = T.TDVPEngine([a.astype(complex) for a in T.neel_state(N)], T.xxz_mpo(N, 1.0, 0.5))
    ctx = env.sysenv.ctx
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    t0 = time.perf_counter()
    for _ in range(steps):
        T.tdvpsweep_(env, -1j * dt, 2, maxdim=maxdim, cutoff=cutoff, outputlevel=0)
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    el = time.perf_counter() - t0
    e_neel = -(N - 1) * 0.5 / 4
    out = {"workload": f"TDVP XXZ N={N} Delta=0.5 from Neel, dt={dt}, {steps} tdvpsweep! calls, maxdim {maxdim}, "
                       f"cutoff {cutoff:g}, complex128",
           "s_per_step": round(el / steps, 4), "steps": steps,
           "energy_drift": abs(env.swdata.energy[-1] - e_neel), "maxchi": env.swdata.maxchi[-1]}
    env.sysenv.close()
    return out


def ttn_instances(T, count, seed=5, total=256):
    """Many-instance TTN ground states: `total` random-field Heisenberg chains
    (h_i ~ U[-2, 2], N ~ U{16..64}, numpy default_rng(seed)), each a
    default binary-tree TTN from default_randomTTN(N, 12) optimised with the
    reference test's schedule (maxdim 128, 5 sweeps, cutoff 1e-14, noise
    1e-2, noisedecay 5, disable_noise_after 2).  The first `count` instances
    run one after another on the library's stream; the projection to all
    instances scales the wall time by the sum of N."""
    rng = np.random.default_rng(seed)
    Ns = [int(n) for n in rng.integers(16, 65, total)]
    hs = [rng.uniform(-2, 2, n) for n in Ns]
    params = T.make_optimize_params(maxdim=[128], nsweeps=[5], cutoff=1e-14, noise=1e-2, noisedecay=5,
                                    disable_noise_after=2)
    res = []
    ctx = T.default_context()
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    t0 = time.perf_counter()
    for i in range(count):
        N = Ns[i]
        os_ = T.OpStrings()
        for j in range(1, N):
            os_ += (1.0, "Sz", j, "Sz", j + 1)
            os_ += (0.5, "S+", j, "S-", j + 1)
            os_ += (0.5, "S-", j, "S+", j + 1)
        for j in range(1, N + 1):
            os_ += (float(hs[i][j - 1]), "Sz", j)
        psi0 = T.default_randomTTN(N, 12, seed=i)
        env = T.StateEnvsTTN(psi0, T.CouplingModel(os_, N))
        sw = T.optimize_(env, params, T.default_sweeppath(psi0), outputlevel=0, seed=i)
        env.close()
        res.append((sw.energy[-1], sw.maxchi[-1]))
    ctx.check(ctx.lib.tn_ctx_synchronize(ctx.h))
    dt = time.perf_counter() - t0
    done_n = sum(Ns[:count])
    return {"workload": f"{total} random-field Heisenberg chains, N~U{{16..64}}, h~U[-2,2] (seed {seed}); "
                        f"binary-tree TTN from default_randomTTN(N, 12), maxdim 128, 5 sweeps, noise 1e-2 "
                        f"(decay 5, off after 2), cutoff 1e-14; FP64",
            "instances_timed": count, "wall_s": round(dt, 3), "instances_per_s": round(count / dt, 4),
            "site_sweeps_per_s": round(5 * done_n / dt, 1),
            "projected_s_for_all": round(dt * sum(Ns) / done_n, 1),
            "energy_per_site": [round(res[i][0] / Ns[i], 10) for i in range(count)],
            "maxchi": [res[i][1] for i in range(count)]}


def cpu_sweep(snap, N, chi, cutoff, kd, mi):
    """The FP64 numpy oracle running the same timed sweep (same snapshot,
    maxdim, cutoff and Krylov settings) on the host cores."""
    from oracle import tn_oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:
        cores = os.cpu_count() or 1
    env = O.StateEnvs([np.array(a) for a in snap], O.heisenberg_mpo(N))
    sw = O.SweepData()
    t0 = time.perf_counter()
    O.fullsweep(env, sw, maxdim=chi, cutoff=cutoff, krylovdim=kd, maxiter=mi)
    return time.perf_counter() - t0, cores, sw.energy[-1]


# HBM bytes per launch of the FP32 matvec GEMM at chi = 4096 (mean of its two
# launches), from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a plain copy:
# profiles/matvec_gemm_chi4096_pmc_traffic.md
GEMM_TRAFFIC = {("fp32", 4096): 17.2e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=7)
    ap.add_argument("--N", type=int, default=100)
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--mv-chi", type=int, default=4096)
    ap.add_argument("--mv-iters", type=int, default=5)
    ap.add_argument("--prec", default="fp32", choices=["fp32", "bf16x3"])
    ap.add_argument("--cutoff", type=float, default=1e-12)
    ap.add_argument("--krylovdim", type=int, default=5)
    ap.add_argument("--maxiter", type=int, default=2)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--mv-only", action="store_true", help="matvec microbenchmark only")
    ap.add_argument("--mv-wide", type=int, default=1, help="also time the w = 20 (J1-J2 cylinder) matvec")
    ap.add_argument("--tdvp-steps", type=int, default=40, help="TDVP steps of the XXZ N=100 run (0: skip)")
    ap.add_argument("--ttn-instances", type=int, default=16, help="TTN instances timed of the 256 (0: skip)")
    ap.add_argument("--ttn-only", action="store_true", help="only the many-instance TTN line")
    ap.add_argument("--tdvp-only", action="store_true", help="only the TDVP XXZ N=100 line")
    ap.add_argument("--copy-ref", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (one MI355X); multi-GPU sharding of the "
              "environments is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.copy_ref:
        # calibration for the FETCH_SIZE / WRITE_SIZE counters: a plain device
        # copy of a known size (reads and writes copy_ref MiB per launch)
        import torch
        n = args.copy_ref * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_ref_mib": args.copy_ref, "launches": 5}), flush=True)
        return

    import tennetlib_amd as T
    T.default_context()

    # 1. matvec microbenchmark at chi = mv_chi
    if args.ttn_only:
        print(json.dumps({"ttn_random_field_batch": ttn_instances(T, max(1, args.ttn_instances))}), flush=True)
        return
    if args.tdvp_only:
        print(json.dumps({"tdvp_xxz_n100": tdvp_run(T, 100, max(1, args.tdvp_steps), 0.05, 512, 1e-12)}), flush=True)
        return
    mv, F_mv, F_gemm = matvec_microbench(T, args.mv_chi, args.mv_iters, ["fp64", "fp32", "bf16x3"])
    mv_wide = None
    if args.mv_wide:
        # the wide-MPO matvec (J1-J2 cylinder bulk, w = 20) at the same chi
        mv_wide, F_wide, _ = matvec_microbench(T, args.mv_chi, max(2, args.mv_iters // 2), ["fp64", "fp32"],
                                               Ws=wide_mpo4())
    if args.mv_only:
        print(json.dumps({"matvec": mv, "matvec_wide": mv_wide, "chi": args.mv_chi}), flush=True)
        return

    # 2. DMRG: warmup ramp, then timed sweeps in the mixed and FP64 forms
    N, chi = args.N, args.chi
    ramp = []
    m = 64
    while len(ramp) < args.warmup:
        ramp.append(min(m, chi))
        m *= 2
    env = T.StateEnvs(neel(N), heis_mpo(N))
    sw = T.SweepData()
    t0 = time.perf_counter()
    for md in ramp:
        T.fullsweep_(env, T.eig_solver, 2, sw, maxdim=md, cutoff=args.cutoff, solver_krylovdim=args.krylovdim,
                     solver_maxiter=args.maxiter, precision=args.prec, outputlevel=0)
        log(f"warmup sweep maxdim={md}: E={sw.energy[-1]:.12f} maxchi={sw.maxchi[-1]} "
            f"t={time.perf_counter() - t0:.1f}s")
    snap = env.psi
    env.close()

    env_m, sw_m, dt_m, tm_m = sweeps(T, snap, N, chi, args.steps, args.prec, args.cutoff, args.krylovdim,
                                     args.maxiter)
    log(f"mixed ({args.prec}) sweep: {dt_m:.3f} s, E={sw_m.energy[-1]:.14f}, maxchi={sw_m.maxchi[-1]}, "
        f"phases(ms)={ {k: round(v, 1) for k, v in tm_m.items()} }")
    E_m = sw_m.energy[-1]
    chi_m = sw_m.maxchi[-1]
    env_m.close()
    fp64 = None
    if not args.skip_fp64:
        env_d, sw_d, dt_d, tm_d = sweeps(T, snap, N, chi, args.steps, "fp64", args.cutoff, args.krylovdim,
                                         args.maxiter)
        log(f"fp64 sweep: {dt_d:.3f} s, E={sw_d.energy[-1]:.14f}, maxchi={sw_d.maxchi[-1]}, "
            f"phases(ms)={ {k: round(v, 1) for k, v in tm_d.items()} }")
        fp64 = dict(sweep_s=dt_d, energy=sw_d.energy[-1], maxchi=sw_d.maxchi[-1],
                    phase_ms={k: round(v, 2) for k, v in tm_d.items()})
        env_d.close()

    tdvp = None
    if args.tdvp_steps > 0:
        tdvp = tdvp_run(T, 100, args.tdvp_steps, 0.05, 512, 1e-12)
        log(f"tdvp: {tdvp}")

    ttn = None
    if args.ttn_instances > 0:
        ttn = ttn_instances(T, args.ttn_instances)
        log(f"ttn: {ttn['instances_timed']} instances in {ttn['wall_s']} s")

    cpu = None
    if not args.skip_cpu:
        t_cpu, cores, e_cpu = cpu_sweep(snap, N, chi, args.cutoff, args.krylovdim, args.maxiter)
        cpu = {"value": round(t_cpu, 3), "unit": "s/sweep", "cores": cores, "kind": "port",
               "sample": f"numpy FP64 oracle (restatement, not the Julia reference), the same sweep: "
                         f"{2 * (N - 1)} two-site updates from the same snapshot at maxdim {chi}, "
                         f"cutoff {args.cutoff:g}; energy {e_cpu:.12f}"}
        log(f"cpu baseline: {t_cpu:.2f} s/sweep on {cores} threads, E={e_cpu:.14f}")

    pk = PEAK["fp32"]
    g = mv[args.prec]
    line = {
        "metric": METRIC,
        "value": round(dt_m, 4),
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_m * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.prec} matvec (MFMA), fp64 Krylov / factorisation / energy refinement",
        "data": "synthetic (Neel start, Heisenberg MPO built in-tree)",
        "config": {"workload": f"two-site DMRG sweep, Heisenberg OBC N={N}, d=2, w=5, maxdim {chi} "
                               f"(ramp {ramp}), cutoff {args.cutoff:g}, krylovdim {args.krylovdim}, "
                               f"maxiter {args.maxiter}; matvec microbench chi={args.mv_chi}",
                   "maxchi": chi_m, "parallelism": "single GPU"},
        "fp64_gpu": fp64,
        "speedup_vs_fp64": (round(fp64["sweep_s"] / dt_m, 3) if fp64 else None),
        "dE_per_site_vs_fp64": (abs(E_m - fp64["energy"]) / N if fp64 else None),
        "energy_mixed": E_m,
        "matvec_chi4096": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in mv.items()},
        "matvec_pct_fp32_peak": round(100 * mv[args.prec]["matvec_tflops"] / pk, 2),
        "matvec_chi4096_w20": ({k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in mv_wide.items()}
                               if mv_wide else None),
        "roofline": {"bound": "mfma", "achieved": round(g["gemm_tflops"], 2), "peak": pk, "unit": "TFLOP/s",
                     "frac": round(g["gemm_tflops"] / pk, 4),
                     "traffic": GEMM_TRAFFIC.get((args.prec, args.mv_chi)),
                     "kernel": f"matvec GEMM ({args.prec}), M=5*{args.mv_chi}, N=4*{args.mv_chi}, "
                               f"K={args.mv_chi}; {F_gemm:.3e} algorithmic flop/launch"},
        "tdvp_xxz_n100": tdvp,
        "ttn_random_field_batch": ttn,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
