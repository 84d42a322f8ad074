"""Many independent TTN ground-state problems on one GPU.

The reference optimises one TTN at a time (``optimize!``,
src/ttn/optimize_ttn.jl:148-223) and threads only over the term ids inside one
contraction (src/ttn/linktensors.jl:85,195).  A batch of independent
instances, such as the random-field chains of the many-instance workload,
needs no communication between instances, so it runs instance-parallel here:
``workers`` processes share the GPU, each with its own library context (its
own stream, workspaces and pinned staging buffers), and each runs the
instances it is handed one after another.  Processes rather than threads,
because the library serialises its entry points within one process.

Each instance's result depends only on its inputs and seed (the TTN's initial
tensors and the subspace-expansion padding are drawn from per-instance
generators), not on which worker ran it.
"""

from __future__ import annotations

import multiprocessing as mp
from typing import Dict, List, Sequence, Tuple

import numpy as np

# processes on the one GPU, each holding a context.  Measured on one MI355X
# with 32 random-field chains (N 16..64, maxdim 128), one run per setting on
# separate machines: 3 workers 1.48 instances/s, 4 workers 2.29 and 3.49,
# 5 workers 1.86, 6 workers 2.35, 8 workers 1.42 and 1.53; one process 0.94.
DEFAULT_WORKERS = 4
MAX_WORKERS = 8


def random_field_chain(N: int, h: Sequence[float]):
    """H = sum_j S_j . S_{j+1} + sum_j h_j Sz_j on an open chain of N spins 1/2,
    as a CouplingModel (OpStrings/CouplingModel, src/base/couplingmodel.jl:153-241):
    the Heisenberg chain of test/test_TTN.jl:10-16 plus one site field per spin."""
    from .ttn import CouplingModel, OpStrings
    os_ = OpStrings()
    for j in range(1, N):
        os_ += (1.0, "Sz", j, "Sz", j + 1)
        os_ += (0.5, "S+", j, "S-", j + 1)
        os_ += (0.5, "S-", j, "S+", j + 1)
    for j in range(1, N + 1):
        os_ += (float(h[j - 1]), "Sz", j)
    return CouplingModel(os_, N)


def ground_state(N: int, h: Sequence[float], seed: int, params: Dict) -> Tuple[List[float], List[int]]:
    """One instance: default_randomTTN(N, 12, seed=seed) on the default binary
    tree, optimised by optimize_(…, seed=seed) with make_optimize_params(**params)
    over default_sweeppath.  Returns every sweep's energy and maximum bond
    dimension."""
    from .ttn import StateEnvsTTN, default_randomTTN, default_sweeppath, make_optimize_params, optimize_
    op = make_optimize_params(**params)  # host-side validation before any device work
    psi0 = default_randomTTN(N, 12, seed=seed)
    env = StateEnvsTTN(psi0, random_field_chain(N, h))
    try:
        sw = optimize_(env, op, default_sweeppath(psi0), outputlevel=0, seed=seed)
    finally:
        env.close()
    return [float(e) for e in sw.energy], [int(c) for c in sw.maxchi]


def _init_worker():
    """One BLAS thread per worker: the workers' host threads issue the GPU work,
    and several processes' BLAS pools on the same cores would compete with them."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return
    threadpool_limits(limits=1)


def _run(task):
    i, N, h, seed, params = task
    return i, ground_state(N, h, seed, params)


def optimize_instances(specs: Sequence[Tuple[int, Sequence[float], int]], params: Dict,
                       workers: int = DEFAULT_WORKERS) -> List[Tuple[List[float], List[int]]]:
    """Ground states of independent random-field chains.

    ``specs``: (N, h, seed) per instance; ``params``: keyword arguments of
    make_optimize_params.  ``workers`` <= 1 runs them in this process on the
    default context; otherwise a pool of ``workers`` spawned processes (at most
    MAX_WORKERS) takes the instances largest first, one at a time, as workers
    free up.  Results come back in the order of ``specs``.  A failure in any
    instance is raised here after the pool has been shut down."""
    tasks = [(i, int(N), np.asarray(h, dtype=np.float64), int(seed), dict(params))
             for i, (N, h, seed) in enumerate(specs)]
    if not tasks:
        return []
    if workers <= 1:
        return [ground_state(N, h, seed, p) for _, N, h, seed, p in tasks]
    nw = min(int(workers), MAX_WORKERS, len(tasks))
    order = sorted(tasks, key=lambda t: -t[1])  # largest N first: the slowest instances start early
    out: List = [None] * len(tasks)
    pool = mp.get_context("spawn").Pool(nw, initializer=_init_worker)
    try:
        for i, res in pool.imap_unordered(_run, order, chunksize=1):
            out[i] = res
        pool.close()
    except BaseException:
        pool.terminate()
        raise
    finally:
        pool.join()
    return out
