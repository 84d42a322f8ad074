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
from concurrent.futures import FIRST_EXCEPTION, ProcessPoolExecutor, wait
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


def _call(i, fn, args):
    return i, fn(*args)


def map_instances(fn, args_list: Sequence[tuple], workers: int = DEFAULT_WORKERS, size=None) -> List:
    """fn(*args) for every tuple of ``args_list``, over ``workers`` spawned
    processes (at most MAX_WORKERS; <= 1: in this process, in order).  ``fn``
    must be importable by name (a module-level function).  ``size(args)``, if
    given, orders the submissions largest first, so the slowest instances start
    early.  Results come back in the order of ``args_list``.  The first failure
    (an instance's exception, or BrokenProcessPool if a worker process died)
    is raised here once the pool has shut down; instances not yet started are
    cancelled."""
    args_list = [tuple(a) for a in args_list]
    if not args_list:
        return []
    if workers <= 1:
        return [fn(*a) for a in args_list]
    nw = min(int(workers), MAX_WORKERS, len(args_list))
    idx = list(range(len(args_list)))
    if size is not None:
        idx.sort(key=lambda i: -size(args_list[i]))
    out: List = [None] * len(args_list)
    with ProcessPoolExecutor(max_workers=nw, mp_context=mp.get_context("spawn"), initializer=_init_worker) as ex:
        futs = [ex.submit(_call, i, fn, args_list[i]) for i in idx]
        done, pending = wait(futs, return_when=FIRST_EXCEPTION)
        for f in pending:
            f.cancel()
        for f in futs:
            if f.done() and not f.cancelled() and f.exception() is not None:
                raise f.exception()
        for f in done:
            i, res = f.result()
            out[i] = res
    return out


def optimize_instances(specs: Sequence[Tuple[int, Sequence[float], int]], params: Dict,
                       workers: int = DEFAULT_WORKERS) -> List[Tuple[List[float], List[int]]]:
    """Ground states of independent random-field chains (ground_state per
    instance, map_instances over the workers, largest N first).

    ``specs``: (N, h, seed) per instance; ``params``: keyword arguments of
    make_optimize_params.  ``workers`` <= 1 runs them in this process on the
    default context.  Results come back in the order of ``specs``."""
    args = [(int(N), np.asarray(h, dtype=np.float64), int(seed), dict(params)) for N, h, seed in specs]
    return map_instances(ground_state, args, workers, size=lambda a: a[0])
