"""tennetlib_amd -- TenNetLib.jl's effective-Hamiltonian hot path on MI355X.

Host-side mirror of the reference's MPS API (StateEnvs, position!, product,
DMRGParams, dmrg!, fullsweep!, TDVPEngine, tdvpsweep!) and TTN API
(StateEnvsTTN, OptimizeParamsTTN, optimize!) over the C ABI of libtnhip.so.  Julia ``f!`` names are spelled ``f_``.
"""

from . import core as _core
from . import ttn as _ttn
from ._lib import (TNHIPError, TN_PREC_BF16X3, TN_PREC_FP32, TN_PREC_FP64, load)  # noqa: F401
from .core import (Context, DeviceArray, DMRGParams, StateEnvs, SweepData, TDVPEngine,  # noqa: F401
                   default_context, dmrg, dmrg1, dmrg2, dmrg_, dynamic_fullsweep_, eig_solver, exp_solver,
                   krylov_extend_, make_dmrg_params, nsite, orthogonalize_, set_nsite_, tdvpsweep_)
from .couplingmodel import coupling_mpo, mpo_sum_ids  # noqa: F401
from .batch import optimize_instances  # noqa: F401
from .ttn import (TTN, CouplingModel, OpStrings, OptimizeParamsTTN, StateEnvsTTN, SweepDataTTN,  # noqa: F401
                  default_graph_sitenodes, default_randomTTN, default_sweeppath, find_eccentric_central_node,
                  make_optimize_params, maxlinkdim, nextnode_in_path, optimize, optimize_, randomTTN,
                  shortest_path, subspace_expand_)


def _dispatch(mps_fn, ttn_fn):
    def f(sysenv, *args, **kwargs):
        return (ttn_fn if isinstance(sysenv, _ttn.StateEnvsTTN) else mps_fn)(sysenv, *args, **kwargs)
    f.__name__ = mps_fn.__name__
    f.__doc__ = f"{mps_fn.__doc__}\n\nOn a StateEnvsTTN: {ttn_fn.__doc__}"
    return f


# Julia dispatches these on StateEnvs / StateEnvsTTN; here on the Python type
position_ = _dispatch(_core.position_, _ttn.position_)
product = _dispatch(_core.product, _ttn.product)
update_position_ = _dispatch(_core.update_position_, _ttn.update_position_)
fullsweep_ = _dispatch(_core.fullsweep_, _ttn.fullsweep_)
getpsi = _dispatch(_core.getpsi, _ttn.getpsi)
from .models import (cylinder_couplings, heisenberg_mpo, hubbard_mpo, hubbard_occupation_state,  # noqa: F401
                     hubbard_qns, j1j2_cylinder_mpo, neel_state, product_state, spin_coupling_mpo, spin_half_qns,
                     xxz_mpo)
