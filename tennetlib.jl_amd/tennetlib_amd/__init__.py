"""tennetlib_amd -- TenNetLib.jl's effective-Hamiltonian hot path on MI355X.

Host-side mirror of the reference's MPS API (StateEnvs, position!, product,
DMRGParams, dmrg!, fullsweep!, TDVPEngine, tdvpsweep!) over the C ABI of
libtnhip.so.  Julia ``f!`` names are spelled ``f_``.
"""

from ._lib import (TNHIPError, TN_PREC_BF16X3, TN_PREC_FP32, TN_PREC_FP64, load)  # noqa: F401
from .core import (Context, DeviceArray, DMRGParams, StateEnvs, SweepData, TDVPEngine,  # noqa: F401
                   default_context, dmrg, dmrg2, dmrg_, eig_solver, exp_solver, fullsweep_, getpsi,
                   make_dmrg_params, nsite, orthogonalize_, position_, product, set_nsite_, tdvpsweep_,
                   update_position_)
from .models import (cylinder_couplings, heisenberg_mpo, j1j2_cylinder_mpo, neel_state, product_state,  # noqa: F401
                     spin_coupling_mpo, xxz_mpo)
