"""MI355X-native effective-Hamiltonian / transfer-matrix path for MPSKit.jl.

Host-side mirror of MPSKit.jl's derivative-operator, transfer-matrix and
environment interface over the C ABI of ``libmpskit_hip.so`` (HIP/CDNA4
kernels for gfx950).  Importing this package does not touch the GPU; the first
call that needs the library loads it and raises if it is missing.
"""
from . import models, tensors
from ._lib import BF16X3, FP32, FP64, EXPORTED, LIB_PATH, MKError, NotConvergedError, SpaceMismatch, load
from .derivatives import (
    BatchedDerivative, autotune_stats, DerivativeOperator, JordanMPO_AC2_Hamiltonian, JordanMPO_AC_Hamiltonian, MPO_AC2_Hamiltonian, gemm,
    JordanMPO_AC2_Terms, JordanMPO_AC_Terms,
    MPO_AC_Hamiltonian, MPO_C_Hamiltonian, MPODerivativeOperator, PreparedDerivative, prepare_operator,
)
from .transfer import (
    ProductTransferMatrix, RegTransferMatrix, SingleTransferMatrix, TransferMatrix, prepare_transfer_left,
    prepare_transfer_right, regularize, regularize_, transfer_left, transfer_right,
)

__all__ = [
    "models", "tensors", "BF16X3", "FP32", "FP64", "EXPORTED", "LIB_PATH", "MKError", "NotConvergedError",
    "SpaceMismatch", "load", "gemm", "BatchedDerivative", "autotune_stats", "DerivativeOperator", "JordanMPO_AC2_Hamiltonian", "JordanMPO_AC_Hamiltonian",
    "MPO_AC2_Hamiltonian", "MPO_AC_Hamiltonian", "MPO_C_Hamiltonian", "MPODerivativeOperator",
    "PreparedDerivative", "prepare_operator", "ProductTransferMatrix", "RegTransferMatrix",
    "JordanMPO_AC2_Terms", "JordanMPO_AC_Terms",
    "SingleTransferMatrix", "TransferMatrix", "prepare_transfer_left", "prepare_transfer_right", "regularize", "regularize_", "transfer_left", "transfer_right",
]
from .algorithms import (
    DMRG, DMRG2, AdaptiveKrylov, Davidson, Lanczos, expectation_value, find_groundstate, find_groundstate_batched,
)
from .learned import MLPPreconditioner, davidson
from .u1 import U1Bond, U1FiniteMPS, find_groundstate_u1, set_charge_blocks
from .environments import FiniteEnvironments, environments
from .states import FiniteMPOHamiltonian, FiniteMPS

__all__ += ["Davidson", "MLPPreconditioner", "davidson", "find_groundstate_batched", "U1Bond", "U1FiniteMPS",
            "find_groundstate_u1", "set_charge_blocks"]
__all__ += ["DMRG", "DMRG2", "AdaptiveKrylov", "Lanczos", "expectation_value", "find_groundstate", "FiniteEnvironments",
            "environments", "FiniteMPOHamiltonian", "FiniteMPS"]

# algorithm modules beyond DMRG2 / VUMPS, reachable as mpskit_amd.<module>
from . import approximate, blocks, boundary, changebonds, ed, excited, idmrg, observables, qp, tdvp, vumps  # noqa: E402,F401

