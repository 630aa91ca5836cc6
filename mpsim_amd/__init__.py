"""mpsim_amd: MI355X-native MPS gate application with mpsim's Python API.

    import mpsim_amd as mpsim
    mps = mpsim.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, maxsvals=1)
    mps.wavefunction()

Arithmetic runs in libmpsx.so (hand-written HIP kernels for gfx950, built by
`python -m mpsim_amd.build`); there is no CPU fallback.
"""
from .core import MPS, MPSOperation, CannotConvertToMPSOperation
from .gates import (
    Node,
    igate,
    xgate,
    ygate,
    zgate,
    hgate,
    rgate,
    cnot,
    cphase,
    swap,
    haar_random_unitary,
    is_unitary,
    is_hermitian,
    is_projector,
    computational_basis_state,
    computational_basis_projector,
)

__all__ = [
    "MPS", "MPSOperation", "CannotConvertToMPSOperation", "Node", "igate", "xgate", "ygate",
    "zgate", "hgate", "rgate", "cnot", "cphase", "swap", "haar_random_unitary", "is_unitary",
    "is_hermitian", "is_projector", "computational_basis_state", "computational_basis_projector",
]
