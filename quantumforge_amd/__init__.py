"""quantumforge_amd: MI355X-native (gfx950) state-vector engine with the API
of the QuantumForge crate's state-vector path.

Host-side circuit description (core), the HIP backend over the C ABI of
include/qf_hip.h (backend), gradients (gradients), workload builders
(workloads) and the torch integration (torch_ops, imported on demand).
"""
from ._native import QuantumError
from .core import (Circuit, DensityMatrix, Gate, NoiseChannel, ParameterizedCircuit, PauliSum, RotationAxis,
                   ShotResult, SingleQubitObservable, SparseOperator, StateVector, rotation_gate)

__all__ = [
    "QuantumError", "Circuit", "DensityMatrix", "Gate", "NoiseChannel", "ParameterizedCircuit", "PauliSum",
    "RotationAxis", "ShotResult", "SingleQubitObservable", "SparseOperator", "StateVector", "rotation_gate", "HipBackend",
    "HipDeviceState", "HipDeviceStateBatch", "HipDeviceDensityMatrix", "Plan",
]


def __getattr__(name):
    # The backend loads libqfhip.so; import it lazily so the pure-host parts
    # (circuit building, workloads) import without the library.
    if name in ("HipBackend", "HipDeviceState", "HipDeviceStateBatch", "HipDeviceDensityMatrix", "Plan"):
        from . import backend

        return getattr(backend, name)
    raise AttributeError(name)
