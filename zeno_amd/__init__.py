"""zeno_amd — an MI355X-native (gfx950 HIP) state-vector backend with the
capabilities of cleoanka/zeno's CPU path.

The compute lives in zeno_amd/libzeno_hip.so (HIP kernels + the C++ host
side, C ABI in include/zeno_hip.h).  Importing this package does not touch
the GPU; constructing a backend or running a circuit does, and fails loudly
if the library or a HIP device is missing.
"""
from .backend import (
    DeviceError,
    HipBackend,
    MemoryBudget,
    Unsupported,
    ZenoError,
    calibration_copy,
    copy_probe,
    device_count,
    device_memory,
)
from .circuit import Circuit, matching_circuit, qft, random_circuit, u3_qft
from .exec import (
    Compiled,
    Counts,
    InvalidCircuit,
    NoiseModel,
    RunOptions,
    RunResult,
    Simulator,
    batch_supported,
    budget_bytes,
    compile,
    format_key,
    max_qubits,
    plan,
    resolve_precision,
    run_batch,
    run_compiled,
    run_program,
    run_trajectories,
    state_bytes,
    trajectories_supported,
)
from . import ir
from .serve import Served, Server
from . import qasm, tune

__all__ = [
    "Circuit", "Compiled", "calibration_copy", "copy_probe", "device_count", "device_memory", "Counts", "DeviceError", "HipBackend", "InvalidCircuit", "MemoryBudget",
    "NoiseModel", "RunOptions", "RunResult", "Simulator", "Unsupported", "ZenoError", "compile",
    "format_key", "ir", "qasm", "tune", "matching_circuit", "max_qubits", "plan", "budget_bytes", "resolve_precision", "batch_supported", "run_batch", "qft", "random_circuit",
    "Served", "Server", "run_compiled", "run_program", "run_trajectories", "state_bytes", "trajectories_supported", "u3_qft",
]
