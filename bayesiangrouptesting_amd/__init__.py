"""MI355X (gfx950) HIP provider for the BayesianGroupTesting lattice hot path.

The numerical work runs in hand-written HIP kernels behind the C ABI in
include/bgt_hip.h (libbgt_hip.so, built in-tree by bayesiangrouptesting_amd.build).
This package mirrors the reference's public API for that path.
"""
from .api import (  # noqa: F401
    DilutionParameters,
    DilutionMode,
    DilutionTable,
    Error,
    InvalidArgument,
    KernelSpec,
    Lattice,
    LatticeType,
    Precision,
    ProgressConfig,
    ProgressSnapshot,
    Provider,
    ReportConfig,
    Runtime,
    Selection,
    SelectorType,
    ShrinkPlan,
    ShrinkingLattice,
    SimulationConfig,
    SimulationMode,
    SimulationOptions,
    SimulationResult,
    SimulationRunStatus,
    SimulationTimings,
    TerminationReason,
    TreeStats,
    TrueStatePolicy,
    Unsupported,
    available_providers,
    batch_round,
    clear_simulation_cancellation,
    exact_upset_mass,
    hip_available,
    lattice_footprint,
    make_shrink_plan,
    request_simulation_cancellation,
    reverse_bits,
    run_simulation,
    simulation_cancellation_reason,
    simulation_cancellation_requested,
)

__all__ = [
    "DilutionParameters", "DilutionMode", "DilutionTable", "Error", "InvalidArgument", "KernelSpec", "Lattice",
    "LatticeType", "Precision", "ProgressConfig", "ProgressSnapshot", "Provider", "ReportConfig", "Runtime",
    "Selection", "SelectorType", "ShrinkPlan", "ShrinkingLattice", "SimulationConfig", "SimulationMode",
    "SimulationOptions", "SimulationResult", "SimulationRunStatus", "SimulationTimings", "TerminationReason",
    "TreeStats", "TrueStatePolicy", "Unsupported", "available_providers", "batch_round",
    "clear_simulation_cancellation", "exact_upset_mass", "hip_available", "lattice_footprint", "make_shrink_plan",
    "request_simulation_cancellation", "reverse_bits", "run_simulation", "simulation_cancellation_reason",
    "simulation_cancellation_requested",
]
