"""MI355X (gfx950) energy-and-force evaluation of Amber/CHARMM systems.

The evaluation runs in hand-written HIP kernels (md_amd/csrc/gpu) behind the
mdgpu C ABI (include/mdgpu.h); the C++ host (md_amd/csrc/host) keeps the
reference's AmberTopology/System API.  This package holds the Python binding
of both libraries, the synthetic workload generator and the differentiable
wrapper.
"""
from .native import (PREC_FP64, PREC_MIXED, SEM_AMBER, SEM_REF, Evaluator, GpuError, Topology,
                     desc_from_arrays, make_params)

__all__ = ["PREC_FP64", "PREC_MIXED", "SEM_AMBER", "SEM_REF", "Evaluator", "GpuError", "Topology",
           "desc_from_arrays", "make_params"]
