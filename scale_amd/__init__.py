"""scale_amd -- MI355X (gfx950) HIP implementation of SCALE-RM's HEVI short
time step (ATMOS_DYN_Tstep_short_fvm_hevi with UD3/CD4 flux stencils, the
vertical implicit solve and the numerical filter) and of FCT tracer advection
(ATMOS_DYN_Tstep_tracer_fvm_heve / ATMOS_DYN_fct), behind the reference's
ATMOS_DYN_Tstep_short / ATMOS_DYN_Tstep_tracer interfaces.

The compute path is libscalehip.so (scale_amd/csrc, built in-tree with
scale_amd/build.py); ``scale_amd.hip`` is its Python front end and
``scale_amd.inputs`` builds the synthetic workloads.
"""
from . import inputs  # noqa: F401

__all__ = ["inputs", "hip"]
