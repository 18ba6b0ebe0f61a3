"""MI355X-native Langevin hot path of LangevinDynamicsForPCP.

Under- and overdamped stepping of a particle over periodic single- and
double-barrier potentials with dichotomic barrier noise, plus first-passage,
transition-path and MSD statistics, as hand-written HIP kernels for gfx950
behind the C ABI in include/ldpcp.h.

  integrator.LangevinEquation   reference-semantics integrator (drop-in path)
  sweep.Sweep                   batched absorbing first-passage engine
  serve.SweepServer             continuous-batching request loop

The shared library is built in-tree (``python -m langevindynamicsforpcp_amd.build``);
nothing here falls back to a CPU path.
"""
from ._native import LdpcpError, load  # noqa: F401

__all__ = ["LdpcpError", "load"]
