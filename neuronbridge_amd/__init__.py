"""MI355X (gfx950) simulation core for NeuronBridge's per-timestep hot path.

The compute lives in libnbx.so (hand-written HIP, C ABI in include/nbx_core.h);
this package is the host-side wrapper. Importing it does not touch the GPU;
the first call that needs libnbx.so raises ImportError if it was not built.
"""
from . import _lib
from ._lib import (EULER, RK4, LIF_VJ, LIF_EXP, IZH, IZH_CUR, LIF_EXP2, LIF_EXP3, POISSON, TRIGGER_RELAY,
                   ENGINE_KERNELS, ENGINE_GRAPH, ENGINE_PERSISTENT,
                   ENGINE_AUTO, NbxError, lib)
from .network import ModelSpec, Net, PopulationsSpec, Synapses, build_fixed_indegree

__all__ = ["PopulationsSpec", "EULER", "RK4", "LIF_VJ", "LIF_EXP", "IZH", "IZH_CUR", "LIF_EXP2", "LIF_EXP3", "POISSON",
           "TRIGGER_RELAY", "ENGINE_KERNELS", "ENGINE_GRAPH",
           "ENGINE_PERSISTENT", "ENGINE_AUTO", "NbxError", "lib", "ModelSpec", "Net", "Synapses",
           "build_fixed_indegree"]
