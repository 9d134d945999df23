"""mcstreamer_amd: MI355X (gfx950) implementation of the per-step particle-mesh
loop of aluque/mcstreamer behind the reference's JuMC / Multigrid / Swarm API.

The compute lives in libmcs.so (HIP kernels + C ABI, include/mcs.h); the
modules here mirror the reference's Julia functions and call it through ctypes.
"""
__all__ = ["jumc", "multigrid", "streamer", "swarm", "lxcat", "gases", "constants"]
