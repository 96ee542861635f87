"""giwaxsim_amd: MI355X (gfx950) forward-scattering path of GIWAXSim.

Drop-in modules mirroring the reference's call surface:
``giwaxsim_amd.tools.{voxelgrids, detector, comparison, utilities}`` and the
config-file entry point ``giwaxsim_amd.simulate_GIWAXS``.  All per-element
work runs in the hand-written HIP kernels of ``lib/libgiwaxsim_amd.so``.
"""

__version__ = "0.1.0"
