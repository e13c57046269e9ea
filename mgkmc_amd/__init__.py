"""mgkmc_amd: MI355X-native (gfx950 HIP) equilibrium-solve hot path of mgkmc.

Drop-in replacements for the reference's solver seams (same names, argument
meaning and error behaviour), backed by libmgkmc_hip.so through ctypes.  There
is no CPU fallback: the HIP library and a ROCm GPU are required at call time.
"""
__version__ = "0.1.0"
