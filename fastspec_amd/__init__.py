"""fastspec_amd: MI355X-native line opacity and formal solution for FastSpec.

The compute path is libfsgpu.so (HIP, gfx950) behind the C-ABI in include/fsgpu.h.
"""
__version__ = "0.1.0"
