"""protlocus_amd: MI355X-native spliced protein-to-DNA DP core.

The product is libnasw_hip.so (gfx950 HIP kernels behind the reference's ns_* C-ABI,
declared in include/nasw-hip.h). This package holds its sources (csrc/), the in-tree build
(build.py) and the seeded synthetic workloads used by the benchmark (synth.py).
"""
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnasw_hip.so")
