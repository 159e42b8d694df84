"""suqa_amd -- MI355X-native (gfx950) SUQA state-vector engine and QMS step.

Product layout:
  include/suqa_hip.h        C-ABI (the drop-in boundary)
  suqa_amd/csrc/            HIP kernels + C-ABI implementation
  suqa_amd/host/            C++ mirror of the reference's suqa:: / qms:: API
  suqa_amd/capi.py          ctypes binding used by tests and bench.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
