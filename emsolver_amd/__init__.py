"""emsolver_amd - MI355X-native SCN TLM time step (HIP, gfx950).

The product is the C-ABI library ``lib/libemsolver_tlm.so`` (declared in
``include/GpuTlmHip.h``) plus the C++ drop-in ``RunGpuTlm`` glue under
``host/``. This package is the Python handle used by tests and bench.py.
"""
__version__ = "0.1.0"
