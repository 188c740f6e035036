"""CPU restatement of the reference TLM step: the checker for the HIP path.

Test infrastructure only. tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this package; the product
package ``emsolver_amd`` never does.
"""
from .oracle_tlm import OracleTlm, OracleConfig, load_library, WF_CW_RAMPED, \
    WF_GAUSS_SINE, WF_GAUSS_PULSE, WF_GAUSS_DERIV  # noqa: F401
