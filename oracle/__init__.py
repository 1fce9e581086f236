"""CPU oracle for loopy BP -- TEST INFRASTRUCTURE ONLY.

Importable only from tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, where it serves as the checker. ``lbp`` restates the
reference's BP iteration in NumPy (FP32 op order or FP64), ``brute``
enumerates tiny graphs exactly, and ``lbp_ref.c`` is a C/OpenMP
restatement used for the timed CPU baseline.
"""
