"""CPU oracle for loopy BP -- TEST INFRASTRUCTURE ONLY.

Importable only from tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, where it serves as the checker. ``lbp`` restates the
reference's BP iteration in NumPy (FP32 op order or FP64), ``brute``
enumerates tiny graphs exactly, ``lbp_cpu.c`` (driven by ``cpu``) is a
C/OpenMP restatement of the Enum iteration used for the timed CPU baseline,
and ``mpnn`` is the FP64 oracle of the warm-start network.
"""
