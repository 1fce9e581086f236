"""MI355X-native loopy belief propagation with PGMax's interface.

Subpackages mirror the reference's (pgmax/__init__.py): ``vgroup``,
``factor``, ``fgroup``, ``fgraph``, ``infer``. ``engine`` drives the HIP
kernels in ``csrc/`` through the C ABI declared in ``include/pgx.h``.
"""

__version__ = "0.1.0"
