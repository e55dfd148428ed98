"""coldcore_amd -- MI355X-native exact-transform engine for covering problems.

The package is the drop-in ``gpu`` backend of the coldcore search core for
its one data-parallel hot path: the separable per-axis DP transforms (exact
gain, loss and coverage-count fields over a mixed-radix product space), the
incremental add/remove updates and the per-candidate exact queries.

* ``csrc/``     HIP kernels for gfx950 and the C ABI (``dct_*`` symbols)
* ``gpu``       ctypes binding, ``GpuBackend`` (reference src/coldcore/gpu.py)
* ``patterns``  v1 / ox support-table builders (reference patterns.py)
* ``space``     mixed-radix indexing (reference space.py)
* ``build``     the hipcc recipe for ``libcoldcore_covering.so``
"""
from .patterns import PROBLEM_IDS, ball_size, lee_ball_size, support_patterns
from .space import ProductSpace

__all__ = ["PROBLEM_IDS", "ProductSpace", "ball_size", "lee_ball_size",
           "support_patterns"]
