"""Cell matrices shared by the CPU and GPU parity tests.

FAMILY_CELLS is the reference's own cpu-vs-ref parity matrix
(reference tests/test_dp_backend.py:21-38); TRANSFORM_CELLS is its GPU
transform matrix (reference tests/test_gpu_fast_paths.py:181-187).
"""
FAMILY_CELLS = [
    ("hamming", [2] * 8, 1),
    ("hamming", [3] * 5, 2),
    ("hamming", [4] * 4, 1),
    ("hamming", [2, 3, 4, 5], 2),
    ("hamming", [6, 6], 2),
    ("lee", [5, 5], 1),
    ("lee", [7, 7], 2),
    ("lee", [4, 5, 6], 3),
    ("lee", [8, 8], 4),
    ("torus_linf", [9, 9], 1),
    ("torus_linf", [7, 7, 7], 1),
    ("torus_linf", [11, 5], 2),
    ("grid_linf", [8, 8], 1),
    ("grid_linf", [5, 6, 4], 2),
    ("grid_linf", [7, 7], 5),
]
# the reference matrix also holds ("hamming", [5, 2, 3], 0): R = 0 is
# rejected by the plugin ABI (dct_init3 needs R >= 1), so it is a CPU-only
# oracle case here
ORACLE_ONLY_CELLS = [("hamming", [5, 2, 3], 0)]

FAMILY_IDS = [f"{p}-{'x'.join(map(str, a))}-R{r}" for p, a, r in FAMILY_CELLS]

TRANSFORM_CELLS = [
    ([2] * 9, 3, 1), ([2] * 10, 2, 1), ([3] * 7, 2, 1), ([3] * 2, 1, 1),
    ([4] * 5, 4, 1), ([5] * 5, 3, 1), ([4, 4, 5, 5, 3], 2, 1),
    ([6] * 4, 2, 1), ([7] * 4, 3, 1), ([8] * 4, 4, 1), ([9] * 4, 3, 1),
    ([10] * 4, 4, 1), ([3] * 4, 4, 1), ([7] * 3, 3, 1),
    ([2, 3, 4, 5, 6, 7], 3, 1), ([10, 9, 8, 7, 6], 5, 1), ([3] * 6, 2, 2),
    ([4] * 4, 1, 3),
]
TRANSFORM_IDS = [f"{'x'.join(map(str, a))}-R{r}-mu{m}" for a, r, m in TRANSFORM_CELLS]
