"""Host-side logic that needs no GPU: the algorithmic flop model behind the
roofline figures, and the grid construction of the density post-processing."""
import math

import numpy as np


def test_flop_model_matches_the_published_class_table():
    """F_prim and F_digest per class (the table the roofline fractions are
    quoted against)."""
    from dft_amd.flops import f_digest, f_prim
    table = {(0, 0, 0, 0): (48, 12), (1, 0, 0, 0): (79, 36), (1, 1, 0, 0): (180, 108),
             (1, 0, 1, 0): (178, 108), (1, 1, 1, 0): (491, 324), (1, 1, 1, 1): (1608, 972),
             (2, 0, 0, 0): (150, 72), (2, 1, 0, 0): (411, 216), (2, 2, 0, 0): (1006, 432),
             (2, 2, 1, 1): (9484, 3888), (2, 1, 2, 1): (9334, 3888), (2, 2, 2, 2): (51358, 15552)}
    for cls, (fp, fd) in table.items():
        assert f_prim(*cls) == fp, cls
        assert f_digest(*cls) == fd, cls


def test_build_flops_columns():
    """build_flops reads K-products (model) or primitive counts (evaluated,
    FP32 records in mixed mode, FP64 ones in FP64 mode) from the stats rows."""
    from dft_amd.flops import build_flops, combo_flops
    row = [10, 20, 300, 400, 500, 600, 0, 700]
    per = [row] + [[0] * 8 for _ in range(20)]
    assert build_flops(per) == (combo_flops(0, 10, 500), combo_flops(0, 20, 600))
    assert build_flops(per, model=False) == (combo_flops(0, 10, 300), combo_flops(0, 20, 400))
    assert build_flops(per, model=False, mode="fp64") == (combo_flops(0, 10, 300), combo_flops(0, 20, 700))


def test_box_grid_and_reference_box():
    """linspace axes of ceil(extent / h) + 1 points, x slowest, FP32; the
    reference's box pads y by -0.75 and z by 1.0 and ends x at max - 1.0
    (shelled_eval.py:27-57)."""
    from dft_amd.grid import box_grid, reference_grid
    g, shape = box_grid([0.0, 0.0, 0.0], [1.0, 0.5, 0.25], 0.1)
    assert shape == (11, 6, 4) and g.dtype == np.float32 and g.shape == (11 * 6 * 4, 3)
    assert np.allclose(g[0], [0, 0, 0]) and np.allclose(g[1], [0, 0, 0.25 / 3])
    assert np.allclose(g[-1], [1.0, 0.5, 0.25])
    c = np.array([[0.0, 0.0, 0.0], [3.0, 2.0, 1.0]])
    g, shape = reference_grid(c, spacing=0.05)
    lo = np.array([0.0, 0.75, -1.0])
    hi = np.array([2.0, 1.25, 2.0])
    assert shape == tuple(math.ceil(e / 0.05) + 1 for e in hi - lo)
    assert np.allclose(g[0], lo, atol=1e-6) and np.allclose(g[-1], hi, atol=1e-6)
