"""The oracle's gridded-visibility mode (gvm_oracle_grid.c) pinned by closed
forms: cell placement of a row and its Hermitian twin, effective weights of
the convolution, compaction order, off-grid rows, idempotence of gridding
a gridded set, agreement with the independent dense gridding comparator,
and the weighting schemes (Uniform sums, Briggs f^2 including the
cumulative first-pass grid, Radial distances, the UV taper)."""
import ctypes as C

import numpy as np
import pytest

import oracle

M = N = 32
DX = -0.05 / 3600.0  # degrees (CDELT1 < 0)
DY = 0.05 / 3600.0
FREQ = np.float32(230e9)
LAM = float(np.float32(2.99792458e8) / FREQ)
RAD = np.pi / 180.0
DU = 1.0 / (M * DX * RAD)  # negative, as the reference's deltau
DV = 1.0 / (N * DY * RAD)
CJ = CK = 16


def _rows(offsets, w, vo):
    """Rows whose (u, v) sit at integer cell offsets (a, b) from the centre."""
    uvw = np.array([[a * DU * LAM, b * DV * LAM, 0.0] for a, b in offsets])
    return uvw, np.asarray(vo, np.float32).reshape(-1, 2), np.asarray(w, np.float32)


def _grid(uvw, vo, w, taps=np.ones((1, 1), np.float32), sup=None):
    taps = np.asarray(taps, np.float32)
    s = int(taps.shape[0] // 2) if sup is None else sup
    return oracle.grid_visibilities(M, N, DX, DY, FREQ, uvw, vo, w, taps, s, s)


def _cells(uvw):
    j = np.rint(uvw[:, 0] / LAM / DU).astype(int) + CJ
    k = np.rint(uvw[:, 1] / LAM / DV).astype(int) + CK
    return k, j


def test_single_row_and_twin_pillbox():
    uvw, vo, w = _rows([(2, -1)], [4.0], [1.5, -0.5])
    gu, gv, gw = _grid(uvw, vo, w)
    assert gu.shape == (2, 3)
    k, j = _cells(gu)
    # row-major order: the twin (row 17) comes after the row (row 15)
    assert list(zip(k, j)) == [(CK - 1, CJ + 2), (CK + 1, CJ - 2)]
    assert np.array_equal(gv, np.array([[1.5, -0.5], [1.5, 0.5]], np.float32))
    assert np.array_equal(gw, np.array([4.0, 4.0], np.float32))
    assert np.allclose(gu[0, :2], uvw[0, :2], rtol=1e-12) and np.all(gu[:, 2] == 0.0)
    assert np.allclose(gu[1, :2], -uvw[0, :2], rtol=1e-12)


def test_same_cell_weighted_mean_and_weight_sum():
    uvw, vo, w = _rows([(3, 4), (3, 4)], [1.0, 3.0], [[2.0, 0.0], [6.0, 4.0]])
    gu, gv, gw = _grid(uvw, vo, w)
    assert gu.shape[0] == 2
    # pillbox: weight (sum w)^2 / sum w = sum w; Vo the weighted mean
    assert np.allclose(gw, 4.0, rtol=1e-7)
    first = np.argmax(_cells(gu)[0] == CK + 4)
    assert np.allclose(gv[first], [5.0, 3.0], rtol=1e-6)


def test_kernel_taps_effective_weight():
    taps = np.array([[0.25, 0.5, 0.25], [0.5, 1.0, 0.5], [0.25, 0.5, 0.25]], np.float32)
    uvw, vo, w = _rows([(5, 5)], [2.0], [1.0, 1.0])
    gu, gv, gw = _grid(uvw, vo, w, taps)
    assert gu.shape[0] == 18  # 3x3 around the row and around its twin
    # a single row: (w k)^2 / (w k^2) = w and Vo unchanged, whatever the tap
    assert np.allclose(gw, 2.0, rtol=1e-6)
    assert np.allclose(np.abs(gv[:, 0]), 1.0, rtol=1e-6)
    # two overlapping rows one cell apart: cell between them has taps 0.5 and 0.5
    uvw, vo, w = _rows([(5, 5), (7, 5)], [2.0, 6.0], [[1.0, 0.0], [3.0, 0.0]])
    gu, gv, gw = _grid(uvw, vo, w, taps)
    k, j = _cells(gu)
    mid = np.flatnonzero((k == CK + 5) & (j == CJ + 6))[0]
    ws, aux = 2 * 0.5 + 6 * 0.5, 2 * 0.25 + 6 * 0.25
    assert np.isclose(gw[mid], ws * ws / aux, rtol=1e-6)
    assert np.isclose(gv[mid, 0], (2 * 0.5 * 1 + 6 * 0.5 * 3) / ws, rtol=1e-6)


def test_off_grid_zero_weight_and_edge_twin():
    # a row far off the grid, a zero-weight row, and a row at column 0 whose
    # twin would land at column N (off the grid)
    uvw, vo, w = _rows([(100, 0), (1, 1), (-CJ, 2)], [1.0, 0.0, 2.0],
                       [[1, 0], [1, 0], [1, 1]])
    gu, gv, gw = _grid(uvw, vo, w)
    k, j = _cells(gu)
    assert list(zip(k, j)) == [(CK + 2, 0)]
    assert np.array_equal(gw, np.array([2.0], np.float32))
    # an empty slice grids to nothing
    e = np.zeros((0, 3))
    gu, gv, gw = _grid(e, np.zeros((0, 2), np.float32), np.zeros(0, np.float32))
    assert gu.shape == (0, 3)


def test_gridding_a_gridded_set_is_idempotent():
    rng = np.random.default_rng(3)
    n = 400
    uvw = np.column_stack([rng.uniform(-12, 12, n) * DU * LAM, rng.uniform(-12, 12, n) * DV * LAM,
                           rng.uniform(-50, 50, n)])
    vo = rng.normal(size=(n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    gu, gv, gw = _grid(uvw, vo, w)
    hu, hv, hw = _grid(gu, gv, gw)
    # every cell maps to itself; a gridded set is Hermitian, so each cell
    # receives itself and its mirror's twin: same Vo, twice the weight
    assert np.array_equal(np.stack(_cells(gu)), np.stack(_cells(hu)))
    assert np.allclose(hv, gv, rtol=1e-6, atol=1e-7)
    assert np.allclose(hw, 2 * gw, rtol=1e-6)


def test_matches_dense_gridding_comparator():
    """The independent dense gridding comparator (OpenMP, 7x7 pillbox) gives
    the same cells and values."""
    rng = np.random.default_rng(4)
    n = 300
    uvw = np.column_stack([rng.uniform(-10, 10, n) * DU * LAM, rng.uniform(-10, 10, n) * DV * LAM,
                           np.zeros(n)])
    vo = rng.normal(size=(n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    gu, gv, gw = _grid(uvw, vo, w, np.ones((7, 7), np.float32))
    ch = (oracle.OChan * 1)(oracle.OChan(FREQ, n, uvw.ctypes.data, vo.ctypes.data, w.ctypes.data))
    re, im, dw = (np.zeros((N, M)) for _ in range(3))
    oracle.lib().orc_do_gridding_pillbox(M, N, DU, DV, ch, 1, 3, 1, re.ctypes.data,
                                         im.ctypes.data, dw.ctypes.data)
    k, j = _cells(gu)
    assert np.count_nonzero(dw) == gu.shape[0]
    assert np.allclose(dw[k, j], gw, rtol=1e-5)
    assert np.allclose(re[k, j], gv[:, 0], rtol=1e-5, atol=1e-6)
    assert np.allclose(im[k, j], gv[:, 1], rtol=1e-5, atol=1e-6)


def _wrows(offsets, w):
    uvw, _, w = _rows(offsets, w, np.zeros(2 * len(offsets)))
    return uvw, w


def _weigh(scheme, slices, **kw):
    return oracle.weighting(scheme, slices, M, N, DX, DY, **kw)


def test_uniform_weights_sum_to_one_per_cell():
    uvw, w = _wrows([(1, 1), (1, 1), (1, 1), (4, -2), (-3, 5), (100, 0)],
                    [1.0, 2.0, 5.0, 3.0, 0.5, 1.0])
    (out,) = _weigh("uniform", [(FREQ, uvw, w)])
    assert np.allclose(out[:3], np.array([1, 2, 5]) / 8.0, rtol=1e-7)
    assert np.allclose(out[3:5], 1.0)
    assert out[5] == 0.0  # off the grid
    # rows with u < 0 are mirrored: (-4, 2) shares the cell of (4, -2)
    uvw, w = _wrows([(4, -2), (-4, 2)], [1.0, 3.0])
    (out,) = _weigh("uniform", [(FREQ, uvw, w)])
    assert np.allclose(out, [0.25, 0.75], rtol=1e-7)


@pytest.mark.parametrize("robust", [-2.0, 0.0, 0.5, 2.0])
@pytest.mark.parametrize("float_totals", [True, False])
def test_briggs_closed_form(robust, float_totals):
    offs = [(a, b) for a in range(1, 6) for b in range(-3, 3)]
    uvw, w = _wrows(offs, np.ones(len(offs)))
    (out,) = _weigh("briggs", [(FREQ, uvw, w)], robust=robust, float_totals=float_totals)
    f2 = 25.0 * 10.0 ** (-2 * robust)  # S2 / S1 = 1: one row of weight 1 per cell
    assert np.allclose(out, 1.0 / (1.0 + f2), rtol=1e-6)


def test_briggs_first_pass_grid_is_cumulative_over_slices():
    offs = [(a, 2) for a in range(1, 9)]
    uvw, w = _wrows(offs, np.ones(len(offs)))
    out = _weigh("briggs", [(FREQ, uvw, w), (FREQ, uvw.copy(), w.copy())], robust=0.0)
    # pass 1: S1 = 2n, S2 = n (slice 0) + 4n (slice 1 sees both slices) = 5n
    # => average 2.5, f^2 = 25 / 2.5 = 10; pass 2 grids each slice alone: G = 1
    for o in out:
        assert np.allclose(o, 1.0 / 11.0, rtol=1e-6)


def test_radial_taper_and_natural():
    uvw, w = _wrows([(3, 4), (-6, 8), (0, 0)], [1.0, 2.0, 3.0])
    (out,) = _weigh("radial", [(FREQ, uvw, w)])
    r = np.hypot(uvw[:, 0] / LAM, uvw[:, 1] / LAM)
    assert np.allclose(out, w * r, rtol=1e-6)
    (out,) = _weigh("natural", [(FREQ, uvw, w)])
    assert np.array_equal(out, w)
    sigma = 3.0 * abs(DU)
    taper = dict(amplitude=1.0, u_0=0.0, v_0=0.0, sigma_maj=sigma, sigma_min=sigma, bpa=0.0)
    (out,) = _weigh("natural", [(FREQ, uvw, w)], taper=taper)
    assert np.allclose(out, w * np.exp(-r ** 2 / (2 * sigma ** 2)), rtol=1e-5)
    # an elliptical taper rotated by 90 degrees swaps the axes
    t2 = dict(taper, sigma_maj=2 * sigma, bpa=float(np.pi / 2))
    (out,) = _weigh("natural", [(FREQ, uvw, w)], taper=t2)
    u, v = uvw[:, 0] / LAM, uvw[:, 1] / LAM
    ref = w * np.exp(-u ** 2 / (2 * sigma ** 2) - v ** 2 / (2 * (2 * sigma) ** 2))
    assert np.allclose(out, ref, rtol=1e-4)
