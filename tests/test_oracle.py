"""Pins the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails it: library routines
(numpy FFT), brute force, closed forms and invariants.  No reference code runs
here; the reference holds no golden vectors for this path.
"""
import numpy as np
import pytest

import oracle
from gpuvmem_amd.problem import small_problem

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- FFT (a4)
@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (64, 128), (256, 256)])
def test_fft_matches_numpy(shape):
    a = RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)
    f = oracle.fft2d(a, +1)
    ref = np.fft.ifft2(a) * a.size  # e^{+2 pi i}, unnormalised
    assert np.abs(f - ref).max() <= 1e-12 * np.abs(ref).max()
    b = oracle.fft2d(a, -1)
    assert np.abs(b - np.fft.fft2(a)).max() <= 1e-12 * np.abs(b).max()


def test_fft_brute_force_and_closed_forms():
    n = 8
    a = RNG.standard_normal((n, n)) + 1j * RNG.standard_normal((n, n))
    k = np.arange(n)
    W = np.exp(2j * np.pi * np.outer(k, k) / n)
    brute = W @ a @ W.T
    assert np.allclose(oracle.fft2d(a, +1), brute, atol=1e-12)
    # delta at the origin -> constant; shifted delta -> pure phase ramp
    d = np.zeros((16, 16), complex)
    d[0, 0] = 3.0
    assert np.allclose(oracle.fft2d(d), 3.0)
    d[:] = 0
    d[2, 5] = 1.0
    kk, ll = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    assert np.allclose(oracle.fft2d(d), np.exp(2j * np.pi * (2 * kk + 5 * ll) / 16))


def test_fft_rejects_non_power_of_two():
    with pytest.raises(ValueError):
        oracle.fft2d(np.zeros((12, 16)))


# ------------------------------------------------------- forward model (a1-a8)
@pytest.fixture(scope="module")
def prob():
    return small_problem(npix=64, n_antennas=8, seed=3)


@pytest.fixture(scope="module")
def orc(prob):
    return oracle.Oracle(prob, separable=False)


def test_point_source_at_phase_centre_is_constant_grid(prob, orc):
    """A point at the (integer) phase centre gives V = flux PB fg everywhere:
    the phase rotation cancels the shift of the transform exactly."""
    img = np.zeros((2, prob.N, prob.M))
    img[0] = prob.minpix
    img[0, int(prob.y0), int(prob.x0)] = 1000.0
    grid = orc.model_grid(img)
    pb = orc.pb()
    floor = -prob.eta * prob.minpix
    ref_img = np.maximum(img[0], floor) * pb * prob.fg_scale
    # direct sum of the image with the phase rotation applied
    ref = np.fft.ifft2(ref_img) * ref_img.size
    js = np.fft.fftfreq(prob.M, 1.0 / prob.M)
    rot = np.exp(-2j * np.pi * (prob.x0 / prob.M * js[None, :] + prob.y0 / prob.N * js[:, None]))
    assert np.abs(grid - ref * rot).max() < 1e-9 * np.abs(ref).max()
    # only the point (the floor is tiny in comparison): nearly constant
    assert np.abs(grid - grid[0, 0]).max() < 1e-2 * abs(grid[0, 0])


def test_calculate_inu_floor_and_spectral_index(prob):
    """calculateInu: I_nu = I0 (nu/nu0)^alpha floored at -eta MINPIX."""
    o = oracle.Oracle(prob)
    img = np.zeros((2, prob.N, prob.M))
    img[0] = -5.0  # below the floor everywhere
    grid = o.model_grid(img)
    pb = o.pb()
    floor = -prob.eta * prob.minpix
    ref = np.fft.ifft2(floor * pb * prob.fg_scale) * prob.M * prob.N
    js = np.fft.fftfreq(prob.M, 1.0 / prob.M)
    rot = np.exp(-2j * np.pi * (prob.x0 / prob.M * js[None, :] + prob.y0 / prob.N * js[:, None]))
    assert np.allclose(grid, ref * rot, atol=1e-9 * np.abs(ref).max())


def _grid_problem(prob, uvw):
    import copy

    q = copy.copy(prob)
    q.uvw = np.ascontiguousarray(uvw, np.float64)
    q.vo = np.zeros((uvw.shape[0], 2), np.float32)
    q.weight = np.ones(uvw.shape[0], np.float32)
    return q


def test_bilinear_degrid_on_grid_points_and_linear_grids(prob):
    """vis_mod: samples on cell corners return the cell; a grid linear in the
    cell indices is interpolated exactly; off-grid samples get weight 0."""
    import ctypes as C

    N = prob.N
    dxr, dyr = prob.deltax * np.pi / 180, prob.deltay * np.pi / 180
    du, dv = 1.0 / (N * dxr), 1.0 / (N * dyr)
    cells = [(3, 5), (10, 2), (0, 0), (N - 1, N - 2)]
    fr = [(0.0, 0.0), (0.25, 0.5), (0.75, 0.125), (0.5, 0.5)]
    uvw = []
    for (i, j), (a, b) in zip(cells, fr):
        # negative u (after the flip) wraps: i + a = u/du + N
        uvw.append(((i + a - N) * du, (j + b) * dv, 0.0))
    uvw.append((-2.5 * N * du, 0.0, 0.0))  # off the grid
    q = _grid_problem(prob, np.array(uvw))
    o = oracle.Oracle(q)
    jj, ii = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")  # [row j (v)][col i (u)]
    grid = (2.0 + 3.0 * ii + 0.5 * jj) + 1j * (1.0 - 0.25 * ii + 2.0 * jj)
    grid = np.ascontiguousarray(grid, np.complex128)
    vm = np.zeros(len(uvw), np.complex128)
    vr = np.zeros(len(uvw), np.complex128)
    L = oracle.lib()
    L.orc_degrid_chi2(C.byref(o.geom), grid.ctypes.data, o.chan, vm.ctypes.data, vr.ctypes.data)
    for k, ((i, j), (a, b)) in enumerate(zip(cells, fr)):
        if i + 1 < N and j + 1 < N:
            expect = (2.0 + 3.0 * (i + a) + 0.5 * (j + b)) + 1j * (1.0 - 0.25 * (i + a) + 2.0 * (j + b))
            assert abs(vm[k] - expect) < 1e-12, k
    # corner case with wrap: (N-1, N-2) + (0.5, 0.5) mixes column N-1 and 0
    i, j = cells[3]
    exp = 0.25 * (grid[j, i] + grid[j + 1, i] + grid[j, 0] + grid[j + 1, 0])
    assert abs(vm[3] - exp) < 1e-12
    assert o.weight[-1] == 0.0 and vm[-1] == 0.0
    assert np.all(o.weight[:-1] == 1.0)


def test_chi2_zero_on_model_data_and_hand_case(prob):
    """Noiseless data made by the same forward model give chi2 = 0; a single
    perturbed visibility gives 0.5 w |delta|^2."""
    o = oracle.Oracle(prob)
    img = prob.initial_image().astype(np.float64)
    img[0] += RNG.random((prob.N, prob.M)) * 20
    _, vr, clipped = o.chi2(img)
    vm = prob.vo[:, 0] + 1j * prob.vo[:, 1] - vr
    import copy

    q = copy.copy(prob)
    q.vo = np.ascontiguousarray(np.stack([vm.real, vm.imag], 1), np.float32)
    o2 = oracle.Oracle(q)
    c0, _, _ = o2.chi2(clipped)
    # float32 rounding of Vo is the only residual
    assert c0 < 1e-10 * np.sum(prob.weight * np.abs(vm) ** 2)
    q.vo = q.vo.copy()
    q.vo[7, 0] += np.float32(0.5)
    o3 = oracle.Oracle(q)
    c1, _, _ = o3.chi2(clipped)
    assert abs(c1 - c0 - 0.5 * prob.weight[7] * 0.25) < 1e-6 * c1


# --------------------------------------------------------------- adjoint (a10)
def _dft_model(prob, orc, img0):
    """Direct-DFT forward model with the w term (what DChi2 is the adjoint of)."""
    N = prob.N
    dxr, dyr = prob.deltax * np.pi / 180, prob.deltay * np.pi / 180
    x = (np.arange(N) - int(prob.x0)) * dxr
    y = (np.arange(N) - int(prob.y0)) * dyr
    X, Y = np.meshgrid(x, y)
    Z = np.sqrt(1 - X ** 2 - Y ** 2) - 1
    pb = orc.pb()
    u, v, w = prob.uvw[:, 0], prob.uvw[:, 1], prob.uvw[:, 2]
    ph = np.exp(2j * np.pi * (np.multiply.outer(u, X) + np.multiply.outer(v, Y) + np.multiply.outer(w, Z)))

    def model(im):
        return np.tensordot(ph, im * pb * prob.fg_scale, axes=([1, 2], [0, 1]))

    return model


def test_dchi2_is_gradient_of_direct_dft_chi2():
    """DChi2 = d/dI of 0.5 sum w |Vo - V_dft(I)|^2 with the direct-DFT model
    (including its w term), checked by central differences on a tiny case."""
    p = small_problem(npix=16, n_antennas=5, seed=9, cell_arcsec=0.2, dump_seconds=600.0)
    o = oracle.Oracle(p, separable=False)
    model = _dft_model(p, o, None)
    vo = p.vo[:, 0] + 1j * p.vo[:, 1]
    w = p.weight.astype(np.float64)
    img = RNG.random((p.N, p.M)) * 1e5

    def chi2(im):
        r = vo - model(im)
        return 0.5 * np.sum(w * np.abs(r) ** 2)

    vr = vo - model(img)
    g = o.dchi2_channel(vr, separable=False)
    mask = p.noise < p.noise_cut
    h = 1.0
    for (i, j) in [(3, 4), (8, 8), (12, 2), (0, 15)]:
        e = np.zeros_like(img)
        e[i, j] = h
        fd = (chi2(img + e) - chi2(img - e)) / (2 * h)
        if mask[i, j]:
            assert abs(g[i, j] - fd) < 1e-6 * abs(fd) + 1e-9, (i, j, g[i, j], fd)
        else:
            assert g[i, j] == 0.0


def test_dchi2_separable_matches_direct(prob, orc):
    img = prob.initial_image().astype(np.float64)
    img[0] += RNG.random((prob.N, prob.M)) * 50
    _, vr, clipped = orc.chi2(img)
    d = orc.dchi2_channel(vr, separable=False)
    s = orc.dchi2_channel(vr, separable=True)
    assert np.linalg.norm(s - d) <= 1e-9 * np.linalg.norm(d)
    pix = [(5, 7), (31, 32), (60, 3)]
    dp = orc.dchi2_pixels(vr, pix)
    assert np.allclose(dp, [d[i, j] for i, j in pix], rtol=1e-12)


def test_dchi2_chain_rule_flags(prob, orc):
    """flag_opt even -> image 0 gets dchi2 (nu = nu0: factor 1), image 1 zero;
    odd -> image 0 zero and image 1 gets dchi2 * I0 (nu/nu0)^a fg ln(nu/nu0) = 0
    at nu = nu0."""
    img = prob.initial_image().astype(np.float64)
    img[0] += RNG.random((prob.N, prob.M)) * 50
    _, vr, clipped = orc.chi2(img)
    d = orc.dchi2_channel(vr)
    r0 = orc.dchi2(clipped, vr, flag_opt=0)
    assert np.allclose(r0[0], d) and np.all(r0[1] == 0)
    r1 = orc.dchi2(clipped, vr, flag_opt=1)
    assert np.all(r1[0] == 0) and np.allclose(r1[1], 0.0)


def test_dchi2_two_channel_chain_rule_at_sampled_pixels():
    """Two channels at different frequencies: the sampled-pixel gradient (and
    the whole-image one) equal the chain rule written out here from each
    channel's own dchi2: sum_c d_c (nu_c/nu0)^alpha for I_nu_0 (even flag);
    sum_c d_c I0 (nu_c/nu0)^alpha fg ln(nu_c/nu0) where I0 > threshold for
    alpha (odd flag), src/functions.cu:3968-4024."""
    p = small_problem(npix=32, n_antennas=5, seed=3)
    q = small_problem(npix=32, n_antennas=5, seed=3, freq_hz=245e9, seed_noise=77)
    slices = [(p.freq, p.uvw, p.vo, p.weight), (q.freq, q.uvw, q.vo, q.weight)]
    p.threshold = 20.0
    p.noise = p.noise.copy()
    p.noise[0, :2] = 2 * p.noise_cut  # two pixels outside the mask
    orc = oracle.Oracle(p, slices=slices)
    rng = np.random.default_rng(8)
    img = p.initial_image().astype(np.float64)
    img[0] += rng.random((p.N, p.M)) * 40
    img[1] = rng.standard_normal((p.N, p.M)) * 0.7
    inside = np.argwhere(p.noise < p.noise_cut)
    pix = inside[rng.choice(len(inside), 12, replace=False)]
    pix = np.vstack([pix, np.argwhere(p.noise >= p.noise_cut)[:2]])
    for flag in (0, 1, 2, 3):
        _, vr, clipped = orc.chi2(img, flag_opt=flag)
        whole = orc.dchi2(clipped, vr, flag_opt=flag, separable=False)
        got = orc.dchi2_pixels_all(clipped, vr, pix, flag_opt=flag)
        hand = np.zeros((len(pix), 2))
        off = 0
        for f, u, v, w in slices:
            one = oracle.Oracle(p, slices=[(f, u, v, w)])
            n = u.shape[0]
            d = one.dchi2_pixels(vr[off:off + n], pix)
            off += n
            r = f / p.nu_0
            for k, (i, j) in enumerate(pix):
                if not p.noise[i, j] < p.noise_cut:
                    continue
                i0, a = clipped[0, i, j], clipped[1, i, j]
                if flag % 2 == 0:
                    hand[k, 0] += d[k] * r ** a
                elif i0 > p.threshold:
                    hand[k, 1] += d[k] * i0 * r ** a * p.fg_scale * np.log(r)
        ref = np.array([[whole[0, i, j], whole[1, i, j]] for i, j in pix])
        assert np.allclose(got, ref, rtol=1e-12, atol=0.0), flag
        assert np.allclose(got, hand, rtol=1e-12, atol=0.0), flag
        k = flag % 2
        assert np.all(got[:, 1 - k] == 0) and np.any(got[:, k] != 0)
        assert np.all(got[-2:] == 0)  # outside the mask


# ---------------------------------------------------------------- priors (a12-a14)
def _const(prob, c):
    img = np.zeros((2, prob.N, prob.M))
    img[0] = c
    return img


def test_entropy_closed_form_and_fd(prob, orc):
    G, c = 0.001, 0.37
    mask = prob.noise < prob.noise_cut
    img = _const(prob, c)
    assert np.isclose(orc.prior_value("entropy", img, prior=G), mask.sum() * c * np.log(c / G), rtol=1e-12)
    g = orc.prior_grad("entropy", img, lam=2.0, prior=G)
    assert np.allclose(g[mask], 2.0 * (np.log(c / G) + 1.0)) and np.all(g[~mask] == 0)
    # eta != -1: S = c ln(c/G + eta + 1), FD of a random positive image
    img[0] = 0.1 + RNG.random((prob.N, prob.M))
    for eta in (-1.0, 0.5):
        f0 = orc.prior_value("entropy", img, prior=G, eta=eta)
        g = orc.prior_grad("entropy", img, prior=G, eta=eta)
        e = np.zeros_like(img)
        e[0, 10, 11] = 1e-5
        fd = (orc.prior_value("entropy", img + e, prior=G, eta=eta) - orc.prior_value("entropy", img - e, prior=G, eta=eta)) / 2e-5
        assert abs(g[10, 11] - fd) < 1e-6 * abs(fd), (eta, f0)


def test_tsv_closed_forms(prob, orc):
    N = prob.N
    mask = prob.noise < prob.noise_cut
    c = 2.5
    img = _const(prob, c)
    border = np.zeros((N, N), bool)
    border[N - 1, :] = True
    border[:, N - 1] = True
    assert np.isclose(orc.prior_value("tsv", img), c * (mask & border).sum())
    g = orc.prior_grad("tsv", img, lam=3.0)
    inner = np.zeros((N, N), bool)
    inner[1:-1, 1:-1] = True
    assert np.allclose(g[inner & mask], 0.0) and np.allclose(g[~inner & mask], 3.0 * c)
    # ramp a*j: interior value a^2 per pixel, interior gradient 0
    a = 0.3
    img[0] = a * np.arange(N)[None, :] * np.ones((N, 1))
    v = orc.prior_value("tsv", img)
    exp = (a * a) * (mask & ~border).sum() + img[0][mask & border].sum()
    assert np.isclose(v, exp)
    # interior gradient = true derivative (central differences)
    img[0] = RNG.random((N, N))
    g = orc.prior_grad("tsv", img)
    for (i, j) in [(5, 6), (20, 30), (40, 41)]:
        e = np.zeros_like(img)
        e[0, i, j] = 1e-4
        fd = (orc.prior_value("tsv", img + e) - orc.prior_value("tsv", img - e)) / 2e-4
        assert abs(g[i, j] - fd) < 1e-6, (i, j)


def test_laplacian_stencil_coefficients(prob, orc):
    """The reference's DL stencil, coefficient by coefficient (delta image),
    including its sign pattern -8 (d - r - u - l), and the constant image."""
    N = prob.N
    img = np.zeros((2, N, N))
    i0, j0 = 20, 30
    img[0, i0, j0] = 1.0
    g = orc.prior_grad("laplacian", img)
    # pixel p's stencil reads the delta through one neighbour:
    assert g[i0, j0] == 20.0
    assert g[i0 - 1, j0] == -8.0  # delta is the "d" neighbour of the pixel above
    assert g[i0 + 1, j0] == 8.0   # "u" of the pixel below
    assert g[i0, j0 - 1] == 8.0   # "r" of the pixel to the left
    assert g[i0, j0 + 1] == 8.0   # "l" of the pixel to the right
    for di, dj in [(-1, -1), (-1, 1), (1, -1), (1, 1)]:
        assert g[i0 + di, j0 + dj] == 2.0
    for di, dj in [(-2, 0), (2, 0), (0, -2), (0, 2)]:
        assert g[i0 + di, j0 + dj] == 1.0
    assert np.count_nonzero(g) == 13
    c = 1.5
    g = orc.prior_grad("laplacian", _const(prob, c), lam=2.0)
    inner = np.zeros((N, N), bool)
    inner[2:-2, 2:-2] = True
    assert np.allclose(g[inner], 2.0 * 48.0 * c) and np.all(g[~inner] == 0)
    # value: interior 0.5 (Dx + Dy)^2 = 0 for a constant, border pixels raw
    border = np.ones((N, N), bool)
    border[1:-1, 1:-1] = False
    assert np.isclose(orc.prior_value("laplacian", _const(prob, c)), c * border.sum())


# ------------------------------------------------------------ line search (a18)
def test_mnbrak_brent_parabola_and_quartic():
    f = lambda x: (x - 3.0) ** 2 + 1.0
    ax, bx, cx, fa, fb, fc = oracle.mnbrak(f, 0.0, 1.0)
    assert min(ax, cx) < 3.0 < max(ax, cx) and fb <= fa and fb <= fc
    xmin, fx = oracle.brent(f, ax, bx, cx, 1e-7)
    assert abs(xmin - 3.0) < 1e-6 and abs(fx - 1.0) < 1e-12
    q = lambda x: (x - 2.0) ** 4 + x
    ax, bx, cx, *_ = oracle.mnbrak(q, 0.0, 1.0)
    xmin, _ = oracle.brent(q, ax, bx, cx, 1e-9)
    assert abs(xmin - (2.0 - 0.25 ** (1.0 / 3.0))) < 1e-4
    # downhill from b to a: mnbrak swaps so that f(b) <= f(a)
    ax, bx, cx, fa, fb, fc = oracle.mnbrak(lambda x: (x + 5.0) ** 2, 0.0, 1.0)
    assert min(ax, cx) < -5.0 < max(ax, cx)


# ------------------------------------------------------------ optimisers (a16-a17)
def _quadratic(n, seed):
    r = np.random.default_rng(seed)
    Q = np.linalg.qr(r.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.linspace(1.0, 20.0, n)) @ Q.T
    b = r.standard_normal(n)
    return A, b


def test_cg_solves_spd_quadratic_in_n_steps():
    n = 12
    A, b = _quadratic(n, 0)
    pp = oracle.PyProblem(n, 1, lambda p: 0.5 * p @ A @ p - b @ p, lambda p, it: A @ p - b)
    x, res, traj = pp.optimize(np.zeros(n), "CG-FRPRMN", iterations=n + 2)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-6
    assert res.iterations <= n + 1
    assert np.all(np.diff(traj) <= 1e-12)


def test_lbfgs_reference_indexing_is_memory_one_for_image0():
    """With a second image whose gradient is zero (alpha at flag_opt 0), the
    reference's storage addressing makes image 0 a memory-1 method: the
    trajectory for memory 10 equals the one for memory 1 and differs from a
    standard (properly indexed) memory-10 L-BFGS."""
    n = 10
    A, b = _quadratic(n, 1)

    def val(p):
        return 0.5 * p[:n] @ A @ p[:n] - b @ p[:n]

    def grad(p, it):
        return np.concatenate([A @ p[:n] - b, np.zeros(n)])

    pp = oracle.PyProblem(n, 2, val, grad)
    x0 = np.zeros(2 * n)
    _, _, t10 = pp.optimize(x0, "CG-LBFGS", iterations=8, lbfgs_k=10)
    _, _, t1 = pp.optimize(x0, "CG-LBFGS", iterations=8, lbfgs_k=1)
    assert np.allclose(t10, t1, rtol=0, atol=1e-12)
    # single image (no addressing effect), memory 1: y = g_new + previous direction
    pp1 = oracle.PyProblem(n, 1, val, lambda p, it: A @ p - b)
    x, res, t = pp1.optimize(np.zeros(n), "CG-LBFGS", iterations=40, lbfgs_k=1)
    assert np.all(np.diff(t) <= 1e-12)
    assert t[-1] < t[0]


def _exact_line_min(f, x, d):
    from scipy.optimize import minimize_scalar

    a = minimize_scalar(lambda t: f(x + t * d), bracket=(0.0, 1.0), tol=1e-14).x
    return x + a * d


def test_cg_gamma_is_polak_ribiere_plus_by_hand():
    """src/frprmn.cu:154-175: gg = sum g^2, dgg = sum (xi + g) xi with xi the
    new gradient and g the old negated gradient, gam = max(0, dgg / gg),
    h <- -xi + gam h.  The path is recomputed here with exact line searches
    on a convex non-quadratic function.  At step 2 Polak-Ribiere and
    Fletcher-Reeves coincide (exact line search makes the gradients
    orthogonal); at step 3 they do not, and the oracle follows PR."""
    r = np.random.default_rng(7)
    n = 3
    Q = np.linalg.qr(r.standard_normal((n, n)))[0]
    A = Q @ np.diag([1.0, 3.0, 7.0]) @ Q.T
    b = r.standard_normal(n)
    wq = np.array([0.5, 1.0, 2.0])
    f = lambda p: 0.5 * p @ A @ p - b @ p + np.sum(wq * p ** 4)
    grad = lambda p: A @ p - b + 4.0 * wq * p ** 3

    def by_hand(x0, iters, rule):
        x, gk = x0.copy(), grad(x0)
        d = -gk
        pts = []
        for _ in range(iters):
            x = _exact_line_min(f, x, d)
            pts.append(x.copy())
            gn = grad(x)
            if rule == "PR+":
                gam = max(0.0, float((gn - gk) @ gn) / float(gk @ gk))
            else:  # Fletcher-Reeves
                gam = float(gn @ gn) / float(gk @ gk)
            d = -gn + gam * d
            gk = gn
        return pts

    x0 = np.array([1.0, -1.0, 0.5])
    pr, fr = by_hand(x0, 3, "PR+"), by_hand(x0, 3, "FR")
    pp = oracle.PyProblem(n, 1, f, lambda p, it: grad(p))
    for it in (1, 2, 3):
        x, res, _ = pp.optimize(x0, "CG-FRPRMN", iterations=it)
        assert res.iterations == it
        assert np.linalg.norm(x - pr[it - 1]) < 1e-6, it
    # the check tells the two rules apart: FR's third point is far away
    assert np.linalg.norm(pr[2] - fr[2]) > 1e-2
    x3, _, _ = pp.optimize(x0, "CG-FRPRMN", iterations=3)
    assert np.linalg.norm(x3 - fr[2]) > 1e-2


def test_lbfgs_y_and_h0_by_hand():
    """src/lbfgs.cu:139-180 with calculateSandY (src/functions.cu:3635-3653):
    y = xi_new - (-xi_old), where xi_old is the search direction copied
    before linmin, so y = g_new + d_old (= g_new - g_old only on the first
    iteration); s = p - p_old; the initial Hessian scale is
    sum_images s.y / y.y (src/lbfgs.cu:270-289).  With one image the
    reference's storage addressing leaves a memory-1 two-loop recursion.
    Recomputed here with exact line searches on a 4-D quadratic: the oracle
    follows it, and the textbook y = g_new - g_old, or an unscaled H0, ends
    measurably elsewhere after three iterations."""
    r = np.random.default_rng(11)
    n = 4
    Q = np.linalg.qr(r.standard_normal((n, n)))[0]
    A = Q @ np.diag([1.0, 2.5, 6.0, 15.0]) @ Q.T
    b = r.standard_normal(n)
    f = lambda p: 0.5 * p @ A @ p - b @ p
    grad = lambda p: A @ p - b

    def by_hand(x0, iters, y_rule="reference", scale_h0=True):
        x, gk = x0.copy(), grad(x0)
        d = -gk
        pts = []
        for _ in range(iters):
            a = -float(grad(x) @ d) / float(d @ A @ d)
            xn = x + a * d
            gn = grad(xn)
            s = xn - x
            y = gn + d if y_rule == "reference" else gn - gk
            rho = 1.0 / float(y @ s)
            alpha = rho * float(s @ gn)
            q = gn - alpha * y
            h0 = float(s @ y) / float(y @ y) if scale_h0 else 1.0
            rr = q * h0
            beta = rho * float(y @ rr)
            rr = rr + (alpha - beta) * s
            d = -rr
            x, gk = xn, gn
            pts.append(x.copy())
        return pts

    x0 = np.zeros(n)
    ref = by_hand(x0, 3)
    textbook = by_hand(x0, 3, y_rule="textbook")
    unscaled = by_hand(x0, 3, scale_h0=False)
    pp = oracle.PyProblem(n, 1, f, lambda p, it: grad(p))
    for it in (1, 2, 3):
        x, res, _ = pp.optimize(x0, "CG-LBFGS", iterations=it, lbfgs_k=10)
        assert res.iterations == it
        assert np.linalg.norm(x - ref[it - 1]) < 1e-6, it
    x3, _, _ = pp.optimize(x0, "CG-LBFGS", iterations=3, lbfgs_k=10)
    assert np.linalg.norm(x3 - textbook[2]) > 1e-2
    assert np.linalg.norm(x3 - unscaled[2]) > 1e-2


def test_objective_gating_and_chi2_first(prob):
    """ObjectiveFunction: priors are 0 at iteration 0 and active after a
    gradient call with iter > 0; Chi2 overwrites dphi (it must be first)."""
    o = oracle.Oracle(prob)
    of = o.objective([("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)],
                     lambdas=(0.01, 1e-3))
    img = prob.initial_image().astype(np.float64)
    img[0] += RNG.random((prob.N, prob.M)) * 5
    v0, clipped = o.value(of, img)
    chi2, vr, _ = o.chi2(img)
    assert np.isclose(v0, chi2, rtol=1e-12)
    g0 = o.gradient(of, clipped, 0)
    assert np.allclose(g0[0], o.dchi2_channel(vr))
    g1 = o.gradient(of, clipped, 1)
    v1, _ = o.value(of, clipped)
    s = o.prior_value("entropy", clipped, prior=0.001)
    t = o.prior_value("tsv", clipped)
    assert np.isclose(v1, chi2 + 0.01 * s + 1e-3 * t, rtol=1e-12)
    ge = o.prior_grad("entropy", clipped, lam=0.01, prior=0.001)
    gt = o.prior_grad("tsv", clipped, lam=1e-3)
    assert np.allclose(g1[0], o.dchi2_channel(vr) + ge + gt)
    o.free(of)


# ------------------------------------------------------------------ set-up (a19)
def test_setup_closed_forms():
    import ctypes as C

    L = oracle.lib()
    # 4 points (+-a, 0), (0, +-a), equal weights: round beam 1/(sqrt 2 pi a)
    a_l = 1.0e5
    freq = np.float32(230e9)
    lam = np.float32(2.99792458e8) / freq
    uvw = np.array([[a_l * lam, 0, 0], [-a_l * lam, 0, 0], [0, a_l * lam, 0], [0, -a_l * lam, 0]], np.float64)
    vo = np.zeros((4, 2), np.float32)
    w = np.full(4, 2.0, np.float32)
    ch = (oracle.OChan * 1)(oracle.OChan(freq, 4, uvw.ctypes.data, vo.ctypes.data, w.ctypes.data))
    out = [C.c_double() for _ in range(5)]
    L.orc_beam_and_noise(ch, 1, -1.0 / 3600, *[C.addressof(o) for o in out])
    bmaj, bmin, bpa, sw, vn = (o.value for o in out)
    expect = 1.0 / np.sqrt(2.0) / np.pi / a_l * 180 / np.pi
    assert np.isclose(bmaj, expect, rtol=1e-6) and np.isclose(bmin, expect, rtol=1e-6)
    assert sw == 8.0 and np.isclose(vn, 0.5 / np.sqrt(8.0))
    # Hermitian flip: u > 0 -> (-u, -v), conj Vo, then metres -> wavelengths
    uvw = np.array([[3.0, 4.0, 5.0], [-1.0, 2.0, 3.0]])
    vo = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    L.orc_hermitian_to_lambda(2, uvw.ctypes.data, vo.ctypes.data, freq)
    assert np.allclose(uvw * lam, [[-3, -4, 5], [-1, 2, 3]], rtol=1e-12)
    assert np.allclose(vo, [[1, -2], [3, 4]])


def test_noise_image_mask_is_pb_threshold(prob):
    """noise = noise_jypix PBmax / PB: fg_scale = noise_jypix (the centre) and
    the mask noise < 10 fg_scale is PB > 0.1 PBmax."""
    import ctypes as C

    o = oracle.Oracle(prob)
    noise = np.zeros((prob.N, prob.M))
    nmin = C.c_double()
    L = oracle.lib()
    L.orc_noise_image(C.byref(o.beam), prob.freq, prob.M, prob.N, prob.deltax, prob.deltay,
                      prob.x0, prob.y0, prob.beam_info["noise_jypix"], noise.ctypes.data,
                      C.addressof(nmin))
    assert np.isclose(nmin.value, prob.beam_info["noise_jypix"], rtol=1e-12)
    pb = o.pb()
    with np.errstate(divide="ignore"):
        assert np.allclose(noise[pb > 0], prob.beam_info["noise_jypix"] / pb[pb > 0], rtol=1e-12)
    # the product's float noise image agrees with the oracle's
    fin = np.isfinite(noise)
    assert np.allclose(prob.noise[fin], noise[fin], rtol=1e-6)
    assert np.isclose(prob.fg_scale, nmin.value, rtol=1e-6)
    assert np.array_equal(prob.noise < prob.noise_cut, pb > 0.1 * pb.max() * (1 + 1e-9)) or \
        abs(int((prob.noise < prob.noise_cut).sum()) - int((pb > 0.1).sum())) <= 4


def test_do_gridding_pillbox_single_visibility():
    import ctypes as C

    M = N = 32
    du = dv = 1000.0
    freq = np.float32(230e9)
    lam = np.float32(2.99792458e8) / freq
    uvw = np.array([[2 * du * lam, -1 * dv * lam, 0.0]])
    vo = np.array([[1.5, -0.5]], np.float32)
    w = np.array([4.0], np.float32)
    ch = (oracle.OChan * 1)(oracle.OChan(freq, 1, uvw.ctypes.data, vo.ctypes.data, w.ctypes.data))
    re = np.zeros((N, M))
    im = np.zeros((N, M))
    gw = np.zeros((N, M))
    oracle.lib().orc_do_gridding_pillbox(M, N, du, dv, ch, 1, 3, 2, re.ctypes.data, im.ctypes.data, gw.ctypes.data)
    # the sample lands at (16+2, 16-1) and its conjugate at (16-2, 16+1), 7x7 boxes each
    assert np.isclose(re[15, 18], 1.5) and np.isclose(im[15, 18], -0.5)
    assert np.isclose(re[17, 14], 1.5) and np.isclose(im[17, 14], 0.5)
    assert np.isclose(gw[15, 18], 4.0)
    # boxes rows 12-18 x cols 15-21 and rows 14-20 x cols 11-17 overlap on 5x3 cells
    assert np.count_nonzero(gw) == 2 * 49 - 5 * 3
    ov = (slice(14, 19), slice(15, 18))
    assert np.allclose(re[ov], 1.5) and np.allclose(im[ov], 0.0)
