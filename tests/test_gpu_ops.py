"""GPU parity of the prior kernels and the optimiser vector operations,
through the C ABI, against the FP64 oracle / the reference kernel semantics."""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpuvmem_amd import native
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.problem import small_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    p = small_problem(npix=128, n_antennas=6)
    ctx = Context(p)
    orc = oracle.Oracle(p)
    yield p, ctx, orc
    ctx.close()


def _img(p, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros((2, p.N, p.M), np.float32)
    img[0] = (0.001 + rng.random((p.N, p.M)) * 3.0).astype(np.float32)
    img[1] = rng.standard_normal((p.N, p.M)).astype(np.float32)
    return img


@pytest.mark.parametrize("kind", ["entropy", "tsv", "laplacian"])
@pytest.mark.parametrize("index", [0, 1])
def test_prior_value_and_gradient(setup, kind, index):
    p, ctx, orc = setup
    print(f"case: prior {kind} image {index}", flush=True)
    img = _img(p, 3)
    if kind == "entropy":
        img[1] = np.abs(img[1]) + 0.01  # entropy needs positive pixels
    d = ctx.image(img)
    v = ctx.prior_value(kind, d, index=index, prior=0.001, eta=-1.0)
    ref = orc.prior_value(kind, img.astype(np.float64), index=index, prior=0.001, eta=-1.0)
    assert abs(v - ref) <= 2e-6 * abs(ref) + 1e-9, (v, ref)
    dgi = ctx.zeros(p.M * p.N)
    ctx.prior_grad(kind, d, dgi, index=index, lam=0.37, prior=0.001, eta=-1.0)
    ctx.sync()
    g = dgi.download(np.float32, (p.N, p.M))
    gref = orc.prior_grad(kind, img.astype(np.float64), index=index, lam=0.37, prior=0.001)
    assert np.linalg.norm(g - gref) <= 1e-6 * np.linalg.norm(gref), kind
    # dphi[index] += dgi
    dphi = ctx.zeros(2 * p.M * p.N)
    native.check(ctx.lib.gvm_add_to_dphi(ctx.ctx, dphi.ptr, dgi.ptr, index), "add")
    native.check(ctx.lib.gvm_add_to_dphi(ctx.ctx, dphi.ptr, dgi.ptr, index), "add")
    out = dphi.download(np.float32, (2, p.N, p.M))
    assert np.array_equal(out[index], 2 * g) and np.all(out[1 - index] == 0)


def test_entropy_eta_nonzero(setup):
    p, ctx, orc = setup
    img = _img(p, 4)
    d = ctx.image(img)
    v = ctx.prior_value("entropy", d, prior=0.01, eta=0.5)
    ref = orc.prior_value("entropy", img.astype(np.float64), prior=0.01, eta=0.5)
    assert abs(v - ref) <= 2e-6 * abs(ref)


def test_cg_vector_ops(setup):
    p, ctx, _ = setup
    lib = ctx.lib
    n = p.M * p.N
    rng = np.random.default_rng(7)
    xi0 = rng.standard_normal(2 * n).astype(np.float32)
    h0 = rng.standard_normal(2 * n).astype(np.float32)
    g0 = rng.standard_normal(2 * n).astype(np.float32)
    xi, h, g = DeviceBuffer.from_array(xi0), DeviceBuffer.from_array(h0), DeviceBuffer.from_array(g0)
    # gg / dgg over both images (getGGandDGG + reductions)
    gg, dgg = C.c_double(), C.c_double()
    native.check(lib.gvm_cg_gg_dgg(ctx.ctx, xi.ptr, g.ptr, 2, C.byref(gg), C.byref(dgg)), "gg")
    gd, xd = g0.astype(np.float64), xi0.astype(np.float64)
    assert np.isclose(gg.value, np.sum(gd * gd), rtol=1e-12)
    assert np.isclose(dgg.value, np.sum((xd + gd) * xd), rtol=1e-12)
    # CGGradCondition max
    pimg = rng.standard_normal(2 * n).astype(np.float32) * 3
    pb = DeviceBuffer.from_array(pimg)
    gm = C.c_double()
    native.check(lib.gvm_cg_grad_condition(ctx.ctx, xi.ptr, pb.ptr, 7.5, 2, C.byref(gm)), "gc")
    ref = np.max(np.abs(xd) * np.maximum(np.abs(pimg.astype(np.float64)), 1.0) / 7.5)
    assert np.isclose(gm.value, ref, rtol=1e-7)
    # newXi on image 1 only
    native.check(lib.gvm_cg_new_xi(ctx.ctx, g.ptr, xi.ptr, h.ptr, 0.25, 1), "newXi")
    go, xo, ho = g.download(), xi.download(), h.download()
    assert np.array_equal(go[:n], g0[:n]) and np.array_equal(xo[:n], xi0[:n])
    assert np.array_equal(go[n:], -xi0[n:])
    assert np.allclose(xo[n:], -xi0[n:] + np.float32(0.25) * h0[n:], rtol=1e-6) and np.array_equal(xo[n:], ho[n:])
    # searchDirection on image 0
    native.check(lib.gvm_cg_search_direction(ctx.ctx, g.ptr, xi.ptr, h.ptr, 0), "sd")
    go, xo, ho = g.download(), xi.download(), h.download()
    assert np.array_equal(go[:n], -xi0[:n]) and np.array_equal(xo[:n], -xi0[:n]) and np.array_equal(ho[:n], -xi0[:n])


def test_line_search_vector_ops(setup):
    p, ctx, _ = setup
    lib = ctx.lib
    n = p.M * p.N
    rng = np.random.default_rng(8)
    pc = rng.standard_normal(2 * n).astype(np.float32)
    xc = rng.standard_normal(2 * n).astype(np.float32)
    P, X, T = DeviceBuffer.from_array(pc), DeviceBuffer.from_array(xc), DeviceBuffer(8 * n)
    minpix, eta, x = 0.25, -1.0, 0.7
    floor = np.float32(-eta * minpix)
    for image, pos in ((0, 1), (1, 0)):
        native.check(lib.gvm_evaluate_xt(ctx.ctx, T.ptr, P.ptr, X.ptr, x, image, pos, minpix, eta), "xt")
    t = T.download()
    v0 = pc[:n] + np.float32(x) * xc[:n]
    assert np.allclose(t[:n], np.where(v0 > floor, v0, floor), rtol=1e-6, atol=1e-7)
    assert np.allclose(t[n:], pc[n:] + np.float32(x) * xc[n:], rtol=1e-6, atol=1e-7)
    # newP with positivity on image 0: xi *= xmin; p += xi or floor and xi = 0
    native.check(lib.gvm_new_p(ctx.ctx, P.ptr, X.ptr, 0.4, 0, 1, minpix, eta), "newP")
    native.check(lib.gvm_new_p(ctx.ctx, P.ptr, X.ptr, 0.4, 1, 0, minpix, eta), "newP")
    po, xo = P.download(), X.download()
    d = xc * np.float32(0.4)
    ok = pc[:n] + d[:n] > floor
    assert np.allclose(po[:n], np.where(ok, pc[:n] + d[:n], floor), rtol=1e-6, atol=1e-7)
    assert np.allclose(xo[:n], np.where(ok, d[:n], 0.0), rtol=1e-6)
    assert np.allclose(po[n:], pc[n:] + d[n:], rtol=1e-6, atol=1e-7)


def test_lbfgs_vector_ops(setup):
    p, ctx, _ = setup
    lib = ctx.lib
    n = p.M * p.N
    rng = np.random.default_rng(9)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    A, B = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
    out = C.c_double()
    native.check(lib.gvm_vec_dot(ctx.ctx, A.ptr, B.ptr, n, C.byref(out)), "dot")
    assert np.isclose(out.value, np.dot(a.astype(np.float64), b.astype(np.float64)), rtol=1e-12)
    native.check(lib.gvm_vec_absmax(ctx.ctx, A.ptr, n, C.byref(out)), "absmax")
    assert out.value == np.abs(a).max()
    native.check(lib.gvm_vec_axpy(ctx.ctx, A.ptr, B.ptr, -0.5, n), "axpy")
    assert np.allclose(A.download(), a - np.float32(0.5) * b, rtol=1e-6, atol=1e-7)
    native.check(lib.gvm_vec_scale(ctx.ctx, B.ptr, B.ptr, -1.0, n), "scale")
    assert np.array_equal(B.download(), -b)
    y, s = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    pn, xn = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    po, xo = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    bufs = [DeviceBuffer.from_array(v) for v in (pn, xn, po, xo)]
    native.check(lib.gvm_lbfgs_s_and_y(ctx.ctx, y.ptr, s.ptr, *[q.ptr for q in bufs], n), "sy")
    assert np.array_equal(y.download(), xn + xo) and np.array_equal(s.download(), pn - po)
    # copy / zero
    native.check(lib.gvm_vec_copy(ctx.ctx, y.ptr, s.ptr, n), "copy")
    assert np.array_equal(y.download(), pn - po)
    native.check(lib.gvm_vec_zero(ctx.ctx, y.ptr, n), "zero")
    assert np.all(y.download() == 0)


def test_enqueued_values_match_synchronous(setup):
    """The asynchronous value forms leave exactly the synchronous results in
    the chosen value slots, read back together with one call."""
    p, ctx, _ = setup
    lib = ctx.lib
    img = _img(p, 6)
    d = ctx.image(img)
    chi2 = ctx.chi2(d)
    ent = ctx.prior_value("entropy", d, index=0, prior=0.001, eta=-1.0)
    tsv = ctx.prior_value("tsv", d, index=1)
    lap = ctx.prior_value("laplacian", d, index=0)
    native.check(lib.gvm_chi2_enqueue(ctx.ctx, d.ptr, 0, 0, 3), "chi2 enqueue")
    native.check(lib.gvm_entropy_value_enqueue(ctx.ctx, d.ptr, 0, 0.001, -1.0, 5), "entropy")
    native.check(lib.gvm_tsv_value_enqueue(ctx.ctx, d.ptr, 1, 1), "tsv")
    native.check(lib.gvm_laplacian_value_enqueue(ctx.ctx, d.ptr, 0, 15), "laplacian")
    out = (C.c_double * 16)()
    native.check(lib.gvm_values_read(ctx.ctx, 16, out), "read")
    assert (out[3], out[5], out[1], out[15]) == (chi2, ent, tsv, lap)
    assert lib.gvm_tsv_value_enqueue(ctx.ctx, d.ptr, 1, 16) == 1  # slot out of range
    assert lib.gvm_values_read(ctx.ctx, 17, out) == 1


def test_dchi2_requires_chi2_first():
    p = small_problem(npix=64, n_antennas=5)
    ctx = Context(p)
    img = ctx.image(p.initial_image())
    res = ctx.zeros(2 * p.M * p.N)
    with pytest.raises(native.GvmError) as e:
        ctx.dchi2(img, res)
    assert e.value.status == 4
    ctx.close()
