"""The committed golden vectors (tests/golden/, written by make_golden.py from
numpy.fft and brute-force FP64 sums) against the C oracle (CPU) and against
the HIP path through the C ABI (GPU)."""
import ast
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle
from gpuvmem_amd.problem import Problem

GOLD = Path(__file__).resolve().parent / "golden"


def _load():
    z = np.load(GOLD / "forward_adjoint_64.npz")
    g = ast.literal_eval(str(z["geometry"][0]))
    p = Problem(M=g["M"], N=g["N"], deltax=g["deltax"], deltay=g["deltay"], x0=g["x0"], y0=g["y0"],
                freq=g["freq"], nu_0=g["nu_0"], uvw=z["uvw"], vo=z["vo"], weight=z["weight"],
                noise=z["noise"], fg_scale=g["fg_scale"], noise_cut=g["noise_cut"],
                beam_kind=g["beam_kind"], antenna_diameter=g["antenna_diameter"],
                pb_factor=g["pb_factor"], pb_cutoff=g["pb_cutoff"], minpix=g["minpix"],
                eta=g["eta"], threshold=g["threshold"])
    return p, z


def test_oracle_fft_matches_golden():
    z = np.load(GOLD / "fft_inverse_8x16.npz")
    x = np.ascontiguousarray(z["x"])
    y = oracle.fft2d(x, +1)
    assert np.allclose(y, z["y"], rtol=0, atol=1e-12 * np.abs(z["y"]).max())


def test_oracle_forward_and_adjoint_match_golden():
    p, z = _load()
    orc = oracle.Oracle(p)
    assert np.allclose(orc.pb(), z["pb"], rtol=1e-12, atol=0)
    v, vr, clipped = orc.chi2(z["image"])
    assert abs(v - float(z["chi2"])) <= 1e-12 * float(z["chi2"])
    assert np.allclose(vr, z["vr"], rtol=0, atol=1e-12 * np.abs(z["vr"]).max())
    assert np.array_equal(clipped, z["clipped"])
    assert np.array_equal(orc.weight, z["weight_after"].astype(np.float32))
    assert np.sum(z["weight_after"] == 0) == len(range(0, p.nvis, 37))
    for flag, key in ((0, "grad_flag0"), (1, "grad_flag1")):
        ref = z[key]
        for sep in (False, True):
            g = orc.dchi2(clipped, vr, flag_opt=flag, separable=sep)
            err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
            assert err < (1e-12 if not sep else 1e-9), (flag, sep, err)


@pytest.mark.gpu
def test_hip_path_matches_golden():
    from gpuvmem_amd.device import Context

    p, z = _load()
    print(f"case: golden 64^2, nvis={p.nvis}", flush=True)
    ref_chi2 = float(z["chi2"])
    for mode in ("f16x3", "f32", "bf16x3", "direct"):
        ctx = Context(p, adjoint=mode)
        d = ctx.image(z["image"])
        v = ctx.chi2(d)
        assert abs(v - ref_chi2) <= 1e-6 * ref_chi2, (mode, v, ref_chi2)
        np.testing.assert_array_equal(d.download(np.float32, (2, p.N, p.M)),
                                      z["clipped"].astype(np.float32))
        for flag, key in ((0, "grad_flag0"), (1, "grad_flag1")):
            res = ctx.zeros(2 * p.M * p.N)
            ctx.dchi2(d, res, flag_opt=flag)
            ctx.sync()
            g = res.download(np.float32, (2, p.N, p.M))
            ref = z[key]
            err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
            assert err < 1e-5, (mode, flag, err)
        ctx.close()
