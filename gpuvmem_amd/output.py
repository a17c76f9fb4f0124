"""Output of the imaging path: FITS images of the objective's images and the
model / residual visibilities in the frame of the observation (the
reference's OCopyFITS, src/MSFITSIO.cu:94-168, and modelToHost,
src/MSFITSIO.cu:1114-1138)."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import native
from .problem import Problem


def fits_header(p: Problem, bunit: str = "JY/PIXEL", niter: int = 0, ra_deg: float = 0.0,
                dec_deg: Optional[float] = None, object_name: Optional[str] = None
                ) -> native.FitsHeader:
    """Header of an image of problem p: the reference pixel is the phase
    centre (CRPIX is 1-based), the cell sizes are CDELT1 / CDELT2, and the
    frequency axis holds nu_0."""
    h = native.FitsHeader()
    h.crpix1, h.crpix2 = p.x0 + 1.0, p.y0 + 1.0
    h.crval1 = ra_deg
    h.crval2 = (p.sim.declination_deg if p.sim is not None else 0.0) if dec_deg is None else dec_deg
    h.cdelt1, h.cdelt2 = p.deltax, p.deltay
    h.crval3, h.cdelt3 = p.nu_0, 1.0
    h.equinox = 2000.0
    h.radesys, h.bunit = b"ICRS", bunit.encode()
    h.object = object_name.encode() if object_name else None
    h.niter = niter
    return h


def write_fits(path: str, image, p: Problem, index: int = 0, fg_scale: Optional[float] = None,
               **header) -> None:
    """Writes image `index` of `image` (a device buffer or a host array of
    shape (image_count, N, M)) times fg_scale (the problem's by default for
    I_nu_0, 1 for alpha) to `path`."""
    fw = native.load_fw()
    if fg_scale is None:
        fg_scale = p.fg_scale if index == 0 else 1.0
    if index == 1:
        header.setdefault("bunit", "")
    h = fits_header(p, **header)
    if hasattr(image, "ptr"):
        ptr, on_device = image.ptr, 1
    else:
        arr = np.ascontiguousarray(image, np.float32)
        ptr, on_device = arr.ctypes.data, 0
    rc = fw.gvm_fits_write_image(path.encode(), ptr, on_device, p.M, p.N, index, fg_scale,
                                 C.byref(h))
    if rc != 0:
        raise OSError(fw.gvm_fits_last_error().decode(errors="replace"))


def read_fits(path: str):
    """Minimal reader for the files write_fits produces: returns (header
    dict, float32 image of shape (NAXIS2, NAXIS1))."""
    raw = open(path, "rb").read()
    hdr, pos = {}, 0
    while True:
        cardtxt = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = cardtxt[:8].strip()
        if key == "END":
            break
        if cardtxt[8:10] == "= ":
            val = cardtxt[10:].split(" / ")[0].strip()
            if val.startswith("'"):
                hdr[key] = val.strip("'").rstrip()
            elif val in ("T", "F"):
                hdr[key] = val == "T"
            else:
                hdr[key] = float(val) if any(c in val for c in ".E") else int(val)
    pos = (pos + 2879) // 2880 * 2880
    n1, n2 = hdr["NAXIS1"], hdr["NAXIS2"]
    data = np.frombuffer(raw[pos:pos + 4 * n1 * n2], dtype=">f4").astype(np.float32)
    return hdr, data.reshape(n2, n1)
