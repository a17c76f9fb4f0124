/* Image output as FITS (host C): replaces the reference's OCopyFITS
 * (src/MSFITSIO.cu:94-168) for the images of the objective.  The reference
 * copies the header of an input FITS template and updates BUNIT, NITER,
 * NAXIS1/2, RADESYS, EQUINOX and CRVAL1/2; this writer builds the primary
 * header from the image geometry with the same keys, then writes the image
 * multiplied by fg_scale as 32-bit IEEE floats (BITPIX = -32, big-endian,
 * 2880-byte records). */
#ifndef GVM_FITSIO_H
#define GVM_FITSIO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double crpix1, crpix2;  /* reference pixel (1-based, FITS convention) */
  double crval1, crval2;  /* RA, Dec of the reference pixel (degrees) */
  double cdelt1, cdelt2;  /* cell sizes (degrees; CDELT1 < 0) */
  double crval3, cdelt3;  /* frequency axis (Hz): nu_0 and channel width, 0 if none */
  float equinox;          /* EQUINOX (2000) */
  const char* radesys;    /* RADESYS ("ICRS", "FK5") */
  const char* bunit;      /* BUNIT ("JY/PIXEL" for I_nu_0, "" for alpha) */
  const char* object;     /* OBJECT, may be NULL */
  int niter;              /* NITER: iterations of the optimiser */
} gvm_fits_header;

/* Writes image `index` of the image-major float[image_count*M*N] array
 * `image` (device pointer if on_device, else host), times fg_scale, to
 * `path` (overwritten).  Returns 0, or 1 on a bad argument / I/O error
 * (gvm_fits_last_error). */
int gvm_fits_write_image(const char* path, const float* image, int on_device, long M, long N,
                         int index, float fg_scale, const gvm_fits_header* hdr);

const char* gvm_fits_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
