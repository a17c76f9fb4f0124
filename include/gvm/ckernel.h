/* Gridding (anti-aliasing) kernels and their correction images (host C):
 * the CKernel family of the reference (include/classes/ckernel.cuh,
 * src/pillBox2D.cu, src/gaussian2D.cu, src/sinc2D.cu, src/gaussianSinc2D.cu,
 * src/pswf_12D.cu).  The taps feed gvm_grid_visibilities_dev, the correction
 * image gvm_set_gcf (include/gvm_hip.h). */
#ifndef GVM_CKERNEL_H
#define GVM_CKERNEL_H

#ifdef __cplusplus
extern "C" {
#endif

#define GVM_CKERNEL_PILLBOX 0        /* PillBox2D, the reference's default (src/main.cu:152) */
#define GVM_CKERNEL_GAUSSIAN 1       /* Gaussian2D */
#define GVM_CKERNEL_SINC 2           /* Sinc2D */
#define GVM_CKERNEL_GAUSSIAN_SINC 3  /* GaussianSinc2D */
#define GVM_CKERNEL_PSWF 4           /* PSWF_12D */

typedef struct {
  int kind;
  int m, n;    /* rows, columns */
  float amp;   /* amplitude (1) */
  float w;     /* width parameter w */
  float w2;    /* second width (GaussianSinc2D) */
  float alpha; /* exponent (Gaussian2D, GaussianSinc2D) */
} gvm_ckernel_params;

/* The parameters of the kind's default constructor: PillBox2D() is 1 x 1
 * (src/pillBox2D.cu:22-25); the others are 7 x 7 (CKernel(),
 * include/classes/ckernel.cuh:101-114) with w = 1 (Gaussian2D, Sinc2D),
 * 2.52 (GaussianSinc2D, w2 = 1.55) or 6 (PSWF_12D), alpha = 2.
 * Returns 0, or 1 for an unknown kind. */
int gvm_ckernel_defaults(int kind, gvm_ckernel_params* p);

/* CKernel::buildKernel() after setSigmas(|deltau|, |deltav|)
 * (src/mfs.cu:512-513): taps[n*i + j] at x = (j - support_x) sigma_x,
 * y = (i - support_y) sigma_y, supports floor(m / 2) (both from m,
 * CKernel::setSupports).  taps: m*n floats.  Returns 0 or 1. */
int gvm_ckernel_build(const gvm_ckernel_params* p, float sigma_x, float sigma_y, float* taps,
                      int* support_x, int* support_y);

/* CKernel::initializeGCF(M, N, |dx|, |dy|) (include/classes/ckernel.cuh:82-88,
 * src/mfs.cu:515): the correction image gcf[N*i + j] (M*N floats) of a
 * kernel of kind p->kind, with w = M.  PillBox2D: 1; Gaussian2D: the
 * Gaussian at pi x with width 2 w (src/gaussian2D.cu:194-205); PSWF_12D:
 * 1 / kernel (src/pswf_12D.cu:261-271).  Sinc2D and GaussianSinc2D define
 * no correction image in the reference (its GCF would read past the taps):
 * returns 1 for them (and for an unknown kind), 0 otherwise. */
int gvm_ckernel_gcf(const gvm_ckernel_params* p, int M, int N, float dx, float dy, float* gcf);

#ifdef __cplusplus
}
#endif
#endif
