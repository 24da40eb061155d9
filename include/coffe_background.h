/*
 * coffe_background.h -- background tables (input provider of the multipole
 * weights), restating coffe_background_init (background.c:301-697) for
 * w(z) = w0 + wa z/(1+z).  All outputs are host arrays of length `bins` on the
 * grid z_i = 15 i/(bins-1); units Mpc and 1/Mpc as in the reference.
 * Magnification (s) and evolution (f_evo) biases are linear interpolants of the
 * n_bias samples given on bias_z.  Returns 0 on success.
 */
#ifndef COFFE_BACKGROUND_H
#define COFFE_BACKGROUND_H

#ifdef __cplusplus
extern "C" {
#endif

int coffe_background_compute(double Omega0_cdm, double Omega0_baryon, double Omega0_gamma, double w0, double wa,
                             double h, int bins, int n_bias, const double *bias_z, const double *s1, const double *s2,
                             const double *fevo1, const double *fevo2, double *z, double *a, double *Hz,
                             double *conformal_Hz, double *conformal_Hz_prime, double *D1, double *f, double *chi,
                             double *G1, double *G2);

#ifdef __cplusplus
}
#endif
#endif
