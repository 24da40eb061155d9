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

/*
 * Jacobian-vector product of coffe_background_compute (no reference counterpart;
 * the differentiation of the background): par = (Omega0_cdm, Omega0_baryon,
 * Omega0_gamma, w0, wa, h), dir = a direction in the same six parameters.  out and
 * tout are host arrays of 10 x bins: the tables z, a, Hz, conformal_Hz,
 * conformal_Hz_prime, D1, f, chi, G1, G2 (as coffe_background_compute) and their
 * derivatives along dir, exact for the computed tables (fixed grid, RK4 steps and
 * quadrature nodes).  Returns 0 on success.
 */
int coffe_background_jvp(const double *par, const double *dir, int bins, int n_bias, const double *bias_z,
                         const double *s1, const double *s2, const double *fevo1, const double *fevo2, double *out,
                         double *tout);

#ifdef __cplusplus
}
#endif
#endif
