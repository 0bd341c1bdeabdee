/* qcd_oracle.h -- CPU oracle for the Wilson Dslash / Gauss smearing / GCR path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used as the checker by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py.  The product library (qcd_amd/) never links,
 * loads or calls it.
 *
 * Data are host arrays in the reference's AoS layouts (include/qcd.h:90-126):
 *   spinor  double[V][4 spin][3 colour][2]         24 doubles per site
 *   gauge   double[V][4 mu][3 row][3 col][2]      72 doubles per site
 * with site l = t + L0*(x + L1*(y + L2*z)), L = {T, X, Y, Z}.
 */
#ifndef QCD_ORACLE_H
#define QCD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geo orc_geo;

/* neighbour tables as lib/src/qcd_init.c:110-124 (periodic wrap) */
orc_geo *orc_geo_create(const int L[4], const double theta[4]);
void     orc_geo_destroy(orc_geo *g);
long     orc_geo_volume(const orc_geo *g);
void     orc_set_threads(int n);

/* D psi, half-spinor form (lib/src/qcd_wilson.c:114-336) */
void orc_dslash_half(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm);
/* D psi, full-spinor matrix form (lib/src/qcd_wilson.c:570-792) */
void orc_dslash_full(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm);
/* Q psi = i gamma5 D psi (lib/src/qcd_wilson.c:795-814) */
void orc_qtm(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm);

/* Gauss smearing, one iteration (lib/src/qcd_smearing.c:25-107 and :116-199) */
void orc_gauss3d(const orc_geo *g, double *v, const double *u, double alpha, int t);
void orc_gauss3d_all(const orc_geo *g, double *v, const double *u, double alpha);

/* vector BLAS on V-site spinors (lib/src/qcd_blas.c:301-491) */
void   orc_scale(double *v, double r, long V);
void   orc_sub(double *d, const double *a, const double *b, long V);
void   orc_add(double *s, const double *a, const double *b, long V);
void   orc_axpy(double *out, double are, double aim, const double *x, const double *y, long V);
void   orc_mulc(double *v, double are, double aim, long V);
void   orc_dot(double out[2], const double *a, const double *b, long V);
double orc_norm(const double *v, long V);

/* GCR(n) as invert.c:20-113.  op = 0 uses Q (as invert.c), op = 1 uses D.
 * keep_q1 = 1 drops the k = 0 solution term exactly as invert.c:50-53 does.
 * Returns 0, or 1 when the Krylov space breaks down (invert.c:75-79). */
int orc_gcr(const orc_geo *g, double *sol, const double *src, const double *u,
            double kappa, double mu, int n, int op, int keep_q1);
/* Outer true-residual refinement around GCR (invert.c:255-294), solving
 * D x = b.  For op = 0 it solves Q x = i gamma5 b, the same system.
 * Returns 0 when ||b - D x||/||b|| <= max_res within max_outer cycles. */
int orc_invert(const orc_geo *g, double *x, const double *b, const double *u,
               double kappa, double mu, int n, int op, int keep_q1,
               double max_res, int max_outer, int *outer_out, double *relres_out);
/* ||D x - b|| / ||b|| with m = 1/(2 kappa) - 4 (b_minus_Dx.c:152-167) */
double orc_b_minus_Dx(const orc_geo *g, const double *x, const double *b, const double *u,
                      double kappa, double mu);

/* Synthetic inputs.  Gauge: U = reunitarize(exp(i eps H)), H traceless
 * Hermitian from a seeded Gaussian, one counter-based stream per (site, mu).
 * Spinor: i.i.d. N(0,1) real and imaginary parts, one stream per site. */
void orc_gauge_generate(double *u, long V, double eps, uint64_t seed);
void orc_spinor_gaussian(double *v, long V, uint64_t seed);
void orc_point_source(double *v, const int L[4], int t, int x, int y, int z, int spin, int col);
double orc_plaquette(const orc_geo *g, const double *u);

/* Projection of one 3x3 complex matrix (18 doubles, row-major) onto SU(3) in
 * place: the unitary polar factor of M times exp(-i arg(det M) / 3)
 * (lib/src/qcd_blas.c:940-1364). */
void orc_project_su3(double *m);
/* One step of 3d APE smearing of every spatial link, temporal links copied
 * (lib/src/qcd_smearing.c:205-273); out must not alias u. */
void orc_ape3d(const orc_geo *g, double *out, const double *u, double alpha);
/* exp(iQ) of one Hermitian traceless 3x3 matrix in place, Peardon-Morningstar
 * (lib/src/qcd_smearing.c:293-370); 1 + iQ - Q^2/2 when tr(Q^2)/2 < 1e-24,
 * and acos's argument clamped to 1 (both cases where the reference gives NaN). */
void orc_stout_exp(double *q);
/* One stout smearing step, ndim = 3 (spatial, temporal links copied) or 4
 * (lib/src/qcd_smearing.c:374-476); out must not alias u; 1 on bad ndim. */
int orc_stout(const orc_geo *g, double *out, const double *u, double rho, int ndim);

#ifdef __cplusplus
}
#endif

#endif
