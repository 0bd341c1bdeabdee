/* qcd_smearing.h -- 3D Gaussian (Wuppertal) smearing of the MI355X library.
 *
 * Drop-in counterpart of the reference's include/qcd_smearing.h:11-40 and
 * lib/src/qcd_smearing.c:25-199.  One iteration on time-slice t is
 *    v(x) <- ( v(x) + alpha * sum_{nu=1..3} [ U_nu(x) v(x+nu) + U_nu(x-nu)^+ v(x-nu) ] ) / (1 + 6 alpha)
 * with no spin structure and no boundary phase; other time-slices are left
 * unchanged.  The "All" variant updates every time-slice.  Both run on the GPU.
 */
#ifndef H_QCD_SMEARING
#define H_QCD_SMEARING 1

#ifdef __cplusplus
extern "C" {
#endif

/* t += a + b over a 24-real spinor (qcd_smearing.h:11-36) */
#define qcd_SUM_UP_HOPP(t,a,b) \
   do { int qcd_j_; for (qcd_j_ = 0; qcd_j_ < 24; qcd_j_++) (t)[qcd_j_] += (a)[qcd_j_] + (b)[qcd_j_]; } while (0)

int qcd_gaussIteration3d(qcd_vector *v, qcd_gaugeField *u, qcd_real_8 alpha, qcd_uint_4 t);
/* gaugeCom selects whether the reference exchanges gauge faces; with one rank
 * it has no effect. */
int qcd_gaussIteration3dAll(qcd_vector *v, qcd_gaugeField *u, qcd_real_8 alpha, qcd_uint_2 gaugeCom);
/* One step of 3d APE smearing (include/qcd_smearing.h:41, lib/src/qcd_smearing.c:205-273):
 * spatial links -> P_SU3[ U + alpha * sum of the 4 spatial staples ], temporal
 * links copied.  Runs on the GPU; apeu may alias u. */
int qcd_apeSmear3d(qcd_gaugeField *apeu, qcd_gaugeField *u, qcd_real_8 alpha);
/* One stout smearing step (include/qcd_smearing.h:42, lib/src/qcd_smearing.c:374-476):
 * U' = exp(iQ) U with Q the traceless Hermitian part of rho * (staples) U^+,
 * ndim = 3 (spatial links, temporal copied) or 4.  Runs on the GPU; u_out
 * may alias u.  The reference exits the process on another ndim; this
 * returns 1 with a message.  Where the reference's exponential divides 0 by
 * 0 (Q = 0, e.g. the unit gauge field) the result here is exact. */
int qcd_stoutsmearing(qcd_gaugeField *u_out, qcd_gaugeField *u, qcd_real_8 rho, int ndim);

#ifdef __cplusplus
}
#endif

#endif
