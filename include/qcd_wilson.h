/* qcd_wilson.h -- the twisted-mass Wilson Dirac operator of the MI355X library.
 *
 * Drop-in counterpart of the reference's include/qcd_wilson.h.  The operator
 * entry points (include/qcd_wilson.h:634-637) run on the GPU.  The per-site
 * spinor macros that caller code uses (qcd_wilson.h:17-624) are kept with
 * their names and argument meaning; here they expand to small inline helpers.
 *
 * Spinor convention (qcd_wilson.h:11-15): a site spinor is 24 reals indexed
 * i = (re 0 / im 1) + 2*colour + 6*spin.  A half spinor is the first 12.
 * Gamma basis (lib/src/qcd_gamma.c:46-88), chiral, gamma5 = diag(1,1,-1,-1).
 *
 * (1 + s*gamma_mu) psi, s = +1 or -1, is fixed by its two upper spins
 *    h0 = psi_0 + s*a_mu*psi_{p_mu},   h1 = psi_1 + s*b_mu*psi_{q_mu}
 * and the lower spins are     psi_2 <- s*c_mu*h_{r_mu},  psi_3 <- s*d_mu*h_{w_mu}
 * with a..d in {+1,+i,-1,-i}, stored below as powers of i.
 *
 * Operator (lib/src/qcd_wilson.c:114-336, m = 1/(2 kappa) - 4):
 *  D psi(x) = (m+4) psi(x) + i tm gamma5 psi(x)
 *   - 1/2 sum_mu [ e^{+i theta_mu/L_mu} U_mu(x)      (1-gamma_mu) psi(x+mu)
 *                + e^{-i theta_mu/L_mu} U_mu(x-mu)^+ (1+gamma_mu) psi(x-mu) ]
 */
#ifndef H_QCD_WILSON
#define H_QCD_WILSON 1

#ifdef __cplusplus
extern "C" {
#endif

/* powers of i and partner spins of the projector table, per direction mu */
static const int qcd_proj_a_[4] = {2, 3, 2, 3};
static const int qcd_proj_b_[4] = {2, 3, 0, 1};
static const int qcd_proj_c_[4] = {2, 1, 0, 1};
static const int qcd_proj_d_[4] = {2, 1, 2, 3};
static const int qcd_proj_p_[4] = {2, 3, 3, 2};
static const int qcd_proj_q_[4] = {3, 2, 2, 3};
static const int qcd_proj_r_[4] = {0, 1, 1, 0};
static const int qcd_proj_w_[4] = {1, 0, 0, 1};

/* (re,im) * i^p */
static inline void qcd_ipow_mul_(double re, double im, int p, double *ore, double *oim)
{
   switch (p & 3) {
   case 0:  *ore =  re; *oim =  im; break;
   case 1:  *ore = -im; *oim =  re; break;
   case 2:  *ore = -re; *oim = -im; break;
   default: *ore =  im; *oim = -re; break;
   }
}

/* x = (1 + s*gamma_mu) y; full != 0 also writes the 12 lower reals of x */
static inline void qcd_project_(double *x, const double *y, int mu, int s, int full)
{
   int c;
   int sp = (s > 0) ? 0 : 2;
   for (c = 0; c < 3; c++) {
      double re, im;
      qcd_ipow_mul_(y[2*c + 6*qcd_proj_p_[mu]], y[2*c + 1 + 6*qcd_proj_p_[mu]],
                    qcd_proj_a_[mu] + sp, &re, &im);
      x[2*c]     = y[2*c] + re;
      x[2*c + 1] = y[2*c + 1] + im;
      qcd_ipow_mul_(y[2*c + 6*qcd_proj_q_[mu]], y[2*c + 1 + 6*qcd_proj_q_[mu]],
                    qcd_proj_b_[mu] + sp, &re, &im);
      x[6 + 2*c]     = y[6 + 2*c] + re;
      x[6 + 2*c + 1] = y[6 + 2*c + 1] + im;
   }
   if (full)
      for (c = 0; c < 3; c++) {
         qcd_ipow_mul_(x[2*c + 6*qcd_proj_r_[mu]], x[2*c + 1 + 6*qcd_proj_r_[mu]],
                       qcd_proj_c_[mu] + sp, &x[12 + 2*c], &x[13 + 2*c]);
         qcd_ipow_mul_(x[2*c + 6*qcd_proj_w_[mu]], x[2*c + 1 + 6*qcd_proj_w_[mu]],
                       qcd_proj_d_[mu] + sp, &x[18 + 2*c], &x[19 + 2*c]);
      }
}

/* x = U y (dag == 0) or U^+ y (dag != 0) for nspin colour vectors of y;
 * u is a 3x3 complex matrix as 18 reals, row major */
static inline void qcd_link_apply_(const double *u, double *x, const double *y, int nspin, int dag)
{
   int s, r, k;
   for (s = 0; s < nspin; s++)
      for (r = 0; r < 3; r++) {
         double re = 0.0, im = 0.0;
         for (k = 0; k < 3; k++) {
            double ur = dag ? u[2*(3*k + r)] : u[2*(3*r + k)];
            double ui = dag ? -u[2*(3*k + r) + 1] : u[2*(3*r + k) + 1];
            double yr = y[6*s + 2*k], yi = y[6*s + 2*k + 1];
            re += ur*yr - ui*yi;
            im += ur*yi + ui*yr;
         }
         x[6*s + 2*r] = re;
         x[6*s + 2*r + 1] = im;
      }
}

/* x = (mp4 + i mu gamma5) y */
static inline void qcd_mass_terms_(double *x, const double *y, double mp4, double mu)
{
   int i;
   for (i = 0; i < 12; i++) {
      double g = (i < 6) ? mu : -mu;
      x[2*i]     = mp4*y[2*i] - g*y[2*i + 1];
      x[2*i + 1] = mp4*y[2*i + 1] + g*y[2*i];
   }
}

/* x = a - 1/2 (b+c+...+i) where b,d,f,h come from (1+gamma_mu) and c,e,g,i
 * from (1-gamma_mu), mu = 0..3; half != 0 means b..i are half spinors whose
 * lower spins are rebuilt here (qcd_wilson.h:539-564) */
static inline void qcd_sum_op_(double *x, const double *a, const double *const h[8], int half)
{
   int j, n, mu, s;
   double acc[24];
   for (j = 0; j < 24; j++) acc[j] = 0.0;
   for (n = 0; n < 8; n++) {
      mu = n >> 1;
      s = (n & 1) ? -1 : 1;
      if (!half) {
         for (j = 0; j < 24; j++) acc[j] += h[n][j];
      } else {
         double full[24];
         int sp = (s > 0) ? 0 : 2, c;
         for (j = 0; j < 12; j++) full[j] = h[n][j];
         for (c = 0; c < 3; c++) {
            qcd_ipow_mul_(full[2*c + 6*qcd_proj_r_[mu]], full[2*c + 1 + 6*qcd_proj_r_[mu]],
                          qcd_proj_c_[mu] + sp, &full[12 + 2*c], &full[13 + 2*c]);
            qcd_ipow_mul_(full[2*c + 6*qcd_proj_w_[mu]], full[2*c + 1 + 6*qcd_proj_w_[mu]],
                          qcd_proj_d_[mu] + sp, &full[18 + 2*c], &full[19 + 2*c]);
         }
         for (j = 0; j < 24; j++) acc[j] += full[j];
      }
   }
   for (j = 0; j < 24; j++) x[j] = a[j] - 0.5*acc[j];
}

/* x *= (pr + i pi) for n complex entries */
static inline void qcd_phase_mul_(double *x, int n, double pr, double pi)
{
   int j;
   for (j = 0; j < n; j++) {
      double re = x[2*j]*pr - x[2*j + 1]*pi;
      x[2*j + 1] = x[2*j]*pi + x[2*j + 1]*pr;
      x[2*j] = re;
   }
}

#define qcd_ONE_PLUS_GAMMA0(x,y)            qcd_project_((x),(y),0, 1,1)
#define qcd_ONE_PLUS_GAMMA1(x,y)            qcd_project_((x),(y),1, 1,1)
#define qcd_ONE_PLUS_GAMMA2(x,y)            qcd_project_((x),(y),2, 1,1)
#define qcd_ONE_PLUS_GAMMA3(x,y)            qcd_project_((x),(y),3, 1,1)
#define qcd_ONE_PLUS_GAMMA0_HALFSPINOR(x,y) qcd_project_((x),(y),0, 1,0)
#define qcd_ONE_PLUS_GAMMA1_HALFSPINOR(x,y) qcd_project_((x),(y),1, 1,0)
#define qcd_ONE_PLUS_GAMMA2_HALFSPINOR(x,y) qcd_project_((x),(y),2, 1,0)
#define qcd_ONE_PLUS_GAMMA3_HALFSPINOR(x,y) qcd_project_((x),(y),3, 1,0)
#define qcd_ONE_MINUS_GAMMA0(x,y)            qcd_project_((x),(y),0,-1,1)
#define qcd_ONE_MINUS_GAMMA1(x,y)            qcd_project_((x),(y),1,-1,1)
#define qcd_ONE_MINUS_GAMMA2(x,y)            qcd_project_((x),(y),2,-1,1)
#define qcd_ONE_MINUS_GAMMA3(x,y)            qcd_project_((x),(y),3,-1,1)
#define qcd_ONE_MINUS_GAMMA0_HALFSPINOR(x,y) qcd_project_((x),(y),0,-1,0)
#define qcd_ONE_MINUS_GAMMA1_HALFSPINOR(x,y) qcd_project_((x),(y),1,-1,0)
#define qcd_ONE_MINUS_GAMMA2_HALFSPINOR(x,y) qcd_project_((x),(y),2,-1,0)
#define qcd_ONE_MINUS_GAMMA3_HALFSPINOR(x,y) qcd_project_((x),(y),3,-1,0)
#define qcd_MASSTERMS(x,y,mp4,mu)            qcd_mass_terms_((x),(y),(mp4),(mu))
#define qcd_APPLY_U(u,x,y)                   qcd_link_apply_((u),(x),(y),4,0)
#define qcd_APPLY_U_HALFSPINOR(u,x,y)        qcd_link_apply_((u),(x),(y),2,0)
#define qcd_APPLY_U_DAGGER(u,x,y)            qcd_link_apply_((u),(x),(y),4,1)
#define qcd_APPLY_U_DAGGER_HALFSPINOR(u,x,y) qcd_link_apply_((u),(x),(y),2,1)
#define qcd_SUM_OP(x,a,b,c,d,e,f,g,h,i) \
   do { const double *const qcd_h_[8] = {(b),(c),(d),(e),(f),(g),(h),(i)}; \
        qcd_sum_op_((x),(a),qcd_h_,0); } while (0)
#define qcd_SUM_OP_HALFSPINORS(x,a,b,c,d,e,f,g,h,i) \
   do { const double *const qcd_h_[8] = {(b),(c),(d),(e),(f),(g),(h),(i)}; \
        qcd_sum_op_((x),(a),qcd_h_,1); } while (0)
#define qcd_PHASE_MUL(x,t,pr,pi)             do { (t) = 0; qcd_phase_mul_((x),12,(pr),(pi)); } while (0)
#define qcd_PHASE_MUL_HALFSPINOR(x,t,pr,pi)  do { (t) = 0; qcd_phase_mul_((x),6,(pr),(pi)); } while (0)

/* Per-site helpers on one 4 x 3 spinor (include/qcd_wilson.h:627-632,
 * lib/src/qcd_wilson.c:18-103), host code: (1 +- gamma_mu) vec, U vec and
 * U^+ vec on all four spins, (mp4 + i tm gamma5) vec, and vec * phi. */
void qcd_onePlusGamma(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_uint_2 mu);
void qcd_oneMinusGamma(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_uint_2 mu);
void qcd_applyU(qcd_complex_16 result[][3], qcd_complex_16 u[][3], qcd_complex_16 vec[][3]);
void qcd_applyUDagger(qcd_complex_16 result[][3], qcd_complex_16 u[][3], qcd_complex_16 vec[][3]);
void qcd_massTerms(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_real_8 mp4, qcd_real_8 tm);
void qcd_phaseMul(qcd_complex_16 vec[][3], qcd_complex_16 phi);

/* Operators (include/qcd_wilson.h:634-637).  All three D spellings compute
 * the same operator; qcd_applyQTMOp computes Q = i gamma5 D
 * (lib/src/qcd_wilson.c:795-814).  Return 0 on success, 1 on invalid input
 * (uninitialised fields or mismatched geometry, lib/src/qcd_wilson.c:139-166)
 * or a device error.  Dv must not alias v. */
int qcd_applyWilsonTMOp(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm);
int qcd_applyWilsonTMOpNoScripts(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm);
int qcd_applyWilsonTMOpNoOptim(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm);
int qcd_applyQTMOp(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm);

/* Extensions without a reference counterpart: D (or Q) applied to all 12
 * columns (nu, c2) of a propagator in one batched device launch, the same
 * result as qcd_copyVectorPropagator + qcd_applyWilsonTMOp (or
 * qcd_applyQTMOp) + qcd_copyPropagatorVector per column
 * (lib/src/qcd_blas.c:182-212).  Return 0 on success, 1 on invalid input
 * (uninitialised fields, mismatched geometry, Dp aliasing p). */
int qcd_hip_applyWilsonTMOpPropagator(qcd_propagator *Dp, qcd_propagator *p, qcd_gaugeField *u, qcd_real_8 m,
                                      qcd_real_8 tm);
int qcd_hip_applyQTMOpPropagator(qcd_propagator *Dp, qcd_propagator *p, qcd_gaugeField *u, qcd_real_8 m,
                                 qcd_real_8 tm);

/* Extension without a reference counterpart: keep u resident in HBM so that
 * the drop-in operators, smearing and BLAS above skip its upload.  Library
 * calls that write u (qcd_apeSmear3d, qcd_stoutsmearing, qcd_projectSU33d,
 * qcd_copyGaugeField, qcd_zeroGaugeField, qcd_getGaugeLime) refresh the
 * device copy themselves; call again after changing u->D by other means;
 * unbind to release the device copy. */
int  qcd_hip_bindGaugeField(qcd_gaugeField *u);
void qcd_hip_unbindGaugeField(void);

#ifdef __cplusplus
}
#endif

#endif
