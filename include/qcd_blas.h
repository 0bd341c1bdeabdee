/* qcd_blas.h -- complex helpers and the vector BLAS of the MI355X library.
 *
 * Drop-in counterpart of the vector half of the reference's include/qcd_blas.h
 * (macros :12-39 and :40-221, prototypes :242-267).  Every function below has
 * the reference's name, signature and meaning (lib/src/qcd_blas.c:137-491).
 * Arithmetic entry points run on the GPU: host arrays are copied to the device,
 * the kernel runs there, and results are copied back.  Pure data movement on
 * host arrays (zero / set / copy / column copies) stays a host copy.
 */
#ifndef H_QCD_BLAS
#define H_QCD_BLAS 1

#ifdef __cplusplus
extern "C" {
#endif

/* ---- complex arithmetic on qcd_complex_16 values (qcd_blas.h:12-39) ---- */
#define qcd_CONJ(x)           ( (qcd_complex_16) {(x).re, -(x).im} )
#define qcd_CMUL(x,y)         ( (qcd_complex_16) {(x).re*(y).re - (x).im*(y).im, (x).re*(y).im + (x).im*(y).re} )
#define qcd_CMULR(x,y)        ( (qcd_real_8) ((x).re*(y).re - (x).im*(y).im) )
#define qcd_CMULI(x,y)        ( (qcd_real_8) ((x).re*(y).im + (x).im*(y).re) )
#define qcd_CADJOINTMUL(x,y)  ( (qcd_complex_16) {(x).re*(y).re + (x).im*(y).im, (x).re*(y).im - (x).im*(y).re} )
#define qcd_CADD(x,y)         ( (qcd_complex_16) {(x).re + (y).re, (x).im + (y).im} )
#define qcd_CADDR(x,y)        ( (qcd_real_8) ((x).re + (y).re) )
#define qcd_CADDI(x,y)        ( (qcd_real_8) ((x).im + (y).im) )
#define qcd_CSUB(x,y)         ( (qcd_complex_16) {(x).re - (y).re, (x).im - (y).im} )
#define qcd_CSUBR(x,y)        ( (qcd_real_8) ((x).re - (y).re) )
#define qcd_CSUBI(x,y)        ( (qcd_real_8) ((x).im - (y).im) )
#define qcd_CSCALE(x,a)       ( (qcd_complex_16) {(x).re*(a), (x).im*(a)} )
#define qcd_CSCALER(x,a)      ( (qcd_real_8) ((x).re*(a)) )
#define qcd_CSCALEI(x,a)      ( (qcd_real_8) ((x).im*(a)) )
#define qcd_NORMSQUARED(x)    ( (qcd_real_8) ((x).re*(x).re + (x).im*(x).im) )
#define qcd_NORM(x)           ( (qcd_real_8) sqrt(qcd_NORMSQUARED(x)) )
#define qcd_ARG(x)            ( (qcd_real_8) atan2((x).im, (x).re) )
/* x / y */
#define qcd_CDEV(x,y)         ( (qcd_complex_16) {((x).re*(y).re + (x).im*(y).im)/qcd_NORMSQUARED(y), \
                                                  ((x).im*(y).re - (x).re*(y).im)/qcd_NORMSQUARED(y)} )
#define qcd_CDEVR(x,y)        ( (qcd_real_8) (((x).re*(y).re + (x).im*(y).im)/qcd_NORMSQUARED(y)) )
#define qcd_CDEVI(x,y)        ( (qcd_real_8) (((x).im*(y).re - (x).re*(y).im)/qcd_NORMSQUARED(y)) )
/* x^a on the principal branch */
#define qcd_CPOW(x,a)         ( (qcd_complex_16) {pow(qcd_NORM(x),(a))*cos(qcd_ARG(x)*(a)), \
                                                  pow(qcd_NORM(x),(a))*sin(qcd_ARG(x)*(a))} )
#define qcd_CPOWR(x,a)        ( (qcd_real_8) (pow(qcd_NORM(x),(a))*cos(qcd_ARG(x)*(a))) )
#define qcd_CPOWI(x,a)        ( (qcd_real_8) (pow(qcd_NORM(x),(a))*sin(qcd_ARG(x)*(a))) )

/* ---- 3x3 complex matrices (qcd_blas.h:40-221): a = b*c, a = b*c^+, a = b^+*c ---- */
#define qcd_MUL3x3(a,b,c)        qcd_mat3_mul_((a),(b),(c),0,0)
#define qcd_MULADJOINT3x3(a,b,c) qcd_mat3_mul_((a),(b),(c),0,1)
#define qcd_ADJOINTMUL3x3(a,b,c) qcd_mat3_mul_((a),(b),(c),1,0)
#define qcd_SU3TRACER(u)         ( (qcd_real_8) ((u)[0][0].re + (u)[1][1].re + (u)[2][2].re) )

/* Helper behind the 3x3 macros: op(b) * op(c) with op = identity or adjoint. */
static inline void qcd_mat3_mul_(qcd_complex_16 a[3][3], qcd_complex_16 b[3][3],
                                 qcd_complex_16 c[3][3], int adj_b, int adj_c)
{
   int i, j, k;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
         qcd_real_8 re = 0.0, im = 0.0;
         for (k = 0; k < 3; k++) {
            qcd_complex_16 p = adj_b ? b[k][i] : b[i][k];
            qcd_complex_16 q = adj_c ? c[j][k] : c[k][j];
            if (adj_b) p.im = -p.im;
            if (adj_c) q.im = -q.im;
            re += p.re * q.re - p.im * q.im;
            im += p.re * q.im + p.im * q.re;
         }
         a[i][j].re = re;
         a[i][j].im = im;
      }
}

/* ---- vector BLAS (qcd_blas.h:242-267; lib/src/qcd_blas.c:137-491) ---- */
/* 3x3 complex matrix helpers (include/qcd_blas.h:224-241, lib/src/qcd_blas.c:19-135):
 * host code; addAdjoint = a + b^+, mulAdjoint = a b^+ */
void qcd_zero3x3(qcd_complex_16 a[3][3]);
void qcd_unit3x3(qcd_complex_16 a[3][3]);
void qcd_minusUnit3x3(qcd_complex_16 a[3][3]);
void qcd_sub3x3(qcd_complex_16 difference[3][3], qcd_complex_16 minuend[3][3], qcd_complex_16 subtrahend[3][3]);
void qcd_add3x3(qcd_complex_16 sum[3][3], qcd_complex_16 summand1[3][3], qcd_complex_16 summand2[3][3]);
void qcd_addAdjoint3x3(qcd_complex_16 sum[3][3], qcd_complex_16 summand1[3][3], qcd_complex_16 summand2[3][3]);
void qcd_mul3x3(qcd_complex_16 product[3][3], qcd_complex_16 factor1[3][3], qcd_complex_16 factor2[3][3]);
void qcd_mulAdjoint3x3(qcd_complex_16 product[3][3], qcd_complex_16 factor1[3][3], qcd_complex_16 factor2[3][3]);
void qcd_scale3x3(qcd_complex_16 a[3][3], qcd_real_8 r);
void qcd_cScale3x3(qcd_complex_16 a[3][3], qcd_complex_16 c);
qcd_complex_16 qcd_trace3x3(qcd_complex_16 a[3][3]);
void qcd_copy3x3(qcd_complex_16 dest[3][3], qcd_complex_16 src[3][3]);
void qcd_dagger3x3(qcd_complex_16 a[3][3]);
void qcd_zeroVector(qcd_vector *vec);
void qcd_zeroPropagator(qcd_propagator *prop);
void qcd_zeroGaugeField(qcd_gaugeField *u);
void qcd_setVector(qcd_vector *vec, qcd_complex_16 c);
void qcd_copyVector(qcd_vector *dest, qcd_vector *src);
void qcd_copyGaugeField(qcd_gaugeField *dest, qcd_gaugeField *src);
void qcd_copyVectorPropagator(qcd_vector *vec, qcd_propagator *prop, qcd_uint_2 nu, qcd_uint_2 c2);
void qcd_copyPropagatorVector(qcd_propagator *prop, qcd_vector *vec, qcd_uint_2 nu, qcd_uint_2 c2);
void qcd_copyVectorPropagator2(qcd_vector *vec, qcd_propagator *prop, qcd_uint_2 mu, qcd_uint_2 c1);
void qcd_copyPropagatorVector2(qcd_propagator *prop, qcd_vector *vec, qcd_uint_2 mu, qcd_uint_2 c1);
void qcd_mulVectorC(qcd_vector *vec, qcd_complex_16 c);
void qcd_mulVectorC3d(qcd_vector *vec, qcd_complex_16 c, qcd_uint_4 t);
void qcd_scaleVector(qcd_vector *vec, qcd_real_8 r);
void qcd_scaleVector3d(qcd_vector *vec, qcd_real_8 r, qcd_uint_4 tt);
void qcd_subVector(qcd_vector *difference, qcd_vector *minuend, qcd_vector *subtrahend);
void qcd_addVector(qcd_vector *sum, qcd_vector *summand1, qcd_vector *summand2);
/* axpy = a*x + y.  Unlike lib/src/qcd_blas.c:414-432, axpy may alias x or y. */
void qcd_axpyVector(qcd_vector *axpy, qcd_complex_16 a, qcd_vector *x, qcd_vector *y);
void qcd_addVector3d(qcd_vector *sum, qcd_vector *summand1, qcd_vector *summand2, qcd_uint_4 tt);
/* sum over all 12V entries of conj(factor1)*factor2 */
qcd_complex_16 qcd_mulAdjointVectorVector(qcd_vector *factor1, qcd_vector *factor2);
/* sqrt of the sum over all 24V reals of x^2 */
qcd_real_8 qcd_normVector(qcd_vector *vec);
/* spatial links projected onto SU(3) in place: the unitary polar factor times
 * exp(-i arg(det)/3) (include/qcd_blas.h:271, lib/src/qcd_blas.c:940-1364);
 * runs on the GPU */
void qcd_projectSU33d(qcd_gaugeField *gf);
/* every link projected onto SU(3) on the device (lib/src/qcd_blas.c:541-939);
 * one matrix on the host (:1376-1769): the unitary polar factor times
 * exp(-i arg(det) / 3), reached by Newton's iteration */
void qcd_projectSU3(qcd_gaugeField *gf);
void qcd_projectSU33x3(qcd_complex_16 g[3][3]);
/* propagator and gauge-field arithmetic (lib/src/qcd_blas.c:192-200, :261-285,
 * :358-390, :495-530), host code.  normGaugeField is the sum over links of
 * |tr U|.  randomGaugeField fills links with drand48() entries (re, then im)
 * and projects them onto SU(3) (:1771-1782). */
void qcd_copyPropagatorPropagator(qcd_propagator *prop_new, qcd_propagator *prop_old);
void qcd_mulPropagatorC(qcd_propagator *prop, qcd_complex_16 c);
void qcd_mulPropagatorC3d(qcd_propagator *prop, qcd_complex_16 c, qcd_uint_4 t);
void qcd_conjPropagator(qcd_propagator *prop);
void qcd_subGaugeField(qcd_gaugeField *difference, qcd_gaugeField *minuend, qcd_gaugeField *subtrahend);
void qcd_addGaugeField(qcd_gaugeField *sum, qcd_gaugeField *summand1, qcd_gaugeField *summand2);
void qcd_scaleGaugeField(qcd_gaugeField *u, qcd_real_8 alpha);
qcd_real_8 qcd_normGaugeField(qcd_gaugeField *u);
void qcd_randomGaugeField(qcd_gaugeField *u);

#ifdef __cplusplus
}
#endif

#endif
