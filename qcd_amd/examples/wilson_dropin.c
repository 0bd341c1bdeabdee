/* wilson_dropin.c -- a C caller of the drop-in API, written the way the
 * reference's drivers are (b_minus_Dx.c:82-167, tests/test_wilson.c:49-79):
 * qcd_initGeometry / qcd_initVector / qcd_initGaugeField, then
 * qcd_applyWilsonTMOp with m = 1/(2 kappa) - 4, the QTM operator and norms.
 *
 * usage: wilson_dropin T X Y Z kappa mu eps seed_u seed_psi out.bin
 * Writes D psi (complex128, [V][4][3]) to out.bin and prints ||D psi||,
 * ||Q psi|| and ||Q psi - i g5 D psi|| / ||D psi||.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <qcd.h>
#include <qcd_hip.h>

int main(int argc, char **argv)
{
   qcd_geometry geo;
   qcd_vector psi, dpsi, qpsi;
   qcd_gaugeField u;
   qcd_uint_2 L[4], P[4] = {1, 1, 1, 1};
   qcd_real_8 theta[4] = {M_PI, 0.0, 0.0, 0.0};
   qcd_real_8 kappa, mu, eps, nd, nq, diff = 0.0;
   unsigned long seed_u, seed_psi, l;
   int s, c;
   FILE *f;

   if (argc != 11) {
      fprintf(stderr, "usage: %s T X Y Z kappa mu eps seed_u seed_psi out.bin\n", argv[0]);
      return EXIT_FAILURE;
   }
   for (s = 0; s < 4; s++) L[s] = (qcd_uint_2)atoi(argv[1 + s]);
   kappa = atof(argv[5]);
   mu = atof(argv[6]);
   eps = atof(argv[7]);
   seed_u = strtoul(argv[8], NULL, 10);
   seed_psi = strtoul(argv[9], NULL, 10);

   if (qcd_initGeometry(&geo, L, P, theta, 0, 1)) return EXIT_FAILURE;
   if (qcd_initVector(&psi, &geo) || qcd_initVector(&dpsi, &geo) || qcd_initVector(&qpsi, &geo) ||
       qcd_initGaugeField(&u, &geo)) {
      fprintf(stderr, "not enough memory\n");
      return EXIT_FAILURE;
   }
   if (qcd_hip_gauge_generate_host(&u.D[0][0][0][0].re, (long)geo.lV, eps, seed_u) ||
       qcd_hip_spinor_gaussian_host(&psi.D[0][0][0].re, (long)geo.lV, seed_psi)) {
      fprintf(stderr, "field generation failed: %s\n", qcd_hip_last_error());
      return EXIT_FAILURE;
   }

   if (qcd_applyWilsonTMOp(&dpsi, &psi, &u, 1.0 / (2.0 * kappa) - 4.0, mu)) return EXIT_FAILURE;
   if (qcd_applyQTMOp(&qpsi, &psi, &u, 1.0 / (2.0 * kappa) - 4.0, mu)) return EXIT_FAILURE;
   nd = qcd_normVector(&dpsi);
   nq = qcd_normVector(&qpsi);
   /* Q = i gamma5 D (lib/src/qcd_wilson.c:805-812) */
   for (l = 0; l < geo.lV; l++)
      for (s = 0; s < 4; s++)
         for (c = 0; c < 3; c++) {
            qcd_complex_16 d = dpsi.D[l][s][c], q = qpsi.D[l][s][c];
            qcd_complex_16 e = (s < 2) ? (qcd_complex_16){-d.im, d.re} : (qcd_complex_16){d.im, -d.re};
            diff += (q.re - e.re) * (q.re - e.re) + (q.im - e.im) * (q.im - e.im);
         }
   printf("|D psi| = %.15e\n|Q psi| = %.15e\n|Q psi - i g5 D psi| / |D psi| = %.3e\n", nd, nq, sqrt(diff) / nd);

   f = fopen(argv[10], "wb");
   if (!f || fwrite(dpsi.D, sizeof(qcd_complex_16), 12 * geo.lV, f) != 12 * geo.lV) {
      fprintf(stderr, "could not write %s\n", argv[10]);
      return EXIT_FAILURE;
   }
   fclose(f);
   qcd_destroyVector(&psi);
   qcd_destroyVector(&dpsi);
   qcd_destroyVector(&qpsi);
   qcd_destroyGaugeField(&u);
   qcd_destroyGeometry(&geo);
   return EXIT_SUCCESS;
}
