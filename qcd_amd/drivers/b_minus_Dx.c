/* b_minus_Dx.c -- residual check of one solution vector on one MI355X, the
 * counterpart of the reference's b_minus_Dx.c:40-175.
 *
 * usage: b_minus_Dx <parameter file>
 * Tags (as the reference): <processors_txyz> (1 1 1 1), <lattice_txyz>,
 * <cfg_name> (ILDG/LIME gauge field), <src_name> (LIME source vector b),
 * <sol_name> (HMC propagator holding x), <src_col>, <src_spin> (which of its
 * 12 vectors), <kappa>, <muT>, <eps>, <out_name>.  Prints the plaquette, |x|,
 * |b|, |Dx|, |Dx - b| and |Dx - b| / |b| with D the twisted-mass Wilson
 * operator at m = 1/(2 kappa) - 4 (b_minus_Dx.c:152-167), computed with the
 * drop-in API on the GPU; fails when the relative residue exceeds eps, and
 * writes x to <out_name> as a LIME propagator vector otherwise.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <qcd.h>

int main(int argc, char **argv)
{
   qcd_geometry geo;
   qcd_gaugeField u;
   qcd_vector vec, vec2, vec3, vec4;
   qcd_uint_2 L[4], P[4];
   qcd_real_8 theta[4] = {M_PI, 0.0, 0.0, 0.0};
   char gauge_name[4096], src_name[4096], sol_name[4096], out_name[4096];
   double kappa, muT, target_eps, residue, plaquette;
   int nu, c2, params_len;
   char *params, *text;

   if (argc != 2) {
      fprintf(stderr, "No input file specified\n");
      return EXIT_FAILURE;
   }
   printf("opening input file %s\n", argv[1]);
   if (!(params = qcd_getParams(argv[1], &params_len))) return EXIT_FAILURE;
   text = malloc(params_len);
   /* qcd_getParam cuts the buffer at the next '<': read each tag from a fresh copy */
#define GET(t) (memcpy(text, params, params_len), qcd_getParam(t, text, params_len))
   if (sscanf(GET("<processors_txyz>"), "%hu %hu %hu %hu", &P[0], &P[1], &P[2], &P[3]) != 4 ||
       sscanf(GET("<lattice_txyz>"), "%hu %hu %hu %hu", &L[0], &L[1], &L[2], &L[3]) != 4) {
      fprintf(stderr, "Error: bad <processors_txyz> or <lattice_txyz>\n");
      return EXIT_FAILURE;
   }
   snprintf(gauge_name, sizeof(gauge_name), "%s", GET("<cfg_name>"));
   snprintf(src_name, sizeof(src_name), "%s", GET("<src_name>"));
   snprintf(sol_name, sizeof(sol_name), "%s", GET("<sol_name>"));
   snprintf(out_name, sizeof(out_name), "%s", GET("<out_name>"));
   if (sscanf(GET("<src_col>"), "%i", &c2) != 1 || sscanf(GET("<src_spin>"), "%i", &nu) != 1 ||
       sscanf(GET("<kappa>"), "%lf", &kappa) != 1 || sscanf(GET("<muT>"), "%lf", &muT) != 1 ||
       sscanf(GET("<eps>"), "%lf", &target_eps) != 1) {
      fprintf(stderr, "Error: the parameter file lacks a required tag\n");
      return EXIT_FAILURE;
   }
#undef GET
   free(text);
   free(params);
   printf(" Got conf name: %s\n Got source name %s\n Got solution name %s\n", gauge_name, src_name, sol_name);
   printf(" Got source col=%i\n Got source spin=%i\n Got kappa=%f\n Got mu=%f\n Got eps=%e\n Got output name %s\n",
          c2, nu, kappa, muT, target_eps, out_name);
   if (qcd_initGeometry(&geo, L, P, theta, 0, 1)) return EXIT_FAILURE;
   if (qcd_initVector(&vec, &geo) || qcd_initVector(&vec2, &geo) || qcd_initVector(&vec3, &geo) ||
       qcd_initVector(&vec4, &geo) || qcd_initGaugeField(&u, &geo)) {
      fprintf(stderr, "not enough memory\n");
      return EXIT_FAILURE;
   }
   if (qcd_getGaugeField(gauge_name, qcd_GF_LIME, &u)) {
      fprintf(stderr, "process 0: Error reading gauge field!\n");
      return EXIT_FAILURE;
   }
   printf("got gauge-field\n");
   plaquette = qcd_calculatePlaquette(&u);
   printf("plaquette = %e\n", plaquette);
   if (plaquette > 1 || plaquette < 0.3) printf("POSSIBLE ERROR! SUSPICIOUS PLAQUETTE! PLEASE CHECK!\n");
   if (qcd_getVector(sol_name, qcd_PROP_HMC, (qcd_uint_2)nu, (qcd_uint_2)c2, &vec)) {
      fprintf(stderr, "process 0: Error reading vector!\n");
      return EXIT_FAILURE;
   }
   printf("|x|=%e\n", qcd_normVector(&vec));
   if (qcd_getVector(src_name, qcd_PROP_LIME, 0, 0, &vec2)) {
      fprintf(stderr, "process 0: Error reading vector!\n");
      return EXIT_FAILURE;
   }
   printf("|b|=%e\n", qcd_normVector(&vec2));
   if (qcd_applyWilsonTMOp(&vec3, &vec, &u, 1.0 / (2.0 * kappa) - 4.0, muT)) return EXIT_FAILURE;
   printf("|Dx|=%e\n", qcd_normVector(&vec3));
   qcd_subVector(&vec4, &vec3, &vec2);
   residue = qcd_normVector(&vec4);
   printf("|Dx-b|=%e\n", residue);
   residue /= qcd_normVector(&vec2);
   printf("residue calculated\n");
   printf("solution nu=%i c=%i: |Dx-b| / |b|= %e\n", nu, c2, residue);
   if (residue > target_eps) {
      printf("POSSIBLE ERROR! HIGH RESIDUE! PLEASE CHECK!\n");
      return EXIT_FAILURE;
   }
   if (qcd_writeVector(out_name, qcd_PROP_LIME, 0, 0, &vec)) return EXIT_FAILURE;
   qcd_destroyVector(&vec);
   qcd_destroyVector(&vec2);
   qcd_destroyVector(&vec3);
   qcd_destroyVector(&vec4);
   qcd_destroyGaugeField(&u);
   qcd_destroyGeometry(&geo);
   return EXIT_SUCCESS;
}
