/* invert.c -- propagator solve driver on one MI355X, the counterpart of the
 * reference's invert.c:114-294.
 *
 * usage: invert <parameter file>
 * Reads the reference's tags <processors_txyz> (must be 1 1 1 1),
 * <lattice_txyz>, <source_type> (HMC_PROPAGATOR), <source>, <solution_type>,
 * <solution>, <cfg_name> (ILDG/LIME gauge field), <N_restart>, <kappa>, <mu>,
 * and these optional ones:
 *    <precision>f16</precision>  inner precision f64 | f32 | f16 (default f16)
 *    <maxRes>1e-10</maxRes>      target relative residue (invert.c hard-codes 1e-8)
 *    <eo>1</eo>                  even-odd preconditioned inner solve (default:
 *                                on when every extent is even)
 *    <batch>12</batch>           right-hand sides solved together (default 12)
 *    <beta>0</beta>              beta written in the solution's header
 * As invert.c does, it solves Q x = src for the 12 source vectors (mu, col) of
 * the propagator file, Q = i g5 D with m = 1/(2 kappa) - 4 (invert.c:262-289),
 * i.e. D x = -i g5 src, and it stops when ||src - Q x|| / ||src|| <= maxRes.
 * The 12 solutions are written to <solution> as an HMC propagator file (the
 * reference computes them and does not store them).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <qcd.h>
#include <qcd_hip.h>

static int tag(char *params, int len, const char *name, char *out, size_t n) {
   char t[128];
   snprintf(t, sizeof(t), "<%s>", name);
   if (!strstr(params, t)) return 0;   /* optional tags may be absent */
   snprintf(out, n, "%s", qcd_getParam(t, params, len));
   return 1;
}

int main(int argc, char **argv)
{
   qcd_geometry geo;
   qcd_gaugeField u;
   qcd_vector src[12], sol[12];
   qcd_uint_2 L[4], P[4];
   qcd_real_8 theta[4] = {M_PI, 0.0, 0.0, 0.0};   /* antiperiodic in time (invert.c:138) */
   char buf[4096], src_type[256], src_name[4096], sol_name[4096], cfg_name[4096];
   double kappa, muTM, maxRes = 1e-10, beta = 0.0;
   unsigned Nrestart;
   int params_len, r, s, c, prec = QCD_HIP_F16, eo = -1, batch = 12, rc = 0;
   char *params, *text;

   if (argc != 2) {
      fprintf(stderr, "No input file specified\n");
      return EXIT_FAILURE;
   }
   printf("opening input file %s\n", argv[1]);
   if (!(params = qcd_getParams(argv[1], &params_len))) return EXIT_FAILURE;
   /* qcd_getParam cuts the buffer at the next '<': read each tag from a fresh copy */
   text = malloc(params_len);
#define GET(name, dst) (memcpy(text, params, params_len), tag(text, params_len, name, dst, sizeof(dst)))
   if (!GET("processors_txyz", buf) || sscanf(buf, "%hu %hu %hu %hu", &P[0], &P[1], &P[2], &P[3]) != 4 ||
       !GET("lattice_txyz", buf) || sscanf(buf, "%hu %hu %hu %hu", &L[0], &L[1], &L[2], &L[3]) != 4 ||
       !GET("source_type", src_type) || !GET("source", src_name) || !GET("solution", sol_name) ||
       !GET("cfg_name", cfg_name) || !GET("N_restart", buf) || sscanf(buf, "%u", &Nrestart) != 1 ||
       !GET("kappa", buf) || sscanf(buf, "%lf", &kappa) != 1 || !GET("mu", buf) || sscanf(buf, "%lf", &muTM) != 1) {
      fprintf(stderr, "Error: the parameter file lacks a required tag\n");
      return EXIT_FAILURE;
   }
   if (GET("precision", buf)) prec = strstr(buf, "64") ? QCD_HIP_F64 : strstr(buf, "32") ? QCD_HIP_F32 : QCD_HIP_F16;
   if (GET("maxRes", buf)) maxRes = atof(buf);
   if (GET("eo", buf)) eo = atoi(buf);
   if (GET("batch", buf)) batch = atoi(buf);
   if (GET("beta", buf)) beta = atof(buf);
#undef GET
   free(text);
   free(params);
   printf("Got source type: %s\nGot source file name: %s\nGot solution file name: %s\nGot conf name: %s\n",
          src_type, src_name, sol_name, cfg_name);
   printf("Got N_restart: %u\nGot kappa: %e\nGot mu: %e\n", Nrestart, kappa, muTM);
   if (strstr(src_type, "HMC_PROPAGATOR") == NULL) {
      fprintf(stderr, "Error: source type %s is not supported (HMC_PROPAGATOR)\n", src_type);
      return EXIT_FAILURE;
   }
   if (qcd_initGeometry(&geo, L, P, theta, 0, 1)) return EXIT_FAILURE;
   printf(" Local lattice: %i x %i x %i x %i\n", geo.lL[0], geo.lL[1], geo.lL[2], geo.lL[3]);
   if (qcd_initGaugeField(&u, &geo)) return EXIT_FAILURE;
   for (r = 0; r < 12; r++)
      if (qcd_initVector(&src[r], &geo) || qcd_initVector(&sol[r], &geo)) {
         printf("not enough memory\n");
         return EXIT_FAILURE;
      }
   if (qcd_getGaugeField(cfg_name, qcd_GF_LIME, &u)) return EXIT_FAILURE;
   printf("gauge-field loaded, plaquette = %.12f\n", qcd_calculatePlaquette(&u));
   for (r = 0; r < 12; r++)
      if (qcd_getVector(src_name, qcd_PROP_HMC, r / 3, r % 3, &src[r])) return EXIT_FAILURE;
   printf("12 source vectors from %s loaded\n", src_name);

   {
      int Li[4] = {L[0], L[1], L[2], L[3]};
      qcd_hip_lattice lat;
      qcd_hip_field g, b, x;
      qcd_hip_solver S;
      qcd_hip_solver_params p;
      qcd_hip_solver_stats st;
      qcd_vector tmp;
      if (qcd_hip_lattice_create(&lat, Li, theta) || qcd_hip_gauge_create(&g, lat, QCD_HIP_F64) ||
          qcd_hip_gauge_upload(g, &u.D[0][0][0][0].re) || qcd_hip_spinor_create(&b, lat, QCD_HIP_F64, 12) ||
          qcd_hip_spinor_create(&x, lat, QCD_HIP_F64, 12) || qcd_initVector(&tmp, &geo)) {
         fprintf(stderr, "Error: %s\n", qcd_hip_last_error());
         return EXIT_FAILURE;
      }
      /* Q x = src  <=>  D x = -i g5 src */
      for (r = 0; r < 12; r++) {
         unsigned long l;
         for (l = 0; l < geo.lV; l++)
            for (s = 0; s < 4; s++)
               for (c = 0; c < 3; c++) {
                  const qcd_complex_16 v = src[r].D[l][s][c];
                  const double sg = s < 2 ? 1.0 : -1.0;   /* -i g5 (re, im) = sg * (im, -re) */
                  tmp.D[l][s][c].re = sg * v.im;
                  tmp.D[l][s][c].im = -sg * v.re;
               }
         if (qcd_hip_spinor_upload(b, r, &tmp.D[0][0][0].re)) return EXIT_FAILURE;
      }
      qcd_hip_solver_default_params(&p);
      p.n_restart = (int)Nrestart;
      p.tol = maxRes;
      p.inner_prec = prec;
      p.op = QCD_HIP_OP_D;
      p.batch = batch;
      p.max_outer = 100000;
      p.eo = eo >= 0 ? eo : (L[0] % 2 == 0 && L[1] % 2 == 0 && L[2] % 2 == 0 && L[3] % 2 == 0);
      printf("inner precision %s, even-odd %d, batch %d, maxRes %e\n",
             prec == QCD_HIP_F64 ? "f64" : prec == QCD_HIP_F32 ? "f32" : "f16", p.eo, p.batch, maxRes);
      if (qcd_hip_solver_create(&S, lat, &p) || qcd_hip_solver_set_gauge(S, g)) {
         fprintf(stderr, "Error: %s\n", qcd_hip_last_error());
         return EXIT_FAILURE;
      }
      rc = qcd_hip_solver_solve(S, x, b, kappa, muTM, &st);
      printf("solved 12 vectors in %.3f s: %d GCR cycles of %u iterations\n", st.seconds, st.outer, Nrestart);
      for (r = 0; r < 12; r++) {
         double rel = 0.0;
         qcd_hip_b_minus_Dx(x, r, b, r, g, kappa, muTM, &rel);
         printf("vector mu = %d, col = %d: relative residue |src - Q x| / |src| = %e\n", r / 3, r % 3, rel);
         if (!(rel <= maxRes)) rc = rc ? rc : 6;
         qcd_hip_spinor_download(x, r, &sol[r].D[0][0][0].re);
      }
      qcd_hip_solver_destroy(S);
      qcd_hip_field_destroy(x);
      qcd_hip_field_destroy(b);
      qcd_hip_field_destroy(g);
      qcd_hip_lattice_destroy(lat);
      qcd_destroyVector(&tmp);
   }
   if (rc) fprintf(stderr, "Error: not converged (%d): %s\n", rc, qcd_hip_last_error());
   if (qcd_writePropagatorHMC(sol_name, sol, beta, kappa, muTM)) return EXIT_FAILURE;
   printf("solutions written to %s\n", sol_name);
   for (r = 0; r < 12; r++) {
      qcd_destroyVector(&src[r]);
      qcd_destroyVector(&sol[r]);
   }
   qcd_destroyGaugeField(&u);
   qcd_destroyGeometry(&geo);
   return rc ? EXIT_FAILURE : EXIT_SUCCESS;
}
