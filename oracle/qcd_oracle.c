/* qcd_oracle.c -- CPU oracle: plain-C restatement of the reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see qcd_oracle.h).  Only tests/, the smoke check
 * and the cpu_baseline leg of bench.py load this; the product never does.
 *
 * Each routine names the reference lines it restates.  The reference itself
 * cannot be compiled in this environment (it needs mpi.h), so this oracle is
 * pinned by closed forms, identities and brute force in tests/.
 */
#include "qcd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_geo {
   int L[4];
   long V;
   long *plus;    /* plus[4*l + mu]  */
   long *minus;   /* minus[4*l + mu] */
   double theta[4];
};

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) { cplx r = {a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx conjc(cplx a) { cplx r = {a.re, -a.im}; return r; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
   if (n > 0) omp_set_num_threads(n);
#else
   (void)n;
#endif
}

/* ------------------------------------------------------------------ geometry
 * lib/src/qcd_init.c:110-124: plus/minus neighbour tables with periodic wrap.
 */
orc_geo *orc_geo_create(const int L[4], const double theta[4])
{
   orc_geo *g = (orc_geo *)calloc(1, sizeof(orc_geo));
   long t, x, y, z, V;
   if (!g) return NULL;
   memcpy(g->L, L, sizeof(g->L));
   memcpy(g->theta, theta, sizeof(g->theta));
   V = (long)L[0] * L[1] * L[2] * L[3];
   g->V = V;
   g->plus = (long *)malloc(sizeof(long) * 4 * V);
   g->minus = (long *)malloc(sizeof(long) * 4 * V);
   if (!g->plus || !g->minus) { orc_geo_destroy(g); return NULL; }
#define LEX(t_, x_, y_, z_) ((t_) + L[0] * ((x_) + L[1] * ((y_) + (long)L[2] * (z_))))
   for (z = 0; z < L[3]; z++)
      for (y = 0; y < L[2]; y++)
         for (x = 0; x < L[1]; x++)
            for (t = 0; t < L[0]; t++) {
               long l = LEX(t, x, y, z);
               g->plus[4*l + 0] = LEX((t + 1) % L[0], x, y, z);
               g->plus[4*l + 1] = LEX(t, (x + 1) % L[1], y, z);
               g->plus[4*l + 2] = LEX(t, x, (y + 1) % L[2], z);
               g->plus[4*l + 3] = LEX(t, x, y, (z + 1) % L[3]);
               g->minus[4*l + 0] = LEX((t - 1 + L[0]) % L[0], x, y, z);
               g->minus[4*l + 1] = LEX(t, (x - 1 + L[1]) % L[1], y, z);
               g->minus[4*l + 2] = LEX(t, x, (y - 1 + L[2]) % L[2], z);
               g->minus[4*l + 3] = LEX(t, x, y, (z - 1 + L[3]) % L[3]);
            }
#undef LEX
   return g;
}

void orc_geo_destroy(orc_geo *g)
{
   if (!g) return;
   free(g->plus);
   free(g->minus);
   free(g);
}

long orc_geo_volume(const orc_geo *g) { return g->V; }

/* --------------------------------------------------- spin projection tables
 * The chiral basis of lib/src/qcd_gamma.c:46-88.  For (1 + s*gamma_mu):
 *   h0 = psi0 + s*A[mu]*psi[P[mu]],  h1 = psi1 + s*B[mu]*psi[Q[mu]]
 *   spin2 = s*C[mu]*h[R[mu]],        spin3 = s*D[mu]*h[W[mu]]
 * (qcd_wilson.h:17-387).  Factors are powers of i.
 */
static const int PA[4] = {2, 3, 2, 3}, PB[4] = {2, 3, 0, 1};
static const int PC[4] = {2, 1, 0, 1}, PD[4] = {2, 1, 2, 3};
static const int PP[4] = {2, 3, 3, 2}, PQ[4] = {3, 2, 2, 3};
static const int PR[4] = {0, 1, 1, 0}, PW[4] = {1, 0, 0, 1};

static inline cplx ipow(cplx a, int p)
{
   cplx r;
   switch (p & 3) {
   case 0: r = a; break;
   case 1: r.re = -a.im; r.im = a.re; break;
   case 2: r.re = -a.re; r.im = -a.im; break;
   default: r.re = a.im; r.im = -a.re; break;
   }
   return r;
}

/* half spinor h[2][3] = upper spins of (1 + s gamma_mu) psi */
static void project_half(cplx h[2][3], const cplx psi[4][3], int mu, int s)
{
   int c, sp = (s > 0) ? 0 : 2;
   for (c = 0; c < 3; c++) {
      h[0][c] = cadd(psi[0][c], ipow(psi[PP[mu]][c], PA[mu] + sp));
      h[1][c] = cadd(psi[1][c], ipow(psi[PQ[mu]][c], PB[mu] + sp));
   }
}

/* full spinor from a half spinor of (1 + s gamma_mu) (qcd_wilson.h:539-564) */
static void reconstruct(cplx f[4][3], const cplx h[2][3], int mu, int s)
{
   int c, sp = (s > 0) ? 0 : 2;
   for (c = 0; c < 3; c++) {
      f[0][c] = h[0][c];
      f[1][c] = h[1][c];
      f[2][c] = ipow(h[PR[mu]][c], PC[mu] + sp);
      f[3][c] = ipow(h[PW[mu]][c], PD[mu] + sp);
   }
}

/* x = U y or U^+ y for nspin colour vectors (qcd_wilson.h:419-509) */
static void link_mul(cplx x[][3], const cplx U[3][3], const cplx y[][3], int nspin, int dag)
{
   int s, r, k;
   for (s = 0; s < nspin; s++)
      for (r = 0; r < 3; r++) {
         cplx acc = {0.0, 0.0};
         for (k = 0; k < 3; k++) {
            cplx uu = dag ? conjc(U[k][r]) : U[r][k];
            acc = cadd(acc, cmul(uu, y[s][k]));
         }
         x[s][r] = acc;
      }
}

/* hopping phases e^{+-i theta_mu / L_mu} (lib/src/qcd_wilson.c:172-186) */
static void phases(const orc_geo *g, cplx pp[4], cplx mp[4], int periodic[4])
{
   int b;
   for (b = 0; b < 4; b++) {
      if (g->theta[b] != 0.0) {
         double a = g->theta[b] / (double)g->L[b];
         pp[b].re = cos(a); pp[b].im = sin(a);
         mp[b].re = cos(a); mp[b].im = -sin(a);
         periodic[b] = 0;
      } else {
         pp[b].re = mp[b].re = 1.0;
         pp[b].im = mp[b].im = 0.0;
         periodic[b] = 1;
      }
   }
}

/* (m+4 + i tm gamma5) psi (qcd_wilson.h:389-417) */
static void mass_terms(cplx d[4][3], const cplx psi[4][3], double mp4, double tm)
{
   int s, c;
   for (s = 0; s < 4; s++) {
      double g = (s < 2) ? tm : -tm;
      for (c = 0; c < 3; c++) {
         d[s][c].re = mp4 * psi[s][c].re - g * psi[s][c].im;
         d[s][c].im = mp4 * psi[s][c].im + g * psi[s][c].re;
      }
   }
}

/* ------------------------------------------------- Dslash, half-spinor form
 * lib/src/qcd_wilson.c:114-336.  For each mu: backward hop
 * U_mu(x-mu)^+ (1+gamma_mu) psi(x-mu) times e^{-i theta/L}, forward hop
 * U_mu(x) (1-gamma_mu) psi(x+mu) times e^{+i theta/L}; then
 * out = diag - 1/2 (sum in the order back0, fwd0, back1, fwd1, ...).
 */
void orc_dslash_half(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm)
{
   cplx pp[4], mp[4];
   int periodic[4];
   long l;
   const double mp4 = 4.0 + m;
   phases(g, pp, mp, periodic);
#pragma omp parallel for schedule(static)
   for (l = 0; l < g->V; l++) {
      cplx hop[8][2][3];
      cplx diag[4][3];
      const cplx (*psi)[3];
      const cplx (*U)[3];
      cplx h[2][3];
      int mu, n, s, c;
      for (mu = 0; mu < 4; mu++) {
         long lm = g->minus[4*l + mu], lp = g->plus[4*l + mu];
         psi = (const cplx (*)[3])(in + 24 * lm);
         project_half(h, psi, mu, +1);
         U = (const cplx (*)[3])(u + 72 * lm + 18 * mu);
         link_mul(hop[2*mu], U, h, 2, 1);
         psi = (const cplx (*)[3])(in + 24 * lp);
         project_half(h, psi, mu, -1);
         U = (const cplx (*)[3])(u + 72 * l + 18 * mu);
         link_mul(hop[2*mu + 1], U, h, 2, 0);
         if (!periodic[mu])
            for (s = 0; s < 2; s++)
               for (c = 0; c < 3; c++) {
                  hop[2*mu][s][c] = cmul(hop[2*mu][s][c], mp[mu]);
                  hop[2*mu + 1][s][c] = cmul(hop[2*mu + 1][s][c], pp[mu]);
               }
      }
      mass_terms(diag, (const cplx (*)[3])(in + 24 * l), mp4, tm);
      {
         cplx full[8][4][3];
         cplx *o = (cplx *)(out + 24 * l);
         for (n = 0; n < 8; n++) reconstruct(full[n], hop[n], n >> 1, (n & 1) ? -1 : 1);
         for (s = 0; s < 4; s++)
            for (c = 0; c < 3; c++) {
               double re = 0.0, im = 0.0;
               for (n = 0; n < 8; n++) { re += full[n][s][c].re; im += full[n][s][c].im; }
               o[3*s + c].re = diag[s][c].re - 0.5 * re;
               o[3*s + c].im = diag[s][c].im - 0.5 * im;
            }
      }
   }
}

/* ------------------------------------------- Dslash, full-spinor matrix form
 * lib/src/qcd_wilson.c:570-792: (1 +- gamma_mu) applied as a 4x4 matrix on the
 * whole spinor, the link on all four spins.  The matrices are written out from
 * the gamma table of lib/src/qcd_gamma.c:46-74 (gamma_0..gamma_3).
 */
static const double GAMMA_RE[4][4][4] = {
   {{0,0,-1,0},{0,0,0,-1},{-1,0,0,0},{0,-1,0,0}},
   {{0,0,0,0},{0,0,0,0},{0,0,0,0},{0,0,0,0}},
   {{0,0,0,-1},{0,0,1,0},{0,1,0,0},{-1,0,0,0}},
   {{0,0,0,0},{0,0,0,0},{0,0,0,0},{0,0,0,0}}};
static const double GAMMA_IM[4][4][4] = {
   {{0,0,0,0},{0,0,0,0},{0,0,0,0},{0,0,0,0}},
   {{0,0,0,-1},{0,0,-1,0},{0,1,0,0},{1,0,0,0}},
   {{0,0,0,0},{0,0,0,0},{0,0,0,0},{0,0,0,0}},
   {{0,0,-1,0},{0,0,0,1},{1,0,0,0},{0,-1,0,0}}};

static void one_pm_gamma(cplx r[4][3], const cplx v[4][3], int mu, int s)
{
   int a, b, c;
   for (a = 0; a < 4; a++)
      for (c = 0; c < 3; c++) {
         cplx acc = v[a][c];
         for (b = 0; b < 4; b++) {
            cplx gm = {s * GAMMA_RE[mu][a][b], s * GAMMA_IM[mu][a][b]};
            acc = cadd(acc, cmul(gm, v[b][c]));
         }
         r[a][c] = acc;
      }
}

void orc_dslash_full(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm)
{
   cplx pp[4], mp[4];
   int periodic[4];
   long l;
   const double mp4 = 4.0 + m;
   phases(g, pp, mp, periodic);
#pragma omp parallel for schedule(static)
   for (l = 0; l < g->V; l++) {
      cplx hop[8][4][3], proj[4][3], diag[4][3];
      int mu, n, s, c;
      for (mu = 0; mu < 4; mu++) {
         long lm = g->minus[4*l + mu], lp = g->plus[4*l + mu];
         one_pm_gamma(proj, (const cplx (*)[3])(in + 24 * lm), mu, +1);
         link_mul(hop[2*mu], (const cplx (*)[3])(u + 72 * lm + 18 * mu), proj, 4, 1);
         one_pm_gamma(proj, (const cplx (*)[3])(in + 24 * lp), mu, -1);
         link_mul(hop[2*mu + 1], (const cplx (*)[3])(u + 72 * l + 18 * mu), proj, 4, 0);
         if (!periodic[mu])
            for (s = 0; s < 4; s++)
               for (c = 0; c < 3; c++) {
                  hop[2*mu][s][c] = cmul(hop[2*mu][s][c], mp[mu]);
                  hop[2*mu + 1][s][c] = cmul(hop[2*mu + 1][s][c], pp[mu]);
               }
      }
      mass_terms(diag, (const cplx (*)[3])(in + 24 * l), mp4, tm);
      {
         cplx *o = (cplx *)(out + 24 * l);
         for (s = 0; s < 4; s++)
            for (c = 0; c < 3; c++) {
               double re = 0.0, im = 0.0;
               for (n = 0; n < 8; n++) { re += hop[n][s][c].re; im += hop[n][s][c].im; }
               o[3*s + c].re = diag[s][c].re - 0.5 * re;
               o[3*s + c].im = diag[s][c].im - 0.5 * im;
            }
      }
   }
}

/* Q = i gamma5 D: spins 0,1 times i, spins 2,3 times -i (qcd_wilson.c:795-814) */
void orc_qtm(const orc_geo *g, double *out, const double *in, const double *u, double m, double tm)
{
   long l;
   orc_dslash_full(g, out, in, u, m, tm);
#pragma omp parallel for schedule(static)
   for (l = 0; l < g->V; l++) {
      int s, c;
      cplx *o = (cplx *)(out + 24 * l);
      for (s = 0; s < 4; s++)
         for (c = 0; c < 3; c++) o[3*s + c] = ipow(o[3*s + c], s < 2 ? 1 : 3);
   }
}

/* ------------------------------------------------------------ Gauss smearing
 * lib/src/qcd_smearing.c:25-107: on slice t, v2 = sum_{nu=1..3} U_nu(x) v(x+nu)
 * + U_nu(x-nu)^+ v(x-nu); v2 *= alpha on the slice; v += v2 everywhere (v2 is
 * zero off the slice); v *= 1/(1+6 alpha) on the slice.
 */
static void gauss_hop_site(const orc_geo *g, cplx acc[4][3], const double *v, const double *u, long i)
{
   int nu, s, c;
   for (s = 0; s < 4; s++)
      for (c = 0; c < 3; c++) acc[s][c].re = acc[s][c].im = 0.0;
   for (nu = 1; nu < 4; nu++) {
      cplx up[4][3], ud[4][3];
      long ip = g->plus[4*i + nu], im = g->minus[4*i + nu];
      link_mul(up, (const cplx (*)[3])(u + 72 * i + 18 * nu), (const cplx (*)[3])(v + 24 * ip), 4, 0);
      link_mul(ud, (const cplx (*)[3])(u + 72 * im + 18 * nu), (const cplx (*)[3])(v + 24 * im), 4, 1);
      for (s = 0; s < 4; s++)
         for (c = 0; c < 3; c++) {
            acc[s][c].re += up[s][c].re + ud[s][c].re;
            acc[s][c].im += up[s][c].im + ud[s][c].im;
         }
   }
}

void orc_gauss3d(const orc_geo *g, double *v, const double *u, double alpha, int t)
{
   const double nrm = 1.0 / (1.0 + alpha * 6.0);
   long V3 = g->V / g->L[0], j;
   double *v2 = (double *)malloc(sizeof(double) * 24 * V3);
   if (!v2) return;
#pragma omp parallel for schedule(static)
   for (j = 0; j < V3; j++) {
      cplx acc[4][3];
      gauss_hop_site(g, acc, v, u, j * g->L[0] + t);
      memcpy(v2 + 24 * j, acc, sizeof(acc));
   }
   for (j = 0; j < V3; j++) {
      double *vs = v + 24 * (j * g->L[0] + t), *hs = v2 + 24 * j;
      int k;
      for (k = 0; k < 24; k++) vs[k] = (vs[k] + hs[k] * alpha) * nrm;
   }
   free(v2);
}

/* lib/src/qcd_smearing.c:116-199: the same on every time-slice at once */
void orc_gauss3d_all(const orc_geo *g, double *v, const double *u, double alpha)
{
   const double nrm = 1.0 / (1.0 + alpha * 6.0);
   long i;
   double *v2 = (double *)malloc(sizeof(double) * 24 * g->V);
   if (!v2) return;
#pragma omp parallel for schedule(static)
   for (i = 0; i < g->V; i++) {
      cplx acc[4][3];
      gauss_hop_site(g, acc, v, u, i);
      memcpy(v2 + 24 * i, acc, sizeof(acc));
   }
   for (i = 0; i < 24 * g->V; i++) v[i] = (v[i] + v2[i] * alpha) * nrm;
   free(v2);
}

/* ------------------------------------------------- vector BLAS, qcd_blas.c
 * Elementwise loops run in parallel over sites.  Reductions follow the
 * reference run on ORC_RANKS MPI ranks: each rank sums its own contiguous
 * block of the vector sequentially (qcd_blas.c:456-491) and the rank sums are
 * combined by MPI_Allreduce (:473, :488), here added in rank order.  The
 * blocks are summed in parallel, and the result does not depend on the
 * number of threads. */
#define ORC_RANKS 64

void orc_scale(double *v, double r, long V)   /* :301-308 */
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 24 * V; i++) v[i] *= r;
}

void orc_sub(double *d, const double *a, const double *b, long V)   /* :345-354 */
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 24 * V; i++) d[i] = a[i] - b[i];
}

void orc_add(double *s, const double *a, const double *b, long V)   /* :403-412 */
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 24 * V; i++) s[i] = a[i] + b[i];
}

/* out = a*x + y (:414-432); safe when out aliases y */
void orc_axpy(double *out, double are, double aim, const double *x, const double *y, long V)
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 12 * V; i++) {
      double xr = x[2*i], xi = x[2*i + 1];
      out[2*i] = y[2*i] + xr * are - xi * aim;
      out[2*i + 1] = y[2*i + 1] + xi * are + xr * aim;
   }
}

void orc_mulc(double *v, double are, double aim, long V)   /* :237-245 */
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 12 * V; i++) {
      double re = v[2*i] * are - v[2*i + 1] * aim;
      v[2*i + 1] = v[2*i] * aim + v[2*i + 1] * are;
      v[2*i] = re;
   }
}

/* sum conj(a) b (:456-475): per-rank sequential sums, then the rank sums */
void orc_dot(double out[2], const double *a, const double *b, long V)
{
   double part[ORC_RANKS][2];
   const long n = 12 * V;
   int r;
#pragma omp parallel for schedule(static)
   for (r = 0; r < ORC_RANKS; r++) {
      long i;
      double re = 0.0, im = 0.0;
      for (i = n * r / ORC_RANKS; i < n * (r + 1) / ORC_RANKS; i++) {
         re += a[2*i] * b[2*i] + a[2*i + 1] * b[2*i + 1];
         im += a[2*i] * b[2*i + 1] - a[2*i + 1] * b[2*i];
      }
      part[r][0] = re;
      part[r][1] = im;
   }
   out[0] = out[1] = 0.0;
   for (r = 0; r < ORC_RANKS; r++) {
      out[0] += part[r][0];
      out[1] += part[r][1];
   }
}

/* sqrt(sum x^2) (:478-491): per-rank sequential sums, then the rank sums */
double orc_norm(const double *v, long V)
{
   double part[ORC_RANKS], s = 0.0;
   const long n = 24 * V;
   int r;
#pragma omp parallel for schedule(static)
   for (r = 0; r < ORC_RANKS; r++) {
      long i;
      double t = 0.0;
      for (i = n * r / ORC_RANKS; i < n * (r + 1) / ORC_RANKS; i++) t += v[i] * v[i];
      part[r] = t;
   }
   for (r = 0; r < ORC_RANKS; r++) s += part[r];
   return sqrt(s);
}

/* ------------------------------------------------------------------- solver */
static void apply_op(const orc_geo *g, int op, double *out, const double *in, const double *u, double kappa, double mu)
{
   double m = 1.0 / (2.0 * kappa) - 4.0;
   if (op == 0) orc_qtm(g, out, in, u, m, mu);
   else orc_dslash_half(g, out, in, u, m, mu);
}

/* invert.c:20-113 (algorithm spec gcr.m:3-48) */
int orc_gcr(const orc_geo *g, double *sol, const double *src, const double *u,
            double kappa, double mu, int n, int op, int keep_q1)
{
   const long V = g->V, N = 24 * V;
   double **chi = (double **)calloc(n, sizeof(double *));
   double **res = (double **)calloc(n, sizeof(double *));
   cplx *a = (cplx *)calloc((size_t)n * n, sizeof(cplx));
   cplx *c = (cplx *)calloc(n, sizeof(cplx));
   int i, j, k, l, rc = 0;
   double d[2];
   if (!chi || !res || !a || !c) { rc = 1; goto done; }
   for (i = 0; i < n; i++) {
      chi[i] = (double *)malloc(sizeof(double) * N);
      res[i] = (double *)malloc(sizeof(double) * N);
      if (!chi[i] || !res[i]) { rc = 1; goto done; }
   }
#define A(k_, j_) a[(k_) * n + (j_)]
   memcpy(res[0], src, sizeof(double) * N);
   apply_op(g, op, chi[0], res[0], u, kappa, mu);
   A(0, 0).re = 1.0 / orc_norm(chi[0], V);
   A(0, 0).im = 0.0;
   orc_scale(chi[0], A(0, 0).re, V);
   orc_dot(d, chi[0], src, V);
   c[0].re = d[0]; c[0].im = d[1];
   memset(sol, 0, sizeof(double) * N);
   if (!keep_q1) orc_axpy(sol, c[0].re * A(0, 0).re, c[0].im * A(0, 0).re, res[0], sol, V);
   for (k = 1; k < n; k++) {
      orc_axpy(res[k], -c[k - 1].re, -c[k - 1].im, chi[k - 1], res[k - 1], V);
      apply_op(g, op, chi[k], res[k], u, kappa, mu);
      for (j = 0; j < k; j++) {          /* modified Gram-Schmidt, invert.c:68-73 */
         orc_dot(d, chi[j], chi[k], V);
         A(k, j).re = -d[0];
         A(k, j).im = -d[1];
         orc_axpy(chi[k], A(k, j).re, A(k, j).im, chi[j], chi[k], V);
      }
      A(k, k).re = 1.0 / orc_norm(chi[k], V);
      A(k, k).im = 0.0;
      if (isnan(A(k, k).re)) { rc = 1; goto done; }
      orc_scale(chi[k], A(k, k).re, V);
      for (j = 0; j < k; j++) { A(k, j).re *= A(k, k).re; A(k, j).im *= A(k, k).re; }
      for (j = 0; j < k; j++) {           /* invert.c:87-94 */
         A(k, j).re *= A(j, j).re;
         A(k, j).im *= A(j, j).re;
         for (l = j + 1; l < k; l++) A(k, j) = cadd(A(k, j), cmul(A(k, l), A(l, j)));
      }
      orc_dot(d, chi[k], src, V);
      c[k].re = d[0]; c[k].im = d[1];
      for (j = 0; j <= k; j++) {          /* invert.c:100-103 */
         cplx f = cmul(c[k], A(k, j));
         orc_axpy(sol, f.re, f.im, res[j], sol, V);
      }
   }
#undef A
done:
   if (chi) for (i = 0; i < n; i++) free(chi[i]);
   if (res) for (i = 0; i < n; i++) free(res[i]);
   free(chi); free(res); free(a); free(c);
   return rc;
}

/* invert.c:255-294 with maxRes and maxIter as parameters */
int orc_invert(const orc_geo *g, double *x, const double *b, const double *u,
               double kappa, double mu, int n, int op, int keep_q1,
               double max_res, int max_outer, int *outer_out, double *relres_out)
{
   const long V = g->V, N = 24 * V;
   double *src = (double *)malloc(sizeof(double) * N);
   double *res = (double *)malloc(sizeof(double) * N);
   double *cor = (double *)malloc(sizeof(double) * N);
   double normsrc, normres;
   int iter = 1, rc = 0;
   if (!src || !res || !cor) { free(src); free(res); free(cor); return 1; }
   memcpy(src, b, sizeof(double) * N);
   if (op == 0) {          /* Q x = i gamma5 b is the same system as D x = b */
      long l;
      for (l = 0; l < V; l++) {
         int s, c;
         cplx *p = (cplx *)(src + 24 * l);
         for (s = 0; s < 4; s++)
            for (c = 0; c < 3; c++) p[3*s + c] = ipow(p[3*s + c], s < 2 ? 1 : 3);
      }
   }
   normsrc = orc_norm(src, V);
   if (orc_gcr(g, x, src, u, kappa, mu, n, op, keep_q1)) rc = 1;
   apply_op(g, op, res, x, u, kappa, mu);
   orc_sub(res, src, res, V);
   normres = orc_norm(res, V) / normsrc;
   while (!rc && normres > max_res && iter < max_outer) {
      if (orc_gcr(g, cor, res, u, kappa, mu, n, op, keep_q1)) { rc = 1; break; }
      orc_add(x, x, cor, V);
      apply_op(g, op, res, x, u, kappa, mu);
      orc_sub(res, src, res, V);
      normres = orc_norm(res, V) / normsrc;
      iter++;
   }
   if (outer_out) *outer_out = iter;
   if (relres_out) *relres_out = normres;
   free(src); free(res); free(cor);
   if (rc) return rc;
   return normres <= max_res ? 0 : 2;
}

double orc_b_minus_Dx(const orc_geo *g, const double *x, const double *b, const double *u,
                      double kappa, double mu)
{
   const long V = g->V;
   double *dx = (double *)malloc(sizeof(double) * 24 * V), r;
   if (!dx) return -1.0;
   orc_dslash_half(g, dx, x, u, 1.0 / (2.0 * kappa) - 4.0, mu);
   orc_sub(dx, dx, b, V);
   r = orc_norm(dx, V) / orc_norm(b, V);
   free(dx);
   return r;
}

/* ------------------------------------------------------- synthetic inputs
 * Counter-based generator: draw k of stream s under seed is
 *   mix(key(seed, s) + k * 0x632BE59BD9B4E019), key = mix(seed) ^ s * 0xD1B54A32D192ED03
 * with mix the splitmix64 finaliser; uniforms in (0,1) from the top 53 bits;
 * normals by Box-Muller.  The product library implements the same recipe.
 */
static inline uint64_t mix64(uint64_t z)
{
   z += 0x9E3779B97F4A7C15ULL;
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
   return z ^ (z >> 31);
}

static inline double uniform01(uint64_t key, uint64_t k)
{
   uint64_t r = mix64(key + k * 0x632BE59BD9B4E019ULL);
   return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static void normals(double *out, int n, uint64_t seed, uint64_t stream)
{
   uint64_t key = mix64(seed) ^ (stream * 0xD1B54A32D192ED03ULL);
   int i;
   for (i = 0; i < n; i += 2) {
      double u1 = uniform01(key, (uint64_t)i), u2 = uniform01(key, (uint64_t)i + 1);
      double r = sqrt(-2.0 * log(u1));
      out[i] = r * cos(6.283185307179586 * u2);
      out[i + 1] = r * sin(6.283185307179586 * u2);
   }
}

static void mat_mul3(cplx c[3][3], const cplx a[3][3], const cplx b[3][3])
{
   int i, j, k;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
         cplx s = {0.0, 0.0};
         for (k = 0; k < 3; k++) s = cadd(s, cmul(a[i][k], b[k][j]));
         c[i][j] = s;
      }
}

/* U = reunitarize(exp(i eps H)), H = (A + A^+)/2 made traceless,
 * A_ij = N(0,1) + i N(0,1).  exp by Taylor series (degree 16) on X/2^s with
 * |X/2^s|_inf <= 1/4, then s squarings; reunitarize by Gram-Schmidt of rows
 * 0,1 and row2 = conj(row0 x row1). */
static void make_link(cplx U[3][3], const double g[18], double eps)
{
   cplx H[3][3], X[3][3], E[3][3], T[3][3];
   double tr, nrm = 0.0, scale = 1.0;
   int i, j, k, s = 0;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
         H[i][j].re = 0.5 * (g[2*(3*i + j)] + g[2*(3*j + i)]);
         H[i][j].im = 0.5 * (g[2*(3*i + j) + 1] - g[2*(3*j + i) + 1]);
      }
   tr = (H[0][0].re + H[1][1].re + H[2][2].re) / 3.0;
   for (i = 0; i < 3; i++) { H[i][i].re -= tr; H[i][i].im = 0.0; }
   for (i = 0; i < 3; i++) {
      double row = 0.0;
      for (j = 0; j < 3; j++) {
         X[i][j].re = -eps * H[i][j].im;   /* X = i eps H */
         X[i][j].im = eps * H[i][j].re;
         row += fabs(X[i][j].re) + fabs(X[i][j].im);
      }
      if (row > nrm) nrm = row;
   }
   while (nrm * scale > 0.25) { scale *= 0.5; s++; }
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) { X[i][j].re *= scale; X[i][j].im *= scale; }
   /* Horner: E = I + X/1 (I + X/2 (I + ... (I + X/16))) */
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) { E[i][j].re = (i == j) ? 1.0 : 0.0; E[i][j].im = 0.0; }
   for (k = 16; k >= 1; k--) {
      mat_mul3(T, X, E);
      for (i = 0; i < 3; i++)
         for (j = 0; j < 3; j++) {
            E[i][j].re = T[i][j].re / (double)k + ((i == j) ? 1.0 : 0.0);
            E[i][j].im = T[i][j].im / (double)k;
         }
   }
   for (k = 0; k < s; k++) {
      mat_mul3(T, E, E);
      memcpy(E, T, sizeof(E));
   }
   {
      double n0 = 0.0, n1 = 0.0;
      cplx p = {0.0, 0.0};
      for (j = 0; j < 3; j++) n0 += E[0][j].re * E[0][j].re + E[0][j].im * E[0][j].im;
      n0 = sqrt(n0);
      for (j = 0; j < 3; j++) { U[0][j].re = E[0][j].re / n0; U[0][j].im = E[0][j].im / n0; }
      for (j = 0; j < 3; j++) p = cadd(p, cmul(conjc(U[0][j]), E[1][j]));
      for (j = 0; j < 3; j++) {
         cplx q = cmul(p, U[0][j]);
         U[1][j].re = E[1][j].re - q.re;
         U[1][j].im = E[1][j].im - q.im;
      }
      for (j = 0; j < 3; j++) n1 += U[1][j].re * U[1][j].re + U[1][j].im * U[1][j].im;
      n1 = sqrt(n1);
      for (j = 0; j < 3; j++) { U[1][j].re /= n1; U[1][j].im /= n1; }
      for (j = 0; j < 3; j++) {
         int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
         cplx x = cmul(U[0][j1], U[1][j2]), y = cmul(U[0][j2], U[1][j1]);
         U[2][j].re = x.re - y.re;
         U[2][j].im = -(x.im - y.im);
      }
   }
}

void orc_gauge_generate(double *u, long V, double eps, uint64_t seed)
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < 4 * V; i++) {
      double g[18];
      normals(g, 18, seed, (uint64_t)i);
      make_link((cplx (*)[3])(u + 18 * i), g, eps);
   }
}

void orc_spinor_gaussian(double *v, long V, uint64_t seed)
{
   long i;
#pragma omp parallel for schedule(static)
   for (i = 0; i < V; i++) normals(v + 24 * i, 24, seed ^ 0x5350494E4F52ULL, (uint64_t)i);
}

void orc_point_source(double *v, const int L[4], int t, int x, int y, int z, int spin, int col)
{
   long V = (long)L[0] * L[1] * L[2] * L[3];
   long l = t + (long)L[0] * (x + (long)L[1] * (y + (long)L[2] * z));
   memset(v, 0, sizeof(double) * 24 * V);
   v[24 * l + 6 * spin + 2 * col] = 1.0;
}

/* lib/src/qcd_observables.c:15-36 */
double orc_plaquette(const orc_geo *g, const double *u)
{
   double sum = 0.0;
   long l;
   for (l = 0; l < g->V; l++) {
      int mu, nu, i;
      for (mu = 0; mu < 3; mu++)
         for (nu = mu + 1; nu < 4; nu++) {
            const cplx (*U1)[3] = (const cplx (*)[3])(u + 72 * l + 18 * mu);
            const cplx (*U2)[3] = (const cplx (*)[3])(u + 72 * g->plus[4*l + mu] + 18 * nu);
            const cplx (*U3)[3] = (const cplx (*)[3])(u + 72 * g->plus[4*l + nu] + 18 * mu);
            const cplx (*U4)[3] = (const cplx (*)[3])(u + 72 * l + 18 * nu);
            cplx P[3][3], T[3][3], A3[3][3], A4[3][3];
            int r, c;
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) { A3[r][c] = conjc(U3[c][r]); A4[r][c] = conjc(U4[c][r]); }
            mat_mul3(P, U1, U2);
            mat_mul3(T, P, A3);
            mat_mul3(P, T, A4);
            for (i = 0; i < 3; i++) sum += P[i][i].re;
         }
   }
   return sum / ((double)g->V * 3.0 * 6.0);
}


/* ------------------------------------------------------------ APE smearing */

static cplx cdivc(cplx a, cplx b)
{
   const double d = b.re * b.re + b.im * b.im;
   cplx r = {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
   return r;
}

static cplx det3(const cplx a[3][3])
{
   cplx d = {0.0, 0.0};
   int c;
   for (c = 0; c < 3; c++) {   /* expansion along row 0, cyclic minors carry their sign */
      cplx t = cmul(a[1][(c + 1) % 3], a[2][(c + 2) % 3]);
      const cplx s = cmul(a[1][(c + 2) % 3], a[2][(c + 1) % 3]);
      t.re -= s.re;
      t.im -= s.im;
      d = cadd(d, cmul(a[0][c], t));
   }
   return d;
}

/* (a^-1)^+ by the adjugate: inv[r][c] = cof[c][r] / det, so
 * (a^-1)^+[r][c] = conj(cof[r][c] / det) */
static void inv_dagger3(cplx out[3][3], const cplx a[3][3])
{
   const cplx d = det3(a);
   int r, c;
   for (r = 0; r < 3; r++)
      for (c = 0; c < 3; c++) {
         const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
         cplx cof = cmul(a[r1][c1], a[r2][c2]);
         const cplx s = cmul(a[r1][c2], a[r2][c1]);
         cof.re -= s.re;
         cof.im -= s.im;
         out[r][c] = conjc(cdivc(cof, d));
      }
}

static double frob2(const cplx a[3][3])
{
   double s = 0.0;
   int r, c;
   for (r = 0; r < 3; r++)
      for (c = 0; c < 3; c++) s += a[r][c].re * a[r][c].re + a[r][c].im * a[r][c].im;
   return s;
}

/* The reference (qcd_blas.c:950-1364) diagonalises H^2 = M^+ M by Cardano's
 * formula and forms M H^-1 exp(-i arg(det M)/3).  The polar factor is unique
 * for non-singular M, so it is computed here by Newton's iteration
 * X <- (g X + X^-+ / g) / 2 with Frobenius-norm scaling g (Higham 1986),
 * which converges quadratically from X = M. */
void orc_project_su3(double *m18)
{
   cplx M[3][3], X[3][3], Y[3][3];
   int r, c, it;
   memcpy(M, m18, sizeof(M));
   memcpy(X, M, sizeof(X));
   for (it = 0; it < 100; it++) {
      double g = 1.0, diff = 0.0;
      inv_dagger3(Y, X);
      if (it < 8) g = sqrt(sqrt(frob2(Y) / frob2(X)));
      for (r = 0; r < 3; r++)
         for (c = 0; c < 3; c++) {
            cplx n = {0.5 * (g * X[r][c].re + Y[r][c].re / g), 0.5 * (g * X[r][c].im + Y[r][c].im / g)};
            diff += (n.re - X[r][c].re) * (n.re - X[r][c].re) + (n.im - X[r][c].im) * (n.im - X[r][c].im);
            X[r][c] = n;
         }
      if (it >= 8 && diff <= 1e-30) break;
   }
   {
      const cplx d = det3(M);
      const double ph = atan2(d.im, d.re) / 3.0;
      const cplx e = {cos(ph), -sin(ph)};
      for (r = 0; r < 3; r++)
         for (c = 0; c < 3; c++) X[r][c] = cmul(X[r][c], e);
   }
   memcpy(m18, X, sizeof(X));
}

/* apeu_mu(x) = P_SU3[ U_mu(x) + alpha sum_{nu != mu, spatial} (forward staple +
 * backward staple) ] for mu = 1, 2, 3; apeu_0 = U_0 (qcd_smearing.c:205-273).
 * Staples are summed in the reference's order: all forward ones
 * U_nu(x) U_mu(x+nu) U_nu(x+mu)^+, then all backward ones
 * U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu). */
void orc_ape3d(const orc_geo *g, double *out, const double *u, double alpha)
{
   long l;
#pragma omp parallel for schedule(static)
   for (l = 0; l < g->V; l++) {
      int mu, nu, r, c;
      memcpy(out + 72 * l, u + 72 * l, 18 * sizeof(double));
      for (mu = 1; mu < 4; mu++) {
         cplx S[3][3], T[3][3], P[3][3], A[3][3], M[3][3];
         memset(S, 0, sizeof(S));
         for (nu = 1; nu < 4; nu++) {
            if (nu == mu) continue;
            const cplx (*Un)[3] = (const cplx (*)[3])(u + 72 * l + 18 * nu);
            const cplx (*Um)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * l + nu] + 18 * mu);
            const cplx (*Ua)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * l + mu] + 18 * nu);
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) A[r][c] = conjc(Ua[c][r]);
            mat_mul3(T, Un, Um);
            mat_mul3(P, T, A);
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) S[r][c] = cadd(S[r][c], P[r][c]);
         }
         for (nu = 1; nu < 4; nu++) {
            if (nu == mu) continue;
            const long lm = g->minus[4 * l + nu];
            const cplx (*Un)[3] = (const cplx (*)[3])(u + 72 * lm + 18 * nu);
            const cplx (*Um)[3] = (const cplx (*)[3])(u + 72 * lm + 18 * mu);
            const cplx (*Ua)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * lm + mu] + 18 * nu);
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) A[r][c] = conjc(Un[c][r]);
            mat_mul3(T, Um, Ua);
            mat_mul3(P, A, T);
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) S[r][c] = cadd(S[r][c], P[r][c]);
         }
         {
            const cplx (*U0)[3] = (const cplx (*)[3])(u + 72 * l + 18 * mu);
            for (r = 0; r < 3; r++)
               for (c = 0; c < 3; c++) {
                  M[r][c].re = U0[r][c].re + alpha * S[r][c].re;
                  M[r][c].im = U0[r][c].im + alpha * S[r][c].im;
               }
         }
         orc_project_su3((double *)M);
         memcpy(out + 72 * l + 18 * mu, M, sizeof(M));
      }
   }
}

/* ---------------------------------------------------------------- stout */
/* sin(w)/w with the reference's series below |w| = 0.05 (qcd_smearing.c:279-290) */
static double stout_xi0(double w)
{
   if (fabs(w) < 0.05) {
      double t = 1.0 - w * w / 42.0;
      double t1 = 1.0 - w * w / 20.0 * t;
      return 1.0 - w * w / 6.0 * t1;
   }
   return sin(w) / w;
}

/* exp(iQ) for Hermitian traceless Q in place: Cayley-Hamilton coefficients of
 * Peardon & Morningstar (hep-lat/0311018 eqs. 19, 29-34) as in
 * qcd_smearing.c:293-370.  Two differences, both where the reference returns
 * NaN: for tr(Q^2)/2 < 1e-24 it divides 0 by 0 (acos(NaN)); here exp(iQ) =
 * 1 + iQ - Q^2/2, exact to double precision at that size.  And for a
 * degenerate spectrum (two equal eigenvalues, theta = 0) rounding can put
 * c0 / c0max above 1; here the argument of acos is clamped to 1. */
void orc_stout_exp(double *q)
{
   cplx (*M)[3] = (cplx (*)[3])q;
   cplx M2[3][3], M3[3][3];
   int i, j;
   double c0, c1, c0max, theta, u, w, xi, fac;
   cplx e2iu, emiu, h0, h1, h2, f0, f1, f2, t, t1;
   int neg = 0;
   mat_mul3(M2, M, M);
   mat_mul3(M3, M2, M);
   c0 = (M3[0][0].re + M3[1][1].re + M3[2][2].re) / 3.0;
   c1 = (M2[0][0].re + M2[1][1].re + M2[2][2].re) / 2.0;
   if (c1 < 1e-24) {
      for (i = 0; i < 3; i++)
         for (j = 0; j < 3; j++) {
            cplx r;
            r.re = (i == j ? 1.0 : 0.0) - M[i][j].im - 0.5 * M2[i][j].re;
            r.im = M[i][j].re - 0.5 * M2[i][j].im;
            M[i][j] = r;
         }
      return;
   }
   if (c0 < 0) { neg = 1; c0 = -c0; }
   c0max = 2.0 * pow(c1 / 3.0, 1.5);
   theta = acos(c0 < c0max ? c0 / c0max : 1.0);   /* the reference: acos(c0 / c0max), NaN when rounding puts it above 1 */
   u = sqrt(c1 / 3.0) * cos(theta / 3.0);
   w = sqrt(c1) * sin(theta / 3.0);
   xi = stout_xi0(w);
   e2iu.re = cos(2 * u); e2iu.im = sin(2 * u);
   emiu.re = cos(u); emiu.im = -sin(u);
   /* h0 = (u^2 - w^2) e^{2iu} + e^{-iu} (8 u^2 cos w + 2iu (3u^2 + w^2) xi0)   eq. (30) */
   t.re = u * u - w * w; t.im = 0.0;
   t1.re = 8 * u * u * cos(w); t1.im = 2 * u * (3 * u * u + w * w) * xi;
   h0 = cadd(cmul(t, e2iu), cmul(emiu, t1));
   /* h1 = 2u e^{2iu} - e^{-iu} (2u cos w - i (3u^2 - w^2) xi0)   eq. (31) */
   t.re = 2 * u; t.im = 0.0;
   t1.re = 2 * u * cos(w); t1.im = -(3 * u * u - w * w) * xi;
   {
      cplx a = cmul(t, e2iu), b = cmul(emiu, t1);
      h1.re = a.re - b.re; h1.im = a.im - b.im;
   }
   /* h2 = e^{2iu} - e^{-iu} (cos w + 3iu xi0)   eq. (32) */
   t.re = cos(w); t.im = 3 * u * xi;
   {
      cplx b = cmul(emiu, t);
      h2.re = e2iu.re - b.re; h2.im = e2iu.im - b.im;
   }
   fac = 9 * u * u - w * w;   /* eq. (29) */
   if (!neg) {
      f0.re = h0.re / fac; f0.im = h0.im / fac;
      f1.re = h1.re / fac; f1.im = h1.im / fac;
      f2.re = h2.re / fac; f2.im = h2.im / fac;
   } else {   /* eq. (34): f_j(-c0) = (-1)^j conj(f_j(c0)) */
      f0.re = h0.re / fac; f0.im = -h0.im / fac;
      f1.re = -h1.re / fac; f1.im = h1.im / fac;
      f2.re = h2.re / fac; f2.im = -h2.im / fac;
   }
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
         cplx r = cadd(cmul(f1, M[i][j]), cmul(f2, M2[i][j]));
         if (i == j) r = cadd(r, f0);
         M[i][j] = r;
      }
}

/* One stout step (qcd_smearing.c:374-476): for the smeared directions mu
 * (1..3 when ndim = 3, 0..3 when ndim = 4; with ndim = 3 the temporal links
 * are copied), C = sum over nu != mu of the same range of the backward staple
 * U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) then the forward staple
 * U_nu(x) U_mu(x+nu) U_nu(x+mu)^+; W = C U_mu(x)^+; A = W - W^+ minus tr(A)/3;
 * Q = -i rho/2 A (Hermitian, traceless); U'_mu(x) = exp(iQ) U_mu(x).
 * out must not alias u; returns 1 for ndim other than 3 or 4. */
int orc_stout(const orc_geo *g, double *out, const double *u, double rho, int ndim)
{
   long l;
   int start;
   if (ndim != 3 && ndim != 4) return 1;
   start = ndim == 3 ? 1 : 0;
#pragma omp parallel for schedule(static)
   for (l = 0; l < g->V; l++) {
      int mu, nu, r, c;
      memcpy(out + 72 * l, u + 72 * l, 18 * sizeof(double));
      for (mu = start; mu < 4; mu++) {
         cplx S[3][3], T[3][3], P[3][3], A[3][3], W[3][3], Q[3][3], E[3][3];
         const cplx (*U0)[3] = (const cplx (*)[3])(u + 72 * l + 18 * mu);
         cplx tr;
         memset(S, 0, sizeof(S));
         for (nu = start; nu < 4; nu++) {
            if (nu == mu) continue;
            {  /* backward: U_nu(x-nu)^+ [U_mu(x-nu) U_nu(x-nu+mu)] */
               const long lm = g->minus[4 * l + nu];
               const cplx (*Un)[3] = (const cplx (*)[3])(u + 72 * lm + 18 * nu);
               const cplx (*Um)[3] = (const cplx (*)[3])(u + 72 * lm + 18 * mu);
               const cplx (*Ua)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * lm + mu] + 18 * nu);
               for (r = 0; r < 3; r++)
                  for (c = 0; c < 3; c++) A[r][c] = conjc(Un[c][r]);
               mat_mul3(T, Um, Ua);
               mat_mul3(P, A, T);
               for (r = 0; r < 3; r++)
                  for (c = 0; c < 3; c++) S[r][c] = cadd(S[r][c], P[r][c]);
            }
            {  /* forward: [U_nu(x) U_mu(x+nu)] U_nu(x+mu)^+ */
               const cplx (*Un)[3] = (const cplx (*)[3])(u + 72 * l + 18 * nu);
               const cplx (*Um)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * l + nu] + 18 * mu);
               const cplx (*Ua)[3] = (const cplx (*)[3])(u + 72 * g->plus[4 * l + mu] + 18 * nu);
               for (r = 0; r < 3; r++)
                  for (c = 0; c < 3; c++) A[r][c] = conjc(Ua[c][r]);
               mat_mul3(T, Un, Um);
               mat_mul3(P, T, A);
               for (r = 0; r < 3; r++)
                  for (c = 0; c < 3; c++) S[r][c] = cadd(S[r][c], P[r][c]);
            }
         }
         for (r = 0; r < 3; r++)
            for (c = 0; c < 3; c++) A[r][c] = conjc(U0[c][r]);
         mat_mul3(W, S, A);
         for (r = 0; r < 3; r++)
            for (c = 0; c < 3; c++) {
               P[r][c].re = W[r][c].re - W[c][r].re;
               P[r][c].im = W[r][c].im + W[c][r].im;
            }
         tr.re = (P[0][0].re + P[1][1].re + P[2][2].re) / 3.0;
         tr.im = (P[0][0].im + P[1][1].im + P[2][2].im) / 3.0;
         for (r = 0; r < 3; r++) { P[r][r].re -= tr.re; P[r][r].im -= tr.im; }
         for (r = 0; r < 3; r++)   /* Q = -i (rho/2) P */
            for (c = 0; c < 3; c++) {
               Q[r][c].re = 0.5 * rho * P[r][c].im;
               Q[r][c].im = -0.5 * rho * P[r][c].re;
            }
         orc_stout_exp((double *)Q);
         mat_mul3(E, Q, U0);
         memcpy(out + 72 * l + 18 * mu, E, sizeof(E));
      }
   }
   return 0;
}
