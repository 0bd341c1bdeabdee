// helpers.cpp -- the reference's small host helpers: per-site spinor operations
// of qcd_wilson.h, the 3x3 matrix helpers and the gauge-field / propagator
// arithmetic of qcd_blas.h, with the reference's names and argument meaning.
// Host code on host structs; the whole-field SU(3) projection runs on the GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

extern "C" {
#include "qcd.h"
#include "qcd_blas.h"
#include "qcd_hip.h"
#include "qcd_wilson.h"
}

namespace qcdhip {
int dropin_gauge_written(const void *host);
}

namespace {

typedef qcd_complex_16 M3[3][3];

// the unitary polar factor of m, m (m^+ m)^-1/2, times exp(-i arg(det m) / 3):
// Newton's iteration x <- (g x + x^-+ / g) / 2 with Frobenius scaling g, from
// x = m (lib/src/qcd_blas.c:1376-1769 reaches the same matrix by a Cardano
// eigen-decomposition of m^+ m)
struct Cd {
  double re, im;
};
inline Cd cmul(Cd a, Cd b) { return Cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cd csub(Cd a, Cd b) { return Cd{a.re - b.re, a.im - b.im}; }

Cd det3(const Cd a[9]) {
  Cd d{0.0, 0.0};
  for (int c = 0; c < 3; c++) {
    const Cd minor = csub(cmul(a[3 + (c + 1) % 3], a[6 + (c + 2) % 3]), cmul(a[3 + (c + 2) % 3], a[6 + (c + 1) % 3]));
    const Cd t = cmul(a[c], minor);
    d = Cd{d.re + t.re, d.im + t.im};
  }
  return d;
}

void polar_su3(Cd m[9]) {
  Cd x[9];
  for (int k = 0; k < 9; k++) x[k] = m[k];
  for (int it = 0; it < 100; it++) {
    const Cd d = det3(x);
    const double rdd = 1.0 / (d.re * d.re + d.im * d.im);
    Cd y[9];
    double nx = 0.0, ny = 0.0;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const Cd cof = csub(cmul(x[3 * r1 + c1], x[3 * r2 + c2]), cmul(x[3 * r1 + c2], x[3 * r2 + c1]));
        const Cd q{(cof.re * d.re + cof.im * d.im) * rdd, (cof.im * d.re - cof.re * d.im) * rdd};   // cof / det
        y[3 * r + c] = Cd{q.re, -q.im};
        nx += x[3 * r + c].re * x[3 * r + c].re + x[3 * r + c].im * x[3 * r + c].im;
        ny += q.re * q.re + q.im * q.im;
      }
    const double g = it < 8 ? std::sqrt(std::sqrt(ny / nx)) : 1.0, rg = 1.0 / g;
    double diff = 0.0;
    for (int k = 0; k < 9; k++) {
      const Cd n{0.5 * (g * x[k].re + y[k].re * rg), 0.5 * (g * x[k].im + y[k].im * rg)};
      diff += (n.re - x[k].re) * (n.re - x[k].re) + (n.im - x[k].im) * (n.im - x[k].im);
      x[k] = n;
    }
    if (diff <= 1e-30) break;
  }
  const Cd d = det3(m);
  const double ph = std::atan2(d.im, d.re) / 3.0;
  const Cd e{std::cos(ph), -std::sin(ph)};
  for (int k = 0; k < 9; k++) m[k] = cmul(x[k], e);
}

}  // namespace

extern "C" {

// ------------------------------------------------ per-site spinor helpers
// lib/src/qcd_wilson.c:18-103 (qcd_wilson.h:627-632)
void qcd_onePlusGamma(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_uint_2 mu) {
  qcd_project_(&result[0][0].re, &vec[0][0].re, mu, 1, 1);
}

void qcd_oneMinusGamma(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_uint_2 mu) {
  qcd_project_(&result[0][0].re, &vec[0][0].re, mu, -1, 1);
}

void qcd_applyU(qcd_complex_16 result[][3], qcd_complex_16 u[][3], qcd_complex_16 vec[][3]) {
  qcd_link_apply_(&u[0][0].re, &result[0][0].re, &vec[0][0].re, 4, 0);
}

void qcd_applyUDagger(qcd_complex_16 result[][3], qcd_complex_16 u[][3], qcd_complex_16 vec[][3]) {
  qcd_link_apply_(&u[0][0].re, &result[0][0].re, &vec[0][0].re, 4, 1);
}

void qcd_massTerms(qcd_complex_16 result[][3], qcd_complex_16 vec[][3], qcd_real_8 mp4, qcd_real_8 tm) {
  qcd_mass_terms_(&result[0][0].re, &vec[0][0].re, mp4, tm);
}

void qcd_phaseMul(qcd_complex_16 vec[][3], qcd_complex_16 phi) { qcd_phase_mul_(&vec[0][0].re, 12, phi.re, phi.im); }

// ------------------------------------------------ 3x3 matrices
// lib/src/qcd_blas.c:19-135
void qcd_zero3x3(qcd_complex_16 a[3][3]) { std::memset(a, 0, sizeof(M3)); }

void qcd_unit3x3(qcd_complex_16 a[3][3]) {
  std::memset(a, 0, sizeof(M3));
  for (int i = 0; i < 3; i++) a[i][i].re = 1.0;
}

void qcd_minusUnit3x3(qcd_complex_16 a[3][3]) {
  std::memset(a, 0, sizeof(M3));
  for (int i = 0; i < 3; i++) a[i][i].re = -1.0;
}

void qcd_sub3x3(qcd_complex_16 difference[3][3], qcd_complex_16 minuend[3][3], qcd_complex_16 subtrahend[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) difference[i][j] = qcd_CSUB(minuend[i][j], subtrahend[i][j]);
}

void qcd_add3x3(qcd_complex_16 sum[3][3], qcd_complex_16 summand1[3][3], qcd_complex_16 summand2[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sum[i][j] = qcd_CADD(summand1[i][j], summand2[i][j]);
}

// sum = summand1 + summand2^+
void qcd_addAdjoint3x3(qcd_complex_16 sum[3][3], qcd_complex_16 summand1[3][3], qcd_complex_16 summand2[3][3]) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = qcd_CADD(summand1[i][j], qcd_CONJ(summand2[j][i]));
  std::memcpy(sum, t, sizeof(M3));
}

void qcd_mul3x3(qcd_complex_16 product[3][3], qcd_complex_16 factor1[3][3], qcd_complex_16 factor2[3][3]) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      qcd_complex_16 s = {0.0, 0.0};
      for (int k = 0; k < 3; k++) s = qcd_CADD(s, qcd_CMUL(factor1[i][k], factor2[k][j]));
      t[i][j] = s;
    }
  std::memcpy(product, t, sizeof(M3));
}

// product = factor1 factor2^+
void qcd_mulAdjoint3x3(qcd_complex_16 product[3][3], qcd_complex_16 factor1[3][3], qcd_complex_16 factor2[3][3]) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      qcd_complex_16 s = {0.0, 0.0};
      for (int k = 0; k < 3; k++) s = qcd_CADD(s, qcd_CMUL(factor1[i][k], qcd_CONJ(factor2[j][k])));
      t[i][j] = s;
    }
  std::memcpy(product, t, sizeof(M3));
}

void qcd_scale3x3(qcd_complex_16 a[3][3], qcd_real_8 r) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = qcd_CSCALE(a[i][j], r);
}

void qcd_cScale3x3(qcd_complex_16 a[3][3], qcd_complex_16 c) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = qcd_CMUL(a[i][j], c);
}

qcd_complex_16 qcd_trace3x3(qcd_complex_16 a[3][3]) {
  qcd_complex_16 r = {a[0][0].re + a[1][1].re + a[2][2].re, a[0][0].im + a[1][1].im + a[2][2].im};
  return r;
}

void qcd_copy3x3(qcd_complex_16 dest[3][3], qcd_complex_16 src[3][3]) { std::memmove(dest, src, sizeof(M3)); }

void qcd_dagger3x3(qcd_complex_16 a[3][3]) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = qcd_CONJ(a[j][i]);
  std::memcpy(a, t, sizeof(M3));
}

// SU(3) projection of one matrix in place (lib/src/qcd_blas.c:1376-1769)
void qcd_projectSU33x3(qcd_complex_16 g[3][3]) {
  Cd m[9];
  for (int k = 0; k < 9; k++) m[k] = Cd{g[k / 3][k % 3].re, g[k / 3][k % 3].im};
  polar_su3(m);
  for (int k = 0; k < 9; k++) g[k / 3][k % 3] = qcd_complex_16{m[k].re, m[k].im};
}

// ------------------------------------------------ propagators
// lib/src/qcd_blas.c:192-200, :261-285, :516-530
void qcd_copyPropagatorPropagator(qcd_propagator *prop_new, qcd_propagator *prop_old) {
  std::memcpy(prop_new->D, prop_old->D, 144 * prop_new->geo->lV * sizeof(qcd_complex_16));
}

void qcd_mulPropagatorC(qcd_propagator *prop, qcd_complex_16 c) {
  qcd_complex_16 *p = &prop->D[0][0][0][0][0];
  for (qcd_uint_8 i = 0; i < 144 * prop->geo->lV; i++) p[i] = qcd_CMUL(p[i], c);
}

// time-slice t only; site = i * lL[0] + t (the reference's lexicographic convention)
void qcd_mulPropagatorC3d(qcd_propagator *prop, qcd_complex_16 c, qcd_uint_4 t) {
  for (qcd_uint_8 i = 0; i < prop->geo->lV3; i++) {
    qcd_complex_16 *p = &prop->D[i * prop->geo->lL[0] + t][0][0][0][0];
    for (int k = 0; k < 144; k++) p[k] = qcd_CMUL(p[k], c);
  }
}

void qcd_conjPropagator(qcd_propagator *prop) {
  qcd_complex_16 *p = &prop->D[0][0][0][0][0];
  for (qcd_uint_8 i = 0; i < 144 * prop->geo->lV; i++) p[i] = qcd_CONJ(p[i]);
}

// ------------------------------------------------ gauge fields
// lib/src/qcd_blas.c:358-390, :495-514; results written into a bound
// (resident) field refresh its device copy
void qcd_subGaugeField(qcd_gaugeField *difference, qcd_gaugeField *minuend, qcd_gaugeField *subtrahend) {
  qcd_complex_16 *d = &difference->D[0][0][0][0];
  const qcd_complex_16 *a = &minuend->D[0][0][0][0], *b = &subtrahend->D[0][0][0][0];
  for (qcd_uint_8 i = 0; i < 36 * difference->geo->lV; i++) d[i] = qcd_CSUB(a[i], b[i]);
  qcdhip::dropin_gauge_written(difference->D);
}

void qcd_addGaugeField(qcd_gaugeField *sum, qcd_gaugeField *summand1, qcd_gaugeField *summand2) {
  qcd_complex_16 *d = &sum->D[0][0][0][0];
  const qcd_complex_16 *a = &summand1->D[0][0][0][0], *b = &summand2->D[0][0][0][0];
  for (qcd_uint_8 i = 0; i < 36 * sum->geo->lV; i++) d[i] = qcd_CADD(a[i], b[i]);
  qcdhip::dropin_gauge_written(sum->D);
}

void qcd_scaleGaugeField(qcd_gaugeField *u, qcd_real_8 alpha) {
  qcd_complex_16 *d = &u->D[0][0][0][0];
  for (qcd_uint_8 i = 0; i < 36 * u->geo->lV; i++) d[i] = qcd_CSCALE(d[i], alpha);
  qcdhip::dropin_gauge_written(u->D);
}

// sum over links of |tr U| (the reference's "norm"; one rank, so no Allreduce)
qcd_real_8 qcd_normGaugeField(qcd_gaugeField *u) {
  double s = 0.0;
  for (qcd_uint_8 i = 0; i < u->geo->lV; i++)
    for (int mu = 0; mu < 4; mu++) {
      qcd_complex_16 tr = {0.0, 0.0};
      for (int a = 0; a < 3; a++) tr = qcd_CADD(tr, u->D[i][mu][a][a]);
      s += qcd_NORM(tr);
    }
  return s;
}

// random links: entries drand48() (re, then im; x, mu, c1, c2 order), then
// SU(3)-projected (lib/src/qcd_blas.c:1771-1782)
void qcd_randomGaugeField(qcd_gaugeField *u) {
  for (qcd_uint_8 x = 0; x < u->geo->lV; x++)
    for (int mu = 0; mu < 4; mu++)
      for (int c1 = 0; c1 < 3; c1++)
        for (int c2 = 0; c2 < 3; c2++) {
          const double re = drand48();
          const double im = drand48();
          u->D[x][mu][c1][c2] = qcd_complex_16{re, im};
        }
  qcd_projectSU3(u);
}

}  // extern "C"
