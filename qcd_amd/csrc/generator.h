// generator.h -- the synthetic-input recipe, shared by host and device code.
//
// Counter-based draws: draw k of stream s under seed is
//   mix(key + k * 0x632BE59BD9B4E019), key = mix(seed) ^ s * 0xD1B54A32D192ED03,
// mix = the splitmix64 finaliser; uniforms in (0,1) from the top 53 bits;
// normals by Box-Muller.  Links U = reunitarize(exp(i eps H)),
// H = (A + A^+)/2 made traceless with A_ij = N(0,1) + i N(0,1); exp by a
// degree-16 Taylor polynomial of X / 2^s (|X/2^s|_inf <= 1/4) and s squarings;
// reunitarisation by Gram-Schmidt of rows 0 and 1, row 2 = conj(row0 x row1).
// Host builds call the C library's log/cos/sin, so host fields are the same
// bits on every machine with the same libm; device builds use the device
// math library (same recipe, last-bit differences).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace qcdgen {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ inline double unit_draw(uint64_t key, uint64_t k) {
  uint64_t r = mix64(key + k * 0x632BE59BD9B4E019ULL);
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline void gaussians(double *out, int n, uint64_t seed, uint64_t stream) {
  const uint64_t key = mix64(seed) ^ (stream * 0xD1B54A32D192ED03ULL);
  for (int i = 0; i < n; i += 2) {
    double u1 = unit_draw(key, (uint64_t)i), u2 = unit_draw(key, (uint64_t)i + 1);
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(6.283185307179586 * u2);
    out[i + 1] = r * sin(6.283185307179586 * u2);
  }
}

struct C {
  double re, im;
};
__host__ __device__ inline C cm(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline C ca(C a, C b) { return {a.re + b.re, a.im + b.im}; }

__host__ __device__ inline void m3mul(C r[3][3], const C a[3][3], const C b[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C s = {0.0, 0.0};
      for (int k = 0; k < 3; k++) s = ca(s, cm(a[i][k], b[k][j]));
      r[i][j] = s;
    }
}

__host__ __device__ inline void su3_link(C U[3][3], const double g[18], double eps) {
  C H[3][3], X[3][3], E[3][3], T[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      H[i][j].re = 0.5 * (g[2 * (3 * i + j)] + g[2 * (3 * j + i)]);
      H[i][j].im = 0.5 * (g[2 * (3 * i + j) + 1] - g[2 * (3 * j + i) + 1]);
    }
  double tr = (H[0][0].re + H[1][1].re + H[2][2].re) / 3.0;
  for (int i = 0; i < 3; i++) {
    H[i][i].re -= tr;
    H[i][i].im = 0.0;
  }
  double nrm = 0.0, scale = 1.0;
  int s = 0;
  for (int i = 0; i < 3; i++) {
    double row = 0.0;
    for (int j = 0; j < 3; j++) {
      X[i][j].re = -eps * H[i][j].im;
      X[i][j].im = eps * H[i][j].re;
      row += fabs(X[i][j].re) + fabs(X[i][j].im);
    }
    nrm = row > nrm ? row : nrm;
  }
  while (nrm * scale > 0.25) {
    scale *= 0.5;
    s++;
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      X[i][j].re *= scale;
      X[i][j].im *= scale;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) E[i][j] = {i == j ? 1.0 : 0.0, 0.0};
  for (int k = 16; k >= 1; k--) {
    m3mul(T, X, E);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        E[i][j].re = T[i][j].re / (double)k + (i == j ? 1.0 : 0.0);
        E[i][j].im = T[i][j].im / (double)k;
      }
  }
  for (int k = 0; k < s; k++) {
    m3mul(T, E, E);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) E[i][j] = T[i][j];
  }
  double n0 = 0.0, n1 = 0.0;
  for (int j = 0; j < 3; j++) n0 += E[0][j].re * E[0][j].re + E[0][j].im * E[0][j].im;
  n0 = sqrt(n0);
  for (int j = 0; j < 3; j++) U[0][j] = {E[0][j].re / n0, E[0][j].im / n0};
  C p = {0.0, 0.0};
  for (int j = 0; j < 3; j++) p = ca(p, cm({U[0][j].re, -U[0][j].im}, E[1][j]));
  for (int j = 0; j < 3; j++) {
    C q = cm(p, U[0][j]);
    U[1][j] = {E[1][j].re - q.re, E[1][j].im - q.im};
  }
  for (int j = 0; j < 3; j++) n1 += U[1][j].re * U[1][j].re + U[1][j].im * U[1][j].im;
  n1 = sqrt(n1);
  for (int j = 0; j < 3; j++) U[1][j] = {U[1][j].re / n1, U[1][j].im / n1};
  for (int j = 0; j < 3; j++) {
    int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    C x = cm(U[0][j1], U[1][j2]), y = cm(U[0][j2], U[1][j1]);
    U[2][j] = {x.re - y.re, -(x.im - y.im)};
  }
}

constexpr uint64_t kSpinorSeedSalt = 0x5350494E4F52ULL;

}  // namespace qcdgen
