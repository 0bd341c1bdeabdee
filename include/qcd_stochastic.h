/* qcd_stochastic.h -- noise vectors (include/qcd_stochastic.h:5,
 * lib/src/qcd_stochastic.c:8-58). */
#ifndef QCD_HIP_STOCHASTIC_H
#define QCD_HIP_STOCHASTIC_H 1

#ifdef __cplusplus
extern "C" {
#endif

/* Z4 noise ({+1, +i, -1, -i}, from qcd_get_rand_4) on every site of time
 * slice t_slice, zero elsewhere; t_slice < 0 fills every slice.  Draws in the
 * reference's loop order (z, y, x, spin, colour; for t_slice < 0 also t
 * outermost) from the stream of qcd_rng_init, and prints the same
 * "time dilution" / "full in time" line.  Host code; vec must be initialized. */
void qcd_z2Vector(qcd_vector *vec, int t_slice);

#ifdef __cplusplus
}
#endif

#endif
