/* qcd_rng.h -- the random-number interface of the reference
 * (include/qcd_rng.h:1-10, lib/src/qcd_rng.c:1-33).
 *
 * Same functions and the same per-process seed derivation as the reference
 * (srand(seed), then rand() once per rank below myid).  The generator itself
 * differs: the reference draws from GSL's ranlux389, which is not part of
 * this library; here every draw is the splitmix64 finaliser of a 64-bit
 * counter keyed by the process seed.  Streams are reproducible for a seed,
 * but their numbers differ from the reference's. */
#ifndef QCD_HIP_RNG_H
#define QCD_HIP_RNG_H 1

#ifdef __cplusplus
extern "C" {
#endif

/* start the process's stream (qcd_rng.c:5-17) */
void qcd_rng_init(unsigned int seed, int myid);
/* release it (qcd_rng.c:19-24) */
void qcd_rng_finalize(void);
/* uniform integer in {0, 1, 2, 3} (qcd_rng.c:26-31) */
unsigned long int qcd_get_rand_4(void);

#ifdef __cplusplus
}
#endif

#endif
