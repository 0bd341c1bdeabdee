/* qcd_io.h -- file formats of the MI355X library's drop-in API.
 *
 * Counterpart of the reference's include/qcd_io.h:12-44 (implementations in
 * lib/src/qcd_io.c) for the path's inputs and outputs, on one rank: parameter
 * files, ILDG/LIME gauge fields, LIME spinors and HMC propagators.  The LIME
 * record layout is read and written directly (no c-lime library).  Host code
 * only.
 */
#ifndef H_QCD_IO
#define H_QCD_IO 1

#ifdef __cplusplus
extern "C" {
#endif

/* include/qcd_io.h:12-17 */
#define qcd_GF_LIME 0
#define qcd_PROP_CMI 0
#define qcd_PROP_HMC 1
#define qcd_PROP_HMCV 2
#define qcd_PROP_LIME 3
#define qcd_SOURCE_LIME 4

/* byte order (qcd_io.c:20-75) */
int  qcd_isBigEndian(void);
void qcd_swap_8(double *Rd, int N);
void qcd_swap_4(float *Rd, int N);
/* parameter files (qcd_io.c:83-128): the whole file, and the text after
 * "<tag>" up to the next '<' (which is overwritten with a NUL) */
char *qcd_getParams(char *fname, int *len);
char *qcd_getParam(char token[], char *params, int len);
/* ILDG gauge fields in LIME files (qcd_io.c:132-603): 64- or 32-bit reads,
 * 64-bit writes; the message is the "xlf-info" record */
int qcd_getLimeMessage(char *fname, qcd_geometry *geo, char **message_ptr);
int qcd_getGaugeLime(char *fname, qcd_gaugeField *u);
int qcd_getGaugeField(char *fname, int type, qcd_gaugeField *u);
int qcd_writeGaugeLime(char *fname, qcd_gaugeField *u, char *message);
/* spinors (qcd_io.c:984-1414, :1766-1974): LIME (ETMC records) and the HMC
 * propagator format (vector 3 nu + c2 of 12) */
int qcd_getVectorLime(char *fname, qcd_vector *v);
int qcd_getVectorHMC(char *fname, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v);
int qcd_getVector(char *fname, int type, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v);
int qcd_writeVectorLime(char *fname, int type, qcd_vector *v);
int qcd_writeVector(char *fname, int type, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v);
/* Writes the 12 vectors vecs[3 nu + c2] as an HMC propagator file, the format
 * qcd_getVectorHMC reads.  No reference counterpart (the reference only reads
 * this format); used by the invert driver to store its solutions. */
int qcd_writePropagatorHMC(char *fname, qcd_vector *vecs, qcd_real_8 beta, qcd_real_8 kappa, qcd_real_8 mu);

#ifdef __cplusplus
}
#endif

#endif
