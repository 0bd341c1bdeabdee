/* qcd.h -- core types and field layouts of the MI355X lattice-QCD library.
 *
 * Drop-in counterpart of the reference's include/qcd.h (include/qcd.h:19-198):
 * the same type names, the same struct member names and order, and the same
 * host data layouts, so that C callers written against the reference compile
 * unchanged against this header.
 *
 * Layouts (all complex128, array-of-structures, t fastest):
 *   site index        l = t + L0*(x + L1*(y + L2*z))          (include/qcd.h:144)
 *   qcd_vector        D[site][spin 4][colour 3]                (include/qcd.h:90-97)
 *   qcd_gaugeField    D[site][mu 4][row 3][col 3]              (include/qcd.h:119-126)
 *                     U_mu(x) is the link from x to x+mu, mu=0 is time.
 *   qcd_propagator    D[site][sink spin][src spin][sink col][src col]
 *                                                              (include/qcd.h:110-117)
 *
 * This library runs one rank on one GPU.  The reference's geometry embeds
 * MPI request and datatype handles (include/qcd.h:75-84); here they are
 * single-rank placeholders, so sizeof(qcd_geometry) differs from an MPI build
 * and callers must be compiled against this header.
 */
#ifndef H_QCD
#define H_QCD 1

#ifdef __cplusplus
extern "C" {
#endif

#define qcd_MAX_STRING_LENGTH 256

typedef float          qcd_real_4;
typedef double         qcd_real_8;
typedef char           qcd_int_1;
typedef short          qcd_int_2;
typedef int            qcd_int_4;
typedef long           qcd_int_8;
typedef unsigned char  qcd_uint_1;
typedef unsigned short qcd_uint_2;
typedef unsigned int   qcd_uint_4;
typedef unsigned long  qcd_uint_8;

typedef struct { qcd_real_8 re; qcd_real_8 im; } qcd_complex_16;
typedef struct { qcd_real_4 re; qcd_real_4 im; } qcd_complex_8;

/* Single-rank stand-ins for the MPI handle types carried by the geometry. */
typedef int qcd_request_slot;
typedef int qcd_datatype_slot;

typedef struct
{
   qcd_uint_8 V;                  /* global 4d volume                          */
   qcd_uint_8 V3;                 /* global spatial volume                     */
   qcd_uint_2 L[4];               /* lattice extents t,x,y,z                   */
   qcd_uint_8 lV;                 /* local volume (== V on one rank)           */
   qcd_uint_8 lV3;                /* local spatial volume                      */
   qcd_uint_2 lL[4];              /* local extents (== L on one rank)          */
   qcd_uint_2 Pos[4];             /* block position in the process grid (0)    */
   qcd_real_8 theta[4];           /* psi(x+L_mu) = exp(i theta_mu) psi(x)      */
   qcd_uint_4 myid;
   qcd_uint_4 nproc;
   qcd_uint_4 Pplus[4];
   qcd_uint_4 Pminus[4];
   qcd_uint_8 (*plus)[4];         /* periodic +mu neighbour of each site       */
   qcd_uint_8 (*minus)[4];        /* periodic -mu neighbour of each site       */
   qcd_uint_8 (*plus3d)[4];       /* same inside one time-slice (index 1..3)   */
   qcd_uint_8 (*minus3d)[4];
   qcd_uint_8 (*eo)[2];           /* even/odd site lists                       */
   qcd_uint_8 startBminus[4];
   qcd_uint_8 startBplus[4];
   qcd_uint_8 startBminus3d[4];
   qcd_uint_8 startBplus3d[4];
   qcd_uint_8 boundarySize;       /* 0 on one rank: no halo                    */
   qcd_uint_8 boundarySize3d;
   qcd_uint_8 *edge;              /* sites on the surface of the local block   */
   qcd_uint_8 edgePoints;
   qcd_uint_8 *edge0;             /* surface sites of one time-slice (3d index)*/
   qcd_uint_8 edge0Points;

   qcd_request_slot requests[32];
   int numOfRequests;
   qcd_datatype_slot stypeV[4];
   qcd_datatype_slot rtypeV[4];
   qcd_datatype_slot stypeU[4];
   qcd_datatype_slot rtypeU[4];
   qcd_datatype_slot stypeT[4];
   qcd_datatype_slot rtypeT[4];
   qcd_datatype_slot stypeP[4];
   qcd_datatype_slot rtypeP[4];

   qcd_uint_2 initialized;
} qcd_geometry;

typedef struct
{
   qcd_complex_16 (*D)[4][3];
   qcd_complex_16 (*Bplus[4])[4][3];
   qcd_complex_16 (*Bminus[4])[4][3];
   qcd_geometry *geo;
   qcd_uint_2 initialized;
} qcd_vector;

typedef struct
{
   qcd_complex_16 (*D)[4][3];
   qcd_complex_16 (*Bplus[4])[4][3];
   qcd_complex_16 (*Bminus[4])[4][3];
   qcd_geometry *geo;
   qcd_uint_2 initialized;
} qcd_vector3d;

typedef struct
{
   qcd_complex_16 (*D)[4][4][3][3];
   qcd_complex_16 (*Bplus[4])[4][4][3][3];
   qcd_complex_16 (*Bminus[4])[4][4][3][3];
   qcd_geometry *geo;
   qcd_uint_2 initialized;
} qcd_propagator;

typedef struct
{
   qcd_complex_16 (*D)[4][3][3];
   qcd_complex_16 (*Bplus[4])[4][3][3];
   qcd_complex_16 (*Bminus[4])[4][3][3];
   qcd_geometry *geo;
   qcd_uint_2 initialized;
} qcd_gaugeField;

/* Site-index helpers (include/qcd.h:144-151): t is the fastest index. */
#define qcd_LEXIC(t,x,y,z,L) ( (qcd_uint_8) ((t)+(L)[0]*((x)+(L)[1]*((y)+(L)[2]*(z)))) )
#define qcd_LEXIC0(x,y,z,L)  ( (qcd_uint_8) ((x)+(L)[1]*((y)+(L)[2]*(z))) )
#define qcd_LEXIC1(t,y,z,L)  ( (qcd_uint_8) ((t)+(L)[0]*((y)+(L)[2]*(z))) )
#define qcd_LEXIC2(t,x,z,L)  ( (qcd_uint_8) ((t)+(L)[0]*((x)+(L)[1]*(z))) )
#define qcd_LEXIC3(t,x,y,L)  ( (qcd_uint_8) ((t)+(L)[0]*((x)+(L)[1]*(y))) )
#define qcd_LEXIC01(y,z,L)   ( (qcd_uint_8) ((y)+(L)[2]*(z)) )
#define qcd_LEXIC02(x,z,L)   ( (qcd_uint_8) ((x)+(L)[1]*(z)) )
#define qcd_LEXIC03(x,y,L)   ( (qcd_uint_8) ((x)+(L)[1]*(y)) )

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

/* Lifecycle (include/qcd.h:186-198, lib/src/qcd_init.c:16-665).
 * qcd_initGeometry accepts only the single-rank grid P={1,1,1,1}, numprocs==1;
 * every other call returns 1 with a message on stderr. */
void qcd_antilexic(qcd_uint_2 x[], qcd_uint_8 l, const qcd_uint_2 dim[]);
int  qcd_initGeometry(qcd_geometry *geo, const qcd_uint_2 L[], const qcd_uint_2 P[],
                      const qcd_real_8 theta[], const int myid, const int numprocs);
void qcd_destroyGeometry(const qcd_geometry *geo);
int  qcd_initVector(qcd_vector *vec, qcd_geometry *geo);
void qcd_destroyVector(qcd_vector *vec);
int  qcd_initVector3d(qcd_vector3d *vec, qcd_geometry *geo);
void qcd_destroyVector3d(qcd_vector3d *vec);
int  qcd_initPropagator(qcd_propagator *prp, qcd_geometry *geo);
void qcd_destroyPropagator(qcd_propagator *prp);
int  qcd_initGaugeField(qcd_gaugeField *u, qcd_geometry *geo);
void qcd_destroyGaugeField(qcd_gaugeField *u);

/* Halo exchange (include/qcd_communication.h:12-21).  With one rank there is
 * nothing to exchange: these succeed and do nothing. */
int  qcd_communicateVectorPMGaugeP(qcd_vector *v, qcd_gaugeField *u);
int  qcd_communicateVectorPM(qcd_vector *v);
void qcd_waitall(qcd_geometry *geo);

#ifdef __cplusplus
}
#endif

#include <qcd_blas.h>
#include <qcd_wilson.h>
#include <qcd_smearing.h>
#include <qcd_io.h>
#include <qcd_observables.h>
#include <qcd_rng.h>
#include <qcd_stochastic.h>

#endif
