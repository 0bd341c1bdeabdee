/* qcd_hip.h -- device-resident extension API of the MI355X lattice-QCD library.
 *
 * The drop-in entry points (qcd_wilson.h, qcd_smearing.h, qcd_blas.h) take
 * host structs and copy data in and out on every call, as the reference's
 * signatures require.  This header is the fast path: opaque handles whose data
 * stay in HBM across calls (gauge fields uploaded once, solver state resident).
 * It has no counterpart in the reference; every entry point below names the
 * reference routine whose semantics it implements.
 *
 * Conventions: functions return 0 on success and non-zero on failure, with a
 * message available from qcd_hip_last_error().  Host arrays use the
 * reference's AoS layouts (include/qcd.h:90-126) as interleaved doubles.
 * All work is ordered on one library stream per process; calls are not
 * thread-safe.  One GPU, one rank.
 */
#ifndef QCD_HIP_H
#define QCD_HIP_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qcd_hip_lattice_s *qcd_hip_lattice;
typedef struct qcd_hip_field_s *qcd_hip_field;

enum { QCD_HIP_F64 = 0, QCD_HIP_F32 = 1, QCD_HIP_F16 = 2 };
enum { QCD_HIP_SPINOR = 0, QCD_HIP_GAUGE = 1 };
/* operator selector: D (lib/src/qcd_wilson.c:114), Q = i gamma5 D (:795) */
enum { QCD_HIP_OP_D = 0, QCD_HIP_OP_Q = 1 };

/* ------------------------------------------------------------ runtime */
const char *qcd_hip_last_error(void);
int    qcd_hip_init(int device);
int    qcd_hip_device_name(char *buf, int len);
int    qcd_hip_synchronize(void);
size_t qcd_hip_bytes_in_use(void);
/* Device blocks freed by the library are cached for reuse (hipFree would
 * synchronise the device); this hands the cache back to the runtime. */
void   qcd_hip_release_cache(void);

/* ------------------------------------------------------------ lattice
 * L = {T, X, Y, Z}; theta are the fermion boundary phases
 * psi(x + L_mu) = e^{i theta_mu} psi(x), applied as e^{+-i theta_mu/L_mu} on
 * every hop in direction mu (lib/src/qcd_wilson.c:172-186). */
int  qcd_hip_lattice_create(qcd_hip_lattice *lat, const int L[4], const double theta[4]);
void qcd_hip_lattice_destroy(qcd_hip_lattice lat);
long qcd_hip_lattice_volume(qcd_hip_lattice lat);

/* ------------------------------------------------------------ fields */
int  qcd_hip_spinor_create(qcd_hip_field *f, qcd_hip_lattice lat, int prec, int nrhs);
int  qcd_hip_gauge_create(qcd_hip_field *f, qcd_hip_lattice lat, int prec);
void qcd_hip_field_destroy(qcd_hip_field f);
/* spinor rhs <-> host double[V][4][3][2] (qcd_vector.D) */
int  qcd_hip_spinor_upload(qcd_hip_field f, int rhs, const double *aos);
int  qcd_hip_spinor_download(qcd_hip_field f, int rhs, double *aos);
/* gauge <-> host double[V][4][3][3][2] (qcd_gaugeField.D) */
int  qcd_hip_gauge_upload(qcd_hip_field f, const double *aos);
int  qcd_hip_gauge_download(qcd_hip_field f, double *aos);
/* host propagator double[V][4 mu][4 nu][3 c1][3 c2][2] (qcd_propagator.D,
 * include/qcd.h:110-117) <-> 12 rhs of a spinor field: rhs r0 + 3 nu + c2
 * holds the column (nu, c2), i.e. the spinor psi[mu][c1] = P[mu][nu][c1][c2]
 * of qcd_copyVectorPropagator (lib/src/qcd_blas.c:182-212), for all 12
 * columns in one pass on the device. */
int  qcd_hip_propagator_upload(qcd_hip_field f, int r0, const double *prop);
int  qcd_hip_propagator_download(qcd_hip_field f, int r0, double *prop);
/* copy with precision change; same lattice, same kind (and nrhs for spinors) */
int  qcd_hip_field_convert(qcd_hip_field dst, qcd_hip_field src);

/* Synthetic inputs, host side (identical bits on every machine with the same
 * libm): U = reunitarize(exp(i eps H)) per link, H a traceless Hermitian
 * matrix from seeded N(0,1) draws; spinor entries i.i.d. N(0,1). */
int qcd_hip_gauge_generate_host(double *aos, long V, double eps, uint64_t seed);
int qcd_hip_spinor_gaussian_host(double *aos, long V, uint64_t seed);
/* The same recipe run on the GPU straight into an FP64 gauge field (device
 * math library: equal to the host field to rounding, not bit for bit). */
int qcd_hip_gauge_generate(qcd_hip_field gauge64, double eps, uint64_t seed);
/* rhs of an FP64 spinor field = delta at site {t,x,y,z} on (spin, col)
 * (source.c:203-208) */
int qcd_hip_point_source(qcd_hip_field f, int rhs, const int site[4], int spin, int col);
/* Z4 noise ({+1, +i, -1, -i}) in vector rhs of an FP64 spinor on time slice
 * t_slice (all slices when t_slice < 0), zero elsewhere, on the device.  The
 * value of (site, spin, colour) is a counter-based draw under seed, so it
 * does not depend on launch order; the stream differs from the host one of
 * qcd_z2Vector (lib/src/qcd_stochastic.c:8-58). */
int qcd_hip_z4_source(qcd_hip_field f, int rhs, int t_slice, uint64_t seed);

/* ------------------------------------------------------------ operators
 * out = D in (op = QCD_HIP_OP_D) or Q in = i gamma5 D in (QCD_HIP_OP_Q),
 * with m the mass parameter (m + 4 = 1/(2 kappa)) and tm the twisted mass;
 * every rhs of in is applied.  out must not alias in.
 * (lib/src/qcd_wilson.c:114-336, :570-792, :795-814) */
int qcd_hip_dslash(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge,
                   double m, double tm, int op);
/* Times `iters` back-to-back applications (after `warmup`) with HIP events
 * on the library stream; returns the mean per application in *avg_ms. */
int qcd_hip_dslash_time(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge,
                        double m, double tm, int op, int warmup, int iters, double *avg_ms);

/* Autotuner: time every traversal (0 = one contiguous range per XCD,
 * 1 = one y-slab per XCD swept along z, 2 | (n << 8) = each lane marches n
 * steps along z carrying the backward z hop in registers, 3 | (n << 8) = the
 * same over tiles of whole t-rows staged in LDS with their x halo rows,
 * 4 | (n << 8) = traversal 2 with the forward z neighbour kept in LDS as the
 * next step's centre, 5 | (n << 8) = traversal 2 with two rhs per lane for
 * FP32/FP16 batches with the D or Q epilogue, 6 = FP16 batches on the
 * matrix cores: MFMA applies each link to the spin-projected neighbour
 * spinors of 4 rhs per wavefront of 16 sites) and occupancy variant (0 = compiler default, 2 or 3 =
 * at least that many waves per SIMD, 16 = scheduling fences between hop
 * directions) on these fields and keep the fastest for their (lattice
 * extents, precisions).  Runs by itself on first use unless the environment
 * sets QCD_HIP_AUTOTUNE=0.  All candidates give bit-identical results. */
int qcd_hip_dslash_tune(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, int tuned[2], double *ms);
int qcd_hip_dslash_set_config(qcd_hip_field in, qcd_hip_field gauge, int map, int occ);
/* The traversal candidates (map values above) the autotuner times: fills up
 * to max entries of maps and returns how many there are. */
int qcd_hip_dslash_candidates(int *maps, int max);

/* ------------------------------------------------------------ smearing
 * niter iterations of 3D Gaussian smearing of every rhs of v on time-slice t
 * (t < 0: every time-slice), in place (lib/src/qcd_smearing.c:25-199).
 * v and gauge must have the same precision.  The iterations run on two
 * library-owned buffers of the slice (the volume when t < 0) per rhs, as many
 * rhs per pass as fit in 4 GiB and at least one (QCD_HIP_GAUSS_ROWS lowers
 * the count), plus one copy of the slice's links; both stay allocated for
 * reuse until the process ends. */
int qcd_hip_gauss_smear(qcd_hip_field v, qcd_hip_field gauge, double alpha, int t, int niter);

/* ------------------------------------------------------------ BLAS (FP64
 * accumulation, deterministic reduction order for a given lattice size) */
int qcd_hip_norm2(qcd_hip_field v, int rhs, double *out);
int qcd_hip_dot(qcd_hip_field a, int ra, qcd_hip_field b, int rb, double out[2]);
/* y[ry] += a x[rx] (lib/src/qcd_blas.c:414-432; y may alias x) */
int qcd_hip_axpy(qcd_hip_field y, int ry, double are, double aim, qcd_hip_field x, int rx);
/* out[ro] = a x[rx] + b y[ry] on all sites, or on time-slice t0 only when
 * t0 >= 0 (lib/src/qcd_blas.c:237-434: scale, scale3d, add, sub, add3d,
 * mulVectorC are special cases); y may be NULL for b = 0; out may alias x, y */
int qcd_hip_lincomb(qcd_hip_field out, int ro, double are, double aim, qcd_hip_field x, int rx,
                    double bre, double bim, qcd_hip_field y, int ry, int t0);

/* ------------------------------------------------------------ solver
 * Propagator solve of invert.c on the GPU: an FP64 outer loop on the true
 * residual (invert.c:255-294) around GCR(n_restart) (invert.c:20-113) run in
 * the inner precision: QCD_HIP_F64 is the plain solver, F32 / F16 (storage,
 * FP32 arithmetic) give mixed-precision defect correction.  Solves D x = b
 * for every rhs of b (with op = Q it solves Q x = i g5 b, the same system, as
 * invert.c does) until ||b - D x|| / ||b|| <= tol, checked in FP64. */
typedef struct qcd_hip_solver_s *qcd_hip_solver;
typedef struct {
   int    n_restart;   /* GCR steps per cycle (invert.c <N_restart>), 1..64 */
   int    max_outer;   /* outer cycles (invert.c maxIter) */
   double tol;         /* target relative residual (invert.c maxRes) */
   int    inner_prec;  /* QCD_HIP_F64, QCD_HIP_F32 or QCD_HIP_F16 */
   int    op;          /* QCD_HIP_OP_Q (as invert.c) or QCD_HIP_OP_D */
   int    keep_q1;     /* 1: drop the first GCR direction from the cycle's
                          solution exactly as invert.c:50-53 does */
   int    batch;       /* right-hand sides solved together (1 = one at a
                          time as invert.c); workspace grows with it */
   int    eo;          /* 1: even-odd preconditioning: the inner GCR solves
                          the Schur complement D^ = M - H_oe M^-1 H_eo on odd
                          sites (half-length vectors) and the outer loop the
                          full D x = b; needs every extent even; op is then D */
   int    graph;       /* 1 (default): each GCR cycle is captured once as a
                          HIP graph and replayed (one launch instead of
                          about 7 per GCR step); 0: launched kernel by kernel.
                          Both run the same kernels in the same order. */
} qcd_hip_solver_params;
typedef struct {
   int    outer;        /* outer cycles (summed over rhs) */
   int    inner_iters;  /* GCR iterations (summed over rhs) */
   double relres;       /* largest final ||b - D x|| / ||b|| over rhs (FP64) */
   double seconds;      /* wall time of the call */
} qcd_hip_solver_stats;

void qcd_hip_solver_default_params(qcd_hip_solver_params *p);
int  qcd_hip_solver_create(qcd_hip_solver *s, qcd_hip_lattice lat, const qcd_hip_solver_params *p);
/* Device bytes qcd_hip_solver_create plus qcd_hip_solver_set_gauge allocate
 * for a lattice of extents L with parameters p (0 for invalid arguments);
 * the serving loop admits requests by this figure. */
size_t qcd_hip_solver_bytes(const int L[4], const qcd_hip_solver_params *p);
void qcd_hip_solver_destroy(qcd_hip_solver s);
/* set (or refresh after re-upload) the FP64 gauge field; builds the inner-
 * precision copy */
int  qcd_hip_solver_set_gauge(qcd_hip_solver s, qcd_hip_field gauge64);
/* x, b: FP64 spinors with equal rhs counts; x is overwritten.  Returns 0 when
 * every rhs converged, 5 on Krylov breakdown, 6 when max_outer ran out. */
int  qcd_hip_solver_solve(qcd_hip_solver s, qcd_hip_field x, qcd_hip_field b, double kappa, double mu,
                          qcd_hip_solver_stats *stats);
int  qcd_hip_invert(qcd_hip_field x, qcd_hip_field b, qcd_hip_field gauge64, double kappa, double mu,
                    const qcd_hip_solver_params *p, qcd_hip_solver_stats *stats);
/* ||D x[rx] - b[rb]|| / ||b[rb]|| in FP64, m = 1/(2 kappa) - 4 (b_minus_Dx.c:152-167) */
int  qcd_hip_b_minus_Dx(qcd_hip_field x, int rx, qcd_hip_field b, int rb, qcd_hip_field gauge64,
                        double kappa, double mu, double *rel);

/* The same residual without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The caller passes an FP64
 * field r that receives b[rb] - D x[rx] in vector rr; the workgroup partial
 * sums live in a static 2.0 MiB device array of the library (one per GPU,
 * held until the process ends), so lattices whose Dslash launch has more than
 * 2^18 workgroups are rejected with an error.  It uses the tuned launch
 * configuration if there is one (no tuning here), synchronises the library
 * stream to return *rel, and calls on one GPU that overlap race on the array.
 * The library must be initialised (any field creation does that). */
int  qcd_hip_b_minus_Dx_noalloc(qcd_hip_field x, int rx, qcd_hip_field b, int rb, qcd_hip_field gauge64,
                                double kappa, double mu, qcd_hip_field r, int rr, double *rel);

/* nsteps of 3d APE smearing (lib/src/qcd_smearing.c:205-273) of the FP64
 * gauge field in into out (distinct fields on one lattice): spatial links ->
 * P_SU3[U + alpha * spatial staples], temporal links copied.  nsteps = 0
 * copies; alpha = 0 projects the spatial links (qcd_blas.c:940-1364). */
int qcd_hip_ape_smear(qcd_hip_field out, qcd_hip_field in, double alpha, int nsteps);

/* nsteps of stout smearing (lib/src/qcd_smearing.c:374-476) of the FP64
 * gauge field in into out (distinct fields on one lattice), ndim 3 (spatial,
 * temporal links copied) or 4; nsteps = 0 copies.  exp(iQ) by the
 * Peardon-Morningstar formula (qcd_smearing.c:293-370), exact at Q = 0. */
int qcd_hip_stout_smear(qcd_hip_field out, qcd_hip_field in, double rho, int ndim, int nsteps);

/* every link of an FP64 gauge field projected onto SU(3) in place
 * (lib/src/qcd_blas.c:541-939; the unitary polar factor times
 * exp(-i arg(det) / 3)) */
int qcd_hip_project_su3(qcd_hip_field gauge64);

/* mean plaquette of an FP64 gauge field (lib/src/qcd_observables.c:15-36) */
int qcd_hip_plaquette(qcd_hip_field gauge64, double *plaq);

/* Even-odd Schur complement D^ = M - H_oe M^-1 H_eo (M = m + 4 + i tm g5, H
 * the hopping term of D) applied to every rhs of in: in and out hold the odd
 * sites of a lattice {T, X, Y, Z} as fields of its half lattice
 * {T/2, X, Y, Z} (odd site l at entry l >> 1); gauge is the full-lattice
 * field.  No reference counterpart (the reference has no even-odd form). */
int qcd_hip_dslash_schur(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, double m, double tm);

/* z steps per lane of the even-odd hops' z-marching kernel, for every later
 * even-odd hop of the process: 0 = chosen per launch (the default, or the
 * value of QCD_HIP_EO_ZLEN), 1 = one output site per lane, n > 1 = z chunks
 * of n planes.  Every choice gives bit-identical results.  No reference
 * counterpart. */
int qcd_hip_eo_set_zlen(int zlen);

/* ------------------------------------------------------------ serving
 * Continuous serving of propagator-solve requests (no reference
 * counterpart; each request is the invert.c solve of its own configuration).
 * Requests arrive at arrival_s seconds after the start (realtime != 0) or all
 * at once (realtime == 0).  `workers` host threads, each with its own streams,
 * take arrived requests in arrival order whenever a worker is free and the
 * request's device footprint fits next to the running ones in
 * `budget_bytes`; running requests overlap on the GPU.  A request generates
 * its gauge field on the device (qcd_hip_gauge_generate recipe), puts nrhs
 * point sources at the origin (rhs r: spin r/3, colour r%3) and solves them as
 * one batch to tol, verified by the FP64 true residual.  solutions[i], when
 * not NULL, receives request i's solutions as nrhs host AoS spinors. */
typedef struct {
   double   arrival_s;
   int      L[4];
   double   kappa;
   double   eps;         /* gauge field U = reunitarize(exp(i eps H)) */
   uint64_t seed;
   int      nrhs;        /* 1 or 12 point sources */
   int      inner_prec;  /* QCD_HIP_F64 plain, QCD_HIP_F32 / F16 mixed */
   int      n_restart;
   double   tol;
} qcd_hip_request;
typedef struct {
   double queued_s;      /* arrival to start of service */
   double service_s;     /* start of service to completion */
   double latency_s;     /* arrival to completion */
   int    rc;            /* 0 = every rhs converged */
   int    outer;
   int    iters;
   double relres;        /* largest FP64 true relative residual over rhs */
   double bytes;         /* device bytes reserved for the request */
   double xnorm2;        /* sum over rhs of |x|^2 (a fingerprint) */
} qcd_hip_result;
int qcd_hip_serve(const qcd_hip_request *req, int n, int workers, size_t budget_bytes, int realtime,
                  qcd_hip_result *res, double *wall_s, double *const *solutions);

#ifdef __cplusplus
}
#endif

#endif
