// dropin.cpp -- the reference's C API on top of the device library.
//
// Same names, signatures, return codes and error messages as the reference
// (include/qcd.h:186-198, include/qcd_wilson.h:634-637,
// include/qcd_smearing.h:39-40, include/qcd_blas.h:242-267).  Host structs stay
// authoritative: each arithmetic call copies its operands to HBM, runs the
// gfx950 kernels and copies the result back.  A gauge field registered with
// qcd_hip_bindGaugeField() stays resident and is not re-uploaded.
// Pure host data movement (zero / set / copy / propagator columns) stays on
// the host, as a memcpy.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "blas_common.h"

extern "C" {
#include "qcd.h"
}

using namespace qcdhip;

namespace {

// -------------------------------------------------------- device cache
struct DropinCache {
  qcd_hip_lattice lat = nullptr;
  int L[4] = {0, 0, 0, 0};
  double theta[4] = {0, 0, 0, 0};
  qcd_hip_field gauge = nullptr;      // scratch F64 gauge
  const void *bound_host = nullptr;   // host data pointer of the bound gauge field
  qcd_hip_field bound = nullptr;      // resident copy of it
};
DropinCache g_dc;

int lattice_for(const qcd_geometry *geo, qcd_hip_lattice *out) {
  int L[4] = {geo->L[0], geo->L[1], geo->L[2], geo->L[3]};
  bool same = g_dc.lat != nullptr;
  for (int d = 0; d < 4 && same; d++) same = g_dc.L[d] == L[d] && g_dc.theta[d] == geo->theta[d];
  if (!same) {
    if (g_dc.gauge && g_dc.lat) {
      qcd_hip_field_destroy(g_dc.gauge);
      g_dc.gauge = nullptr;
    }
    // the previous lattice object may still be referenced by a bound field
    if (g_dc.lat && !(g_dc.bound && g_dc.bound->lat == g_dc.lat)) qcd_hip_lattice_destroy(g_dc.lat);
    g_dc.lat = nullptr;
    QH_TRY(qcd_hip_lattice_create(&g_dc.lat, L, geo->theta));
    for (int d = 0; d < 4; d++) {
      g_dc.L[d] = L[d];
      g_dc.theta[d] = geo->theta[d];
    }
  }
  *out = g_dc.lat;
  return 0;
}

int gauge_for(qcd_gaugeField *u, qcd_hip_lattice lat, qcd_hip_field *out) {
  if (g_dc.bound && g_dc.bound_host == (const void *)u->D && g_dc.bound->lat->V == lat->V) {
    *out = g_dc.bound;
    return 0;
  }
  if (!g_dc.gauge || g_dc.gauge->lat->V != lat->V) {
    if (g_dc.gauge) qcd_hip_field_destroy(g_dc.gauge);
    g_dc.gauge = nullptr;
    QH_TRY(qcd_hip_gauge_create(&g_dc.gauge, lat, QCD_HIP_F64));
  }
  g_dc.gauge->lat = lat;
  QH_TRY(qcd_hip_gauge_upload(g_dc.gauge, (const double *)u->D));
  *out = g_dc.gauge;
  return 0;
}

}  // namespace

namespace qcdhip {
// A drop-in call wrote new links into the host field whose data start at
// `host`: when that field is the bound (resident) one, its device copy is
// refreshed, so later calls on the field see the new links.
int dropin_gauge_written(const void *host) {
  if (!g_dc.bound || !host || g_dc.bound_host != host) return 0;
  return qcd_hip_gauge_upload(g_dc.bound, (const double *)host);
}
}  // namespace qcdhip

namespace {

// scratch spinors 4..7 are reserved for the drop-in layer
int spinor_in(int slot, qcd_hip_lattice lat, const qcd_vector *v, qcd_hip_field *out) {
  QH_TRY(scratch_spinor(4 + slot, lat, QCD_HIP_F64, 1, out));
  if (v) QH_TRY(qcd_hip_spinor_upload(*out, 0, (const double *)v->D));
  return 0;
}

int report(const char *who, int rc) {
  if (rc) fprintf(stderr, "Error in %s! %s\n", who, qcd_hip_last_error());
  return rc ? 1 : 0;
}

// lib/src/qcd_wilson.c:139-166
int check_operator_args(const char *who, qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u) {
  if (!v || !v->initialized) {
    fprintf(stderr, "Error in %s! Vector must be properly initialized\n", who);
    return 1;
  }
  if (!u || !u->initialized) {
    fprintf(stderr, "Error in %s! Gauge field must be properly initialized\n", who);
    return 1;
  }
  const qcd_geometry *a = v->geo, *b = u->geo;
  bool bad = !a || !b || a->initialized != 1 || b->initialized != 1;
  for (int d = 0; d < 4 && !bad; d++)
    bad = a->L[d] != b->L[d] || a->lL[d] != b->lL[d] || a->Pos[d] != b->Pos[d];
  if (bad) {
    fprintf(stderr, "Error in %s! Geometry not initialized or missmatched\n", who);
    return 1;
  }
  if (!Dv || !Dv->D) {
    fprintf(stderr, "Error in %s! Output vector has no data\n", who);
    return 1;
  }
  if (Dv->D == v->D) {
    fprintf(stderr, "Error in %s! Output must not alias the input\n", who);
    return 1;
  }
  return 0;
}

int apply_operator(const char *who, int op, qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, double m, double tm) {
  if (check_operator_args(who, Dv, v, u)) return 1;
  qcd_hip_lattice lat;
  qcd_hip_field g, in, out;
  // boundary phases come from the gauge field's geometry
  // (lib/src/qcd_wilson.c:174-177 reads u->geo->theta)
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = spinor_in(0, lat, v, &in);
  if (!rc) rc = spinor_in(1, lat, nullptr, &out);
  if (!rc) rc = qcd_hip_dslash(out, in, g, m, tm, op);
  if (!rc) rc = qcd_hip_spinor_download(out, 0, (double *)Dv->D);
  return report(who, rc);
}

bool on_rank(const qcd_geometry *geo, unsigned tt) {
  return tt >= (unsigned)geo->Pos[0] * geo->lL[0] && tt < (unsigned)(geo->Pos[0] + 1) * geo->lL[0];
}

// out = a*x + b*y on the device, copied back into `res` (host)
int lincomb_host(const char *who, qcd_vector *res, qcd_complex_16 a, qcd_vector *x, qcd_complex_16 b, qcd_vector *y,
                 int t0) {
  qcd_hip_lattice lat;
  qcd_hip_field fx, fy = nullptr, fo;
  int rc = lattice_for(x->geo, &lat);
  if (!rc) rc = spinor_in(0, lat, x, &fx);
  if (!rc && y) rc = spinor_in(1, lat, y, &fy);
  // slice-restricted updates leave the other sites of res unchanged
  if (!rc) rc = spinor_in(2, lat, t0 >= 0 ? res : nullptr, &fo);
  if (!rc) rc = qcd_hip_lincomb(fo, 0, a.re, a.im, fx, 0, b.re, b.im, fy, 0, t0);
  if (!rc) rc = qcd_hip_spinor_download(fo, 0, (double *)res->D);
  return report(who, rc);
}

}  // namespace

extern "C" {

// ============================================================ geometry
// lib/src/qcd_init.c:16-23
void qcd_antilexic(qcd_uint_2 x[], qcd_uint_8 l, const qcd_uint_2 dim[]) {
  x[0] = (qcd_uint_2)(l % dim[0]);
  l /= dim[0];
  x[1] = (qcd_uint_2)(l % dim[1]);
  l /= dim[1];
  x[2] = (qcd_uint_2)(l % dim[2]);
  x[3] = (qcd_uint_2)(l / dim[2]);
}

// lib/src/qcd_init.c:25-392, single rank: neighbour tables with periodic wrap,
// surface-site lists and the even/odd split; no halo, no MPI datatypes.
int qcd_initGeometry(qcd_geometry *geo, const qcd_uint_2 L[], const qcd_uint_2 P[], const qcd_real_8 theta[],
                     const int myid, const int numprocs) {
  if (P[0] != 1 || P[1] != 1 || P[2] != 1 || P[3] != 1 || numprocs != 1) {
    fprintf(stderr, "process %i: Error in qcd_initGeometry! This library runs one rank on one GPU: "
                    "need P = 1 1 1 1 and one process, got %i x %i x %i x %i and %i\n",
            myid, P[0], P[1], P[2], P[3], numprocs);
    return 1;
  }
  if (L[0] < 2 || L[1] < 2 || L[2] < 2 || L[3] < 2) {
    fprintf(stderr, "process %i: Error in qcd_initGeometry!\nprocess %i: Local lattices must be at least 2x2x2x2\n",
            myid, myid);
    return 1;
  }
  if (((unsigned long)L[0] * L[1] * L[2] * L[3]) % 2) {
    fprintf(stderr, "process %i: Error in qcd_initGeometry!\nprocess %i: Local volume must be multiple of 2.\n", myid,
            myid);
    return 1;
  }
  std::memset(geo, 0, sizeof(*geo));
  for (int d = 0; d < 4; d++) {
    geo->L[d] = geo->lL[d] = L[d];
    geo->theta[d] = theta[d];
  }
  geo->V = geo->lV = (qcd_uint_8)L[0] * L[1] * L[2] * L[3];
  geo->V3 = geo->lV3 = (qcd_uint_8)L[1] * L[2] * L[3];
  geo->myid = 0;
  geo->nproc = 1;
  const qcd_uint_8 V = geo->lV, V3 = geo->lV3;
  geo->plus = (qcd_uint_8(*)[4])malloc(4 * V * sizeof(qcd_uint_8));
  geo->minus = (qcd_uint_8(*)[4])malloc(4 * V * sizeof(qcd_uint_8));
  geo->plus3d = (qcd_uint_8(*)[4])malloc(4 * V3 * sizeof(qcd_uint_8));
  geo->minus3d = (qcd_uint_8(*)[4])malloc(4 * V3 * sizeof(qcd_uint_8));
  geo->eo = (qcd_uint_8(*)[2])malloc(2 * (V / 2) * sizeof(qcd_uint_8));
  geo->edgePoints = V - (qcd_uint_8)(L[0] - 2) * (L[1] - 2) * (L[2] - 2) * (L[3] - 2);
  geo->edge0Points = V3 - (qcd_uint_8)(L[1] - 2) * (L[2] - 2) * (L[3] - 2);
  geo->edge = (qcd_uint_8 *)malloc((geo->edgePoints ? geo->edgePoints : 1) * sizeof(qcd_uint_8));
  geo->edge0 = (qcd_uint_8 *)malloc((geo->edge0Points ? geo->edge0Points : 1) * sizeof(qcd_uint_8));
  if (!geo->plus || !geo->minus || !geo->plus3d || !geo->minus3d || !geo->eo || !geo->edge || !geo->edge0) {
    fprintf(stderr, "process %i: Error in qcd_initGeometry! Out of memory\n", myid);
    qcd_destroyGeometry(geo);
    return 1;
  }
  qcd_uint_8 ne = 0, ne0 = 0, nev = 0, nod = 0;
  for (qcd_uint_8 z = 0; z < L[3]; z++)
    for (qcd_uint_8 y = 0; y < L[2]; y++)
      for (qcd_uint_8 x = 0; x < L[1]; x++) {
        const qcd_uint_8 j = qcd_LEXIC0(x, y, z, L);
        geo->plus3d[j][0] = geo->minus3d[j][0] = j;
        geo->plus3d[j][1] = qcd_LEXIC0((x + 1) % L[1], y, z, L);
        geo->plus3d[j][2] = qcd_LEXIC0(x, (y + 1) % L[2], z, L);
        geo->plus3d[j][3] = qcd_LEXIC0(x, y, (z + 1) % L[3], L);
        geo->minus3d[j][1] = qcd_LEXIC0((x + L[1] - 1) % L[1], y, z, L);
        geo->minus3d[j][2] = qcd_LEXIC0(x, (y + L[2] - 1) % L[2], z, L);
        geo->minus3d[j][3] = qcd_LEXIC0(x, y, (z + L[3] - 1) % L[3], L);
        const bool s3 = x == 0 || y == 0 || z == 0 || x == L[1] - 1u || y == L[2] - 1u || z == L[3] - 1u;
        if (s3) geo->edge0[ne0++] = j;
        for (qcd_uint_8 t = 0; t < L[0]; t++) {
          const qcd_uint_8 l = qcd_LEXIC(t, x, y, z, L);
          geo->plus[l][0] = qcd_LEXIC((t + 1) % L[0], x, y, z, L);
          geo->plus[l][1] = qcd_LEXIC(t, (x + 1) % L[1], y, z, L);
          geo->plus[l][2] = qcd_LEXIC(t, x, (y + 1) % L[2], z, L);
          geo->plus[l][3] = qcd_LEXIC(t, x, y, (z + 1) % L[3], L);
          geo->minus[l][0] = qcd_LEXIC((t + L[0] - 1) % L[0], x, y, z, L);
          geo->minus[l][1] = qcd_LEXIC(t, (x + L[1] - 1) % L[1], y, z, L);
          geo->minus[l][2] = qcd_LEXIC(t, x, (y + L[2] - 1) % L[2], z, L);
          geo->minus[l][3] = qcd_LEXIC(t, x, y, (z + L[3] - 1) % L[3], L);
          if (s3 || t == 0 || t == L[0] - 1u) geo->edge[ne++] = l;
          if ((t + x + y + z) % 2 == 0) geo->eo[nev++][0] = l;
          else geo->eo[nod++][1] = l;
        }
      }
  geo->initialized = 1;
  return 0;
}

void qcd_destroyGeometry(const qcd_geometry *geo) {
  free(geo->plus);
  free(geo->minus);
  free(geo->plus3d);
  free(geo->minus3d);
  free(geo->edge);
  free(geo->edge0);
  free(geo->eo);
}

// lib/src/qcd_init.c:418-665: host allocation of lV sites (no halo at one rank)
#define QCD_INIT_FIELD(FN, DFN, TYPE, NAME, ELEMS)                                             \
  int FN(TYPE *f, qcd_geometry *geo) {                                                         \
    if (!geo || !geo->initialized) {                                                           \
      fprintf(stderr, "error in " #FN "! Geometry must be properly initialized.\n");          \
      return 1;                                                                                \
    }                                                                                          \
    std::memset(f, 0, sizeof(*f));                                                             \
    f->D = (decltype(f->D))malloc((size_t)(ELEMS) * sizeof(qcd_complex_16));                  \
    if (!f->D) {                                                                               \
      fprintf(stderr, "process %i: Error in " #FN "! Out of memory\n", geo->myid);            \
      return 1;                                                                                \
    }                                                                                          \
    f->geo = geo;                                                                              \
    f->initialized = 1;                                                                        \
    return 0;                                                                                  \
  }                                                                                            \
  void DFN(TYPE *f) {                                                                          \
    if (!f->initialized) {                                                                     \
      fprintf(stderr, "Error in " #DFN "! " NAME " not initialized\n");                       \
      return;                                                                                  \
    }                                                                                          \
    free(f->D);                                                                                \
    f->D = nullptr;                                                                            \
    f->initialized = 0;                                                                        \
  }

QCD_INIT_FIELD(qcd_initVector, qcd_destroyVector, qcd_vector, "Vector", 12 * geo->lV)
QCD_INIT_FIELD(qcd_initVector3d, qcd_destroyVector3d, qcd_vector3d, "Vector", 12 * geo->lV3)
QCD_INIT_FIELD(qcd_initPropagator, qcd_destroyPropagator, qcd_propagator, "Propagator", 144 * geo->lV)
QCD_INIT_FIELD(qcd_initGaugeField, qcd_destroyGaugeField, qcd_gaugeField, "GaugeField", 36 * geo->lV)
#undef QCD_INIT_FIELD

// include/qcd_communication.h: nothing to exchange with one rank
int qcd_communicateVectorPMGaugeP(qcd_vector *v, qcd_gaugeField *u) { return 0; }
int qcd_communicateVectorPM(qcd_vector *v) { return 0; }
void qcd_waitall(qcd_geometry *geo) {}

// ============================================================ operators
int qcd_applyWilsonTMOp(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm) {
  return apply_operator("qcd_ApplyWilsonTMOp", QCD_HIP_OP_D, Dv, v, u, m, tm);
}
int qcd_applyWilsonTMOpNoScripts(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm) {
  return apply_operator("qcd_ApplyWilsonTMOp", QCD_HIP_OP_D, Dv, v, u, m, tm);
}
int qcd_applyWilsonTMOpNoOptim(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm) {
  return apply_operator("qcd_ApplyWilsonTMOp", QCD_HIP_OP_D, Dv, v, u, m, tm);
}
int qcd_applyQTMOp(qcd_vector *Dv, qcd_vector *v, qcd_gaugeField *u, qcd_real_8 m, qcd_real_8 tm) {
  return apply_operator("qcd_applyQTMOp", QCD_HIP_OP_Q, Dv, v, u, m, tm);
}

// D (or Q) applied to all 12 columns of a propagator in one batched launch:
// the columns are gathered on the device into 12 rhs of a library field
// (scratch slot 7), the operator runs over the 12 rhs together, and the
// results are scattered back into Dp (scratch slot 3).  The reference applies
// the operator column by column after qcd_copyVectorPropagator
// (lib/src/qcd_blas.c:182-212, lib/src/qcd_wilson.c:114-336).
static int apply_propagator(const char *who, int op, qcd_propagator *Dp, qcd_propagator *p, qcd_gaugeField *u,
                            double m, double tm) {
  if (!p || !p->initialized || !Dp || !Dp->initialized || !u || !u->initialized) {
    fprintf(stderr, "Error in %s! Propagators and gauge field must be properly initialized\n", who);
    return 1;
  }
  const qcd_geometry *a = p->geo, *b = u->geo, *c = Dp->geo;
  bool bad = !a || !b || !c;
  for (int d = 0; d < 4 && !bad; d++) bad = a->L[d] != b->L[d] || a->L[d] != c->L[d] || a->lL[d] != b->lL[d];
  if (bad) {
    fprintf(stderr, "Error in %s! Geometry not initialized or missmatched\n", who);
    return 1;
  }
  if (Dp->D == p->D) {
    fprintf(stderr, "Error in %s! Output must not alias the input\n", who);
    return 1;
  }
  qcd_hip_lattice lat;
  qcd_hip_field g, in = nullptr, out = nullptr;
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = scratch_spinor(7, lat, QCD_HIP_F64, 12, &in);
  if (!rc) rc = scratch_spinor(3, lat, QCD_HIP_F64, 12, &out);
  if (!rc && (in->nrhs != 12 || out->nrhs != 12)) rc = 1;   // the slots only ever hold 12-rhs fields
  if (!rc) rc = qcd_hip_propagator_upload(in, 0, (const double *)p->D);
  if (!rc) rc = qcd_hip_dslash(out, in, g, m, tm, op);
  if (!rc) rc = qcd_hip_propagator_download(out, 0, (double *)Dp->D);
  return report(who, rc);
}

int qcd_hip_applyWilsonTMOpPropagator(qcd_propagator *Dp, qcd_propagator *p, qcd_gaugeField *u, qcd_real_8 m,
                                      qcd_real_8 tm) {
  return apply_propagator("qcd_hip_applyWilsonTMOpPropagator", QCD_HIP_OP_D, Dp, p, u, m, tm);
}

int qcd_hip_applyQTMOpPropagator(qcd_propagator *Dp, qcd_propagator *p, qcd_gaugeField *u, qcd_real_8 m,
                                 qcd_real_8 tm) {
  return apply_propagator("qcd_hip_applyQTMOpPropagator", QCD_HIP_OP_Q, Dp, p, u, m, tm);
}

// Keep u resident in HBM: later drop-in calls with this field skip the
// upload until qcd_hip_unbindGaugeField().  The caller must rebind after
// changing u->D.
int qcd_hip_bindGaugeField(qcd_gaugeField *u) {
  if (!u || !u->initialized || !u->geo) {
    fprintf(stderr, "Error in qcd_hip_bindGaugeField! Gauge field must be properly initialized\n");
    return 1;
  }
  qcd_hip_unbindGaugeField();
  qcd_hip_lattice lat;
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = qcd_hip_gauge_create(&g_dc.bound, lat, QCD_HIP_F64);
  if (!rc) rc = qcd_hip_gauge_upload(g_dc.bound, (const double *)u->D);
  if (!rc) g_dc.bound_host = u->D;
  return report("qcd_hip_bindGaugeField", rc);
}

void qcd_hip_unbindGaugeField(void) {
  if (g_dc.bound) {
    qcd_hip_lattice l = g_dc.bound->lat;
    qcd_hip_field_destroy(g_dc.bound);
    if (l != g_dc.lat) qcd_hip_lattice_destroy(l);
  }
  g_dc.bound = nullptr;
  g_dc.bound_host = nullptr;
}

// ============================================================ smearing
// lib/src/qcd_smearing.c:25-107
int qcd_gaussIteration3d(qcd_vector *v, qcd_gaugeField *u, qcd_real_8 alpha, qcd_uint_4 t) {
  if (!v || !v->initialized) {
    fprintf(stderr, "Error in qcd_gaussIteration3d! Vector mus be properly initialized\n");
    return 1;
  }
  if (!u || !u->initialized) {
    fprintf(stderr, "Error in qcd_gaussIteration3d! Gauge field must be properly initialized\n");
    return 1;
  }
  if (!on_rank(v->geo, t)) return 0;   // the slice is not on this rank: nothing to do
  qcd_hip_lattice lat;
  qcd_hip_field g, f;
  int rc = lattice_for(v->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = spinor_in(0, lat, v, &f);
  if (!rc) rc = qcd_hip_gauss_smear(f, g, alpha, (int)t, 1);
  if (!rc) rc = qcd_hip_spinor_download(f, 0, (double *)v->D);
  return report("qcd_gaussIteration3d", rc);
}

// lib/src/qcd_smearing.c:116-199
int qcd_gaussIteration3dAll(qcd_vector *v, qcd_gaugeField *u, qcd_real_8 alpha, qcd_uint_2 gaugeCom) {
  if (!v || !v->initialized) {
    fprintf(stderr, "Error in qcd_gaussIteration3dAll! Vector mus be properly initialized\n");
    return 1;
  }
  if (!u || !u->initialized) {
    fprintf(stderr, "Error in qcd_gaussIteration3dAll! Gauge field must be properly initialized\n");
    return 1;
  }
  qcd_hip_lattice lat;
  qcd_hip_field g, f;
  int rc = lattice_for(v->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = spinor_in(0, lat, v, &f);
  if (!rc) rc = qcd_hip_gauss_smear(f, g, alpha, -1, 1);
  if (!rc) rc = qcd_hip_spinor_download(f, 0, (double *)v->D);
  return report("qcd_gaussIteration3dAll", rc);
}

// lib/src/qcd_smearing.c:205-273 (one step; apeu may alias u here)
int qcd_apeSmear3d(qcd_gaugeField *apeu, qcd_gaugeField *u, qcd_real_8 alpha) {
  if (!u || !u->initialized || !apeu || !apeu->initialized) {
    fprintf(stderr, "Error in qcd_apeSmear3d! Gauge fields must be properly initialized\n");
    return 1;
  }
  qcd_hip_lattice lat;
  qcd_hip_field g, out = nullptr;
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = qcd_hip_gauge_create(&out, lat, QCD_HIP_F64);
  if (!rc) rc = qcd_hip_ape_smear(out, g, alpha, 1);
  if (!rc) rc = qcd_hip_gauge_download(out, (double *)apeu->D);
  if (!rc) rc = dropin_gauge_written(apeu->D);
  qcd_hip_field_destroy(out);
  return report("qcd_apeSmear3d", rc);
}

// lib/src/qcd_smearing.c:374-476 (one step; u_out may alias u)
int qcd_stoutsmearing(qcd_gaugeField *u_out, qcd_gaugeField *u, qcd_real_8 rho, int ndim) {
  if (!u || !u->initialized || !u_out || !u_out->initialized) {
    fprintf(stderr, "Error in qcd_stoutsmearing! Gauge fields must be properly initialized\n");
    return 1;
  }
  if (ndim != 3 && ndim != 4) {
    fprintf(stderr, "ndim should be either 3 or 4");
    return 1;
  }
  qcd_hip_lattice lat;
  qcd_hip_field g, out = nullptr;
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = qcd_hip_gauge_create(&out, lat, QCD_HIP_F64);
  if (!rc) rc = qcd_hip_stout_smear(out, g, rho, ndim, 1);
  if (!rc) rc = qcd_hip_gauge_download(out, (double *)u_out->D);
  if (!rc) rc = dropin_gauge_written(u_out->D);
  qcd_hip_field_destroy(out);
  return report("qcd_stoutsmearing", rc);
}

// lib/src/qcd_blas.c:940-1364: spatial links projected onto SU(3) in place
// (an APE step with alpha = 0)
void qcd_projectSU33d(qcd_gaugeField *gf) {
  if (!gf || !gf->initialized) return;
  qcd_hip_lattice lat;
  qcd_hip_field g, out = nullptr;
  int rc = lattice_for(gf->geo, &lat);
  if (!rc) rc = gauge_for(gf, lat, &g);
  if (!rc) rc = qcd_hip_gauge_create(&out, lat, QCD_HIP_F64);
  if (!rc) rc = qcd_hip_ape_smear(out, g, 0.0, 1);
  if (!rc) rc = qcd_hip_gauge_download(out, (double *)gf->D);
  if (!rc) rc = dropin_gauge_written(gf->D);
  qcd_hip_field_destroy(out);
  report("qcd_projectSU33d", rc);
}

// lib/src/qcd_blas.c:541-939: every link (all four directions) projected
// onto SU(3) in place, on the device
void qcd_projectSU3(qcd_gaugeField *gf) {
  if (!gf || !gf->initialized) return;
  qcd_hip_lattice lat;
  qcd_hip_field g;
  int rc = lattice_for(gf->geo, &lat);
  if (!rc) rc = gauge_for(gf, lat, &g);   // a bound field is projected in its resident copy
  if (!rc) rc = qcd_hip_project_su3(g);
  if (!rc) rc = qcd_hip_gauge_download(g, (double *)gf->D);
  report("qcd_projectSU3", rc);
}

// lib/src/qcd_observables.c:15-36
qcd_real_8 qcd_calculatePlaquette(qcd_gaugeField *u) {
  if (!u || !u->initialized) {
    fprintf(stderr, "Error in qcd_calculatePlaquette! Gauge field must be properly initialized\n");
    return 0.0;
  }
  qcd_hip_lattice lat;
  qcd_hip_field g;
  double plaq = 0.0;
  int rc = lattice_for(u->geo, &lat);
  if (!rc) rc = gauge_for(u, lat, &g);
  if (!rc) rc = qcd_hip_plaquette(g, &plaq);
  report("qcd_calculatePlaquette", rc);
  return rc ? 0.0 : plaq;
}

// ============================================================ BLAS
// host data movement: lib/src/qcd_blas.c:137-233
void qcd_zeroVector(qcd_vector *vec) { std::memset(vec->D, 0, 12 * vec->geo->lV * sizeof(qcd_complex_16)); }
void qcd_zeroPropagator(qcd_propagator *p) { std::memset(p->D, 0, 144 * p->geo->lV * sizeof(qcd_complex_16)); }
void qcd_zeroGaugeField(qcd_gaugeField *u) {
  std::memset(u->D, 0, 36 * u->geo->lV * sizeof(qcd_complex_16));
  report("qcd_zeroGaugeField", dropin_gauge_written(u->D));
}

void qcd_setVector(qcd_vector *vec, qcd_complex_16 c) {
  qcd_complex_16 *p = &vec->D[0][0][0];
  for (qcd_uint_8 i = 0; i < 12 * vec->geo->lV; i++) p[i] = c;
}

void qcd_copyVector(qcd_vector *dest, qcd_vector *src) {
  std::memcpy(dest->D, src->D, 12 * dest->geo->lV * sizeof(qcd_complex_16));
}

void qcd_copyGaugeField(qcd_gaugeField *dest, qcd_gaugeField *src) {
  std::memcpy(dest->D, src->D, 36 * dest->geo->lV * sizeof(qcd_complex_16));
  report("qcd_copyGaugeField", dropin_gauge_written(dest->D));
}

void qcd_copyVectorPropagator(qcd_vector *vec, qcd_propagator *prop, qcd_uint_2 nu, qcd_uint_2 c2) {
  for (qcd_uint_8 i = 0; i < vec->geo->lV; i++)
    for (int mu = 0; mu < 4; mu++)
      for (int c1 = 0; c1 < 3; c1++) vec->D[i][mu][c1] = prop->D[i][mu][nu][c1][c2];
}

void qcd_copyPropagatorVector(qcd_propagator *prop, qcd_vector *vec, qcd_uint_2 nu, qcd_uint_2 c2) {
  for (qcd_uint_8 i = 0; i < vec->geo->lV; i++)
    for (int mu = 0; mu < 4; mu++)
      for (int c1 = 0; c1 < 3; c1++) prop->D[i][mu][nu][c1][c2] = vec->D[i][mu][c1];
}

void qcd_copyVectorPropagator2(qcd_vector *vec, qcd_propagator *prop, qcd_uint_2 mu, qcd_uint_2 c1) {
  for (qcd_uint_8 i = 0; i < vec->geo->lV; i++)
    for (int nu = 0; nu < 4; nu++)
      for (int c2 = 0; c2 < 3; c2++) vec->D[i][nu][c2] = prop->D[i][mu][nu][c1][c2];
}

void qcd_copyPropagatorVector2(qcd_propagator *prop, qcd_vector *vec, qcd_uint_2 mu, qcd_uint_2 c1) {
  for (qcd_uint_8 i = 0; i < vec->geo->lV; i++)
    for (int nu = 0; nu < 4; nu++)
      for (int c2 = 0; c2 < 3; c2++) prop->D[i][mu][nu][c1][c2] = vec->D[i][nu][c2];
}

// arithmetic on the GPU: lib/src/qcd_blas.c:237-491
static const qcd_complex_16 kZero = {0.0, 0.0};
static const qcd_complex_16 kOne = {1.0, 0.0};

void qcd_mulVectorC(qcd_vector *vec, qcd_complex_16 c) { lincomb_host("qcd_mulVectorC", vec, c, vec, kZero, nullptr, -1); }

void qcd_mulVectorC3d(qcd_vector *vec, qcd_complex_16 c, qcd_uint_4 t) {
  if (!on_rank(vec->geo, t)) return;
  lincomb_host("qcd_mulVectorC3d", vec, c, vec, kZero, nullptr, (int)t);
}

void qcd_scaleVector(qcd_vector *vec, qcd_real_8 r) {
  lincomb_host("qcd_scaleVector", vec, qcd_complex_16{r, 0.0}, vec, kZero, nullptr, -1);
}

void qcd_scaleVector3d(qcd_vector *vec, qcd_real_8 r, qcd_uint_4 tt) {
  if (!on_rank(vec->geo, tt)) return;
  lincomb_host("qcd_scaleVector3d", vec, qcd_complex_16{r, 0.0}, vec, kZero, nullptr, (int)tt);
}

void qcd_subVector(qcd_vector *difference, qcd_vector *minuend, qcd_vector *subtrahend) {
  lincomb_host("qcd_subVector", difference, kOne, minuend, qcd_complex_16{-1.0, 0.0}, subtrahend, -1);
}

void qcd_addVector(qcd_vector *sum, qcd_vector *summand1, qcd_vector *summand2) {
  lincomb_host("qcd_addVector", sum, kOne, summand1, kOne, summand2, -1);
}

void qcd_axpyVector(qcd_vector *axpy, qcd_complex_16 a, qcd_vector *x, qcd_vector *y) {
  lincomb_host("qcd_axpyVector", axpy, a, x, kOne, y, -1);
}

void qcd_addVector3d(qcd_vector *sum, qcd_vector *summand1, qcd_vector *summand2, qcd_uint_4 tt) {
  if (!on_rank(sum->geo, tt)) return;
  lincomb_host("qcd_addVector3d", sum, kOne, summand1, kOne, summand2, (int)tt);
}

qcd_complex_16 qcd_mulAdjointVectorVector(qcd_vector *factor1, qcd_vector *factor2) {
  qcd_complex_16 r = {NAN, NAN};
  qcd_hip_lattice lat;
  qcd_hip_field a, b;
  double d[2];
  int rc = lattice_for(factor1->geo, &lat);
  if (!rc) rc = spinor_in(0, lat, factor1, &a);
  if (!rc) rc = spinor_in(1, lat, factor2, &b);
  if (!rc) rc = qcd_hip_dot(a, 0, b, 0, d);
  if (report("qcd_mulAdjointVectorVector", rc) == 0) r = qcd_complex_16{d[0], d[1]};
  return r;
}

qcd_real_8 qcd_normVector(qcd_vector *vec) {
  qcd_hip_lattice lat;
  qcd_hip_field a;
  double n2 = NAN;
  int rc = lattice_for(vec->geo, &lat);
  if (!rc) rc = spinor_in(0, lat, vec, &a);
  if (!rc) rc = qcd_hip_norm2(a, 0, &n2);
  report("qcd_normVector", rc);
  return std::sqrt(n2);
}

}  // extern "C"
