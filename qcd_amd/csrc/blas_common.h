// blas_common.h -- declarations shared by the BLAS, smearing and solver files.
#pragma once

#include "dslash_common.h"

namespace qcdhip {

// grid of the reduction kernels: depends only on the volume
int reduce_grid(long V);
// sum n_sums partial arrays of length nparts into out (device), one workgroup
int finalize(const double *partials, int nparts, int n_sums, double *out, hipStream_t st);

// out[ro] = a x[rx] + b y[ry] (all sites, or slice t0 >= 0), asynchronous
int lincomb_launch(qcd_hip_field out, int ro, double are, double aim, qcd_hip_field x, int rx, double bre, double bim,
                   qcd_hip_field y, int ry, int t0, hipStream_t st);
// cached scratch spinor number `slot` (resized on demand), owned by the library
int scratch_spinor(int slot, qcd_hip_lattice lat, int prec, int nrhs, qcd_hip_field *out);
// n iterations of Gauss smearing of every rhs of v on slice t0 (all when < 0)
int gauss_iterate(qcd_hip_field v, qcd_hip_field g, qcd_hip_field tmp, double alpha, int t0, int niter,
                  hipStream_t st);

}  // namespace qcdhip
