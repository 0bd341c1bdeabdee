// serve.cpp -- continuous serving of propagator-solve requests on one GPU.
//
// Requests (lattice, kappa, gauge seed and eps, 1 or 12 point sources) arrive
// over time.  A pool of worker threads, each driving its own HIP streams,
// takes arrived requests in arrival order as soon as a worker is free and the
// request's device footprint fits in the memory budget; requests therefore
// enter while others are on the GPU and their kernels overlap on the device.
// Each request keeps its gauge field and solver state resident in HBM for
// its lifetime: the field is generated on the device, the right-hand sides of
// the request are solved as one batch (solver.hip), and the FP64 true
// residual ||b - D x|| / ||b|| of every rhs is the acceptance test.
// Per-request arithmetic does not depend on what else runs (deterministic
// reductions, launch configurations tuned before the clock starts), so a
// result is the same alone and in the mix.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "blas_common.h"

using namespace qcdhip;

namespace {

using Clock = std::chrono::steady_clock;

size_t vec_bytes(long V, int prec) {
  const long st = round_up(V, kTile);
  return (size_t)st * 12 * spinor_comp_bytes(prec) + (prec == QCD_HIP_F16 ? (size_t)st * 4 : 0);
}

// the solver parameters a request is served with: one batch of all its rhs,
// even-odd preconditioned inner solve when every extent is even
void request_params(const qcd_hip_request &q, qcd_hip_solver_params *p) {
  qcd_hip_solver_default_params(p);
  p->n_restart = q.n_restart;
  p->tol = q.tol;
  p->inner_prec = q.inner_prec;
  p->batch = q.nrhs;
  p->max_outer = 20000;
  p->eo = (q.L[0] % 2 == 0 && q.L[1] % 2 == 0 && q.L[2] % 2 == 0 && q.L[3] % 2 == 0) ? 1 : 0;
}

// device bytes a request holds while it runs: its gauge field, x and b, and
// the solver workspace exactly as qcd_hip_solver_create sizes it
size_t request_bytes(const qcd_hip_request &r) {
  const long V = (long)r.L[0] * r.L[1] * r.L[2] * r.L[3];
  const long st = round_up(V, kTile);
  qcd_hip_solver_params p;
  request_params(r, &p);
  size_t b = (size_t)st * 36 * 16;                                     // FP64 gauge field
  b += 2 * (size_t)r.nrhs * vec_bytes(V, QCD_HIP_F64);                  // x, b
  b += solver_device_bytes(r.L, &p);
  return b;
}

int run_request(const qcd_hip_request &q, qcd_hip_result *res, double *solution, hipStream_t ws) {
  qcd_hip_lattice lat = nullptr;
  qcd_hip_field g = nullptr, b = nullptr, x = nullptr;
  qcd_hip_solver S = nullptr;
  const double theta[4] = {M_PI, 0.0, 0.0, 0.0};
  int rc = qcd_hip_lattice_create(&lat, q.L, theta);
  if (!rc) rc = qcd_hip_gauge_create(&g, lat, QCD_HIP_F64);
  if (!rc) rc = qcd_hip_spinor_create(&b, lat, QCD_HIP_F64, q.nrhs);
  if (!rc) rc = qcd_hip_spinor_create(&x, lat, QCD_HIP_F64, q.nrhs);
  if (!rc) rc = gauge_generate_device(g, q.eps, q.seed, ws);
  for (int r = 0; !rc && r < q.nrhs; r++) rc = point_source_device(b, r, 0, r / 3, r % 3, ws);
  if (!rc && hipStreamSynchronize(ws) != hipSuccess) rc = 2;
  qcd_hip_solver_params p;
  request_params(q, &p);
  if (!rc) rc = qcd_hip_solver_create(&S, lat, &p);
  if (!rc) rc = qcd_hip_solver_set_gauge(S, g);
  qcd_hip_solver_stats st;
  std::memset(&st, 0, sizeof(st));
  if (!rc) rc = qcd_hip_solver_solve(S, x, b, q.kappa, 0.0, &st);
  res->rc = rc;
  res->outer = st.outer;
  res->iters = st.inner_iters;
  res->relres = st.relres;
  if (x && (rc == 0 || rc == 6)) {
    double s = 0.0;
    for (int r = 0; r < q.nrhs; r++) {
      double n2 = 0.0;
      if (qcd_hip_norm2(x, r, &n2) == 0) s += n2;
    }
    res->xnorm2 = s;
    if (solution)
      for (int r = 0; r < q.nrhs; r++)
        qcd_hip_spinor_download(x, r, solution + (size_t)r * 24 * lat->V);
  }
  qcd_hip_solver_destroy(S);
  qcd_hip_field_destroy(x);
  qcd_hip_field_destroy(b);
  qcd_hip_field_destroy(g);
  qcd_hip_lattice_destroy(lat);
  return rc;
}

// Tune every (lattice, precision) the requests will use before timing.
int pretune(const qcd_hip_request *req, int n) {
  const double theta[4] = {M_PI, 0.0, 0.0, 0.0};
  for (int i = 0; i < n; i++) {
    bool seen = false;
    for (int j = 0; j < i && !seen; j++)
      seen = !std::memcmp(req[j].L, req[i].L, sizeof(req[i].L)) && req[j].inner_prec == req[i].inner_prec;
    if (seen) continue;
    qcd_hip_lattice lat = nullptr;
    qcd_hip_field g = nullptr, gi = nullptr, a = nullptr, c = nullptr, ai = nullptr, ci = nullptr;
    int rc = qcd_hip_lattice_create(&lat, req[i].L, theta);
    if (!rc) rc = qcd_hip_gauge_create(&g, lat, QCD_HIP_F64);
    if (!rc) rc = qcd_hip_spinor_create(&a, lat, QCD_HIP_F64, 1);
    if (!rc) rc = qcd_hip_spinor_create(&c, lat, QCD_HIP_F64, 1);
    if (!rc) rc = qcd_hip_dslash_tune(c, a, g, nullptr, nullptr);
    if (!rc && req[i].inner_prec != QCD_HIP_F64) {
      rc = qcd_hip_gauge_create(&gi, lat, req[i].inner_prec);
      if (!rc) rc = qcd_hip_spinor_create(&ai, lat, req[i].inner_prec, 1);
      if (!rc) rc = qcd_hip_spinor_create(&ci, lat, req[i].inner_prec, 1);
      if (!rc) rc = qcd_hip_dslash_tune(ci, ai, gi, nullptr, nullptr);
    }
    for (qcd_hip_field f : {g, gi, a, c, ai, ci}) qcd_hip_field_destroy(f);
    qcd_hip_lattice_destroy(lat);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

int qcd_hip_serve(const qcd_hip_request *req, int n, int workers, size_t budget, int realtime,
                  qcd_hip_result *res, double *wall_s, double *const *solutions) {
  QH_TRY(ensure_init());
  if (!req || !res || n < 0 || workers < 1 || workers > 64) {
    set_error("serve: bad arguments (workers 1..64)");
    return 1;
  }
  for (int i = 0; i < n; i++) {
    const qcd_hip_request &q = req[i];
    if (q.nrhs < 1 || q.nrhs > 64 || q.n_restart < 1 || q.n_restart > kMaxKrylov || q.tol <= 0.0 ||
        q.kappa <= 0.0 || q.inner_prec < QCD_HIP_F64 || q.inner_prec > QCD_HIP_F16) {
      set_error("serve: request %d has bad parameters", i);
      return 1;
    }
    if (request_bytes(q) > budget) {
      set_error("serve: request %d needs %zu bytes, more than the whole budget %zu", i, request_bytes(q), budget);
      return 1;
    }
  }
  std::memset(res, 0, sizeof(qcd_hip_result) * (size_t)n);
  QH_TRY(pretune(req, n));
  QH_CHECK(hipDeviceSynchronize());

  std::mutex mu;
  std::condition_variable cv;
  int next = 0;            // next request in arrival order
  size_t reserved = 0;     // bytes of the requests being served
  int failed = 0;
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
  const int device = ctx()->device;
  const auto t0 = Clock::now();
  auto since = [&](Clock::time_point t) { return std::chrono::duration<double>(t - t0).count(); };

  auto worker = [&]() {
    (void)hipSetDevice(device);
    hipStream_t ws = nullptr;
    if (hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu);
      failed = 1;
      return;
    }
    for (;;) {
      int i = -1;
      size_t need = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (next >= n) break;
          const int cand = order[next];
          const double due = realtime ? req[cand].arrival_s : 0.0;
          const double now = since(Clock::now());
          need = request_bytes(req[cand]);
          if (now >= due && reserved + need <= budget) {
            i = cand;
            next++;
            reserved += need;
            break;
          }
          if (now < due)
            cv.wait_for(lk, std::chrono::duration<double>(std::min(due - now, 0.01)));
          else
            cv.wait(lk);   // memory is held by running requests: wait for one to finish
        }
      }
      if (i < 0) break;
      const double start = since(Clock::now());
      const double arrival = realtime ? req[i].arrival_s : 0.0;
      qcd_hip_result r;
      std::memset(&r, 0, sizeof(r));
      run_request(req[i], &r, solutions ? solutions[i] : nullptr, ws);
      const double end = since(Clock::now());
      r.queued_s = start - arrival;
      r.service_s = end - start;
      r.latency_s = end - arrival;
      r.bytes = (double)need;
      {
        std::lock_guard<std::mutex> lk(mu);
        res[i] = r;
        reserved -= need;
      }
      cv.notify_all();
    }
    (void)hipStreamDestroy(ws);
  };
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; w++) pool.emplace_back(worker);
  for (auto &t : pool) t.join();
  if (wall_s) *wall_s = since(Clock::now());
  if (failed) {
    set_error("serve: a worker could not create its stream");
    return 2;
  }
  for (int i = 0; i < n; i++)
    if (res[i].rc) {
      set_error("serve: request %d ended with code %d", i, res[i].rc);
      return 7;
    }
  return 0;
}

}  // extern "C"
