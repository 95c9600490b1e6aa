// One right-preconditioned Arnoldi step of GMRES as a single host call.
//
// A step of the device-stepped GMRES (solvers/linear_solver.py:438-525: z = P v,
// w = A z, w masked, orthogonalisation and Givens update, residual estimate to the
// host) issues about ten kernels.  Driving each of them from the host language costs
// more host time per step than the small and middle-sized problems of a solve server
// spend on the device, so the step is composed here, in one call that queues the
// whole sequence on the run's stream:
//
//   z = P V[k-1]          iga_fd_apply (zero outside the FD's free box), or a copy
//   w = A z               iga_fused_apply (coefficient sweep + gather)
//   w[constrained] = 0    iga_box_mask
//   Arnoldi / Givens      iga_gmres_step
//   hist[k-1] -> host     a 16-byte copy into pinned host memory and an event
//
// The context fixes everything that does not change between steps (operator,
// preconditioner, box, the run's vectors and least-squares state); the basis is
// passed per step because it grows while the run proceeds.

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace iga {
namespace {

constexpr int kEventRing = 16;

}  // namespace
}  // namespace iga

struct iga_arnoldi_s {
  int dtype = IGA_F64;
  int64_t n = 0;
  int m = 0;
  iga_fused_t op = nullptr;
  const void* packed = nullptr;
  const void* nurbs_w = nullptr;
  void* op_work = nullptr;
  iga_fd_t fd = nullptr;
  void* fd_work = nullptr;
  int ndim = 0;  // 0: no mask
  int64_t box_n[8] = {0}, box_lo[8] = {0}, box_len[8] = {0};
  void* z = nullptr;
  void* w = nullptr;
  double* ls = nullptr;
  void* work = nullptr;
  double* hist_host = nullptr;
  int64_t hist_off = 0;
  hipEvent_t ev[iga::kEventRing] = {};
  int ev_step[iga::kEventRing];
};

namespace iga {
namespace {

// Queue the copy of hist entry j (g and the subdiagonal of Krylov dimension j) to
// the host and the event the host polls for it.
void record(iga_arnoldi_s* c, int j, hipStream_t s) {
  IGA_HIP(hipMemcpyAsync(c->hist_host + 2 * (int64_t)j, c->ls + c->hist_off + 2 * (int64_t)j, 2 * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  const int slot = j % kEventRing;
  if (!c->ev[slot]) IGA_HIP(hipEventCreateWithFlags(&c->ev[slot], hipEventDisableTiming));
  IGA_HIP(hipEventRecord(c->ev[slot], s));
  c->ev_step[slot] = j;
}

hipEvent_t event_of(iga_arnoldi_s* c, int j) {
  IGA_CHECK_ARG(j >= 0, "bad step");
  const int slot = j % kEventRing;
  IGA_CHECK_ARG(c->ev[slot] && c->ev_step[slot] == j,
                "the event of this step was not recorded or was reused (more than 16 steps in flight)");
  return c->ev[slot];
}

}  // namespace
}  // namespace iga

using namespace iga;

extern "C" {

int iga_arnoldi_create(iga_arnoldi_t* out, int dtype, int64_t n, int m, iga_fused_t op, const void* packed,
                       const void* nurbs_w, void* op_work, iga_fd_t fd, void* fd_work, int ndim, const int64_t* box_n,
                       const int64_t* box_lo, const int64_t* box_len, void* z, void* w, double* ls, void* work,
                       double* hist_host) {
  return guard([&] {
    IGA_CHECK_ARG(out && op && packed && op_work && z && w && ls && work && hist_host, "bad argument");
    IGA_CHECK_ARG(n >= 1 && m >= 1, "bad size");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    IGA_CHECK_ARG(ndim >= 0 && ndim <= 8, "box dimension above 8");
    IGA_CHECK_ARG(ndim == 0 || (box_n && box_lo && box_len), "box arrays missing");
    IGA_CHECK_ARG(!fd || fd_work, "FD workspace missing");
    int64_t total = 1;
    for (int d = 0; d < ndim; ++d) {
      IGA_CHECK_ARG(box_lo[d] >= 0 && box_len[d] >= 0 && box_lo[d] + box_len[d] <= box_n[d], "box outside the grid");
      total *= box_n[d];
    }
    IGA_CHECK_ARG(ndim == 0 || total == n, "box grid does not match the vector length");
    int64_t offs[3];
    IGA_CHECK(iga_gmres_layout(m, offs));
    auto* c = new iga_arnoldi_s();
    c->dtype = dtype;
    c->n = n;
    c->m = m;
    c->op = op;
    c->packed = packed;
    c->nurbs_w = nurbs_w;
    c->op_work = op_work;
    c->fd = fd;
    c->fd_work = fd_work;
    c->ndim = ndim;
    for (int d = 0; d < ndim; ++d) {
      c->box_n[d] = box_n[d];
      c->box_lo[d] = box_lo[d];
      c->box_len[d] = box_len[d];
    }
    c->z = z;
    c->w = w;
    c->ls = ls;
    c->work = work;
    c->hist_host = hist_host;
    c->hist_off = offs[2];
    for (int i = 0; i < kEventRing; ++i) c->ev_step[i] = -1;
    *out = c;
  });
}

int iga_arnoldi_start(iga_arnoldi_t c, const void* r, void* V, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(c && r && V, "bad argument");
    hipStream_t s = as_stream(stream);
    IGA_CHECK(iga_gmres_start(c->dtype, c->n, c->m, r, V, c->ls, c->work, stream));
    record(c, 0, s);
  });
}

int iga_arnoldi_step(iga_arnoldi_t c, int k, void* V, int64_t ldv, int preconditioned, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(c && V && ldv >= c->n, "bad argument");
    IGA_CHECK_ARG(k >= 1 && k <= c->m + 1, "step outside the Krylov dimension");
    hipStream_t s = as_stream(stream);
    const size_t es = c->dtype == IGA_F64 ? 8 : 4;
    const void* v = (const char*)V + (size_t)(k - 1) * ldv * es;
    if (!preconditioned) {
      if (c->fd)
        IGA_CHECK(iga_fd_apply(c->fd, v, c->z, 1, c->fd_work, stream));
      else
        IGA_HIP(hipMemcpyAsync(c->z, v, (size_t)c->n * es, hipMemcpyDeviceToDevice, s));
    }
    IGA_CHECK(iga_fused_apply(c->op, c->packed, c->z, c->w, 1, c->nurbs_w, 1.0, 0, c->op_work, stream));
    if (c->ndim > 0)
      IGA_CHECK(iga_box_mask(c->dtype, c->ndim, c->box_n, c->box_lo, c->box_len, 1, c->w, stream));
    IGA_CHECK(iga_gmres_step(c->dtype, c->n, k, c->m, V, ldv, c->w, c->ls, c->work, stream));
    if (k >= 2) record(c, k - 1, s);
  });
}

int iga_arnoldi_query(iga_arnoldi_t c, int j) {
  int done = 0;
  const int rc = guard([&] {
    IGA_CHECK_ARG(c, "bad argument");
    const hipError_t e = hipEventQuery(event_of(c, j));
    if (e == hipSuccess)
      done = 1;
    else if (e != hipErrorNotReady)
      IGA_HIP(e);
  });
  return rc == IGA_OK ? done : rc;
}

int iga_arnoldi_wait(iga_arnoldi_t c, int j) {
  return guard([&] {
    IGA_CHECK_ARG(c, "bad argument");
    IGA_HIP(hipEventSynchronize(event_of(c, j)));
  });
}

int iga_arnoldi_destroy(iga_arnoldi_t c) {
  return guard([&] {
    if (!c) return;
    for (int i = 0; i < kEventRing; ++i)
      if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);  // teardown: nothing to report to
    delete c;
  });
}

}  // extern "C"
