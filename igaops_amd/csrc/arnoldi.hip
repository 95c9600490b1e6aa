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
//   w = A z, masked       the fused apply, its gather writing zeros outside the box
//                         (a caller that batches these two over several solves hands
//                         in z, or w, instead)
//   Arnoldi / Givens      the step of iga_gmres_step, whose small dense kernel also
//                         publishes the residual estimate to host-visible memory
//
// The host polls that memory (a tag identifies the solve, the entry index the step),
// so a step needs neither a copy nor an event.  The context fixes everything that
// does not change between steps (operator, preconditioner, box, the run's vectors and
// least-squares state, the host record); the basis is passed per step because it
// grows while the run proceeds.

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "internal.hpp"

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
  int masked = 0;
  int64_t box_lo[3] = {0, 0, 0}, box_len[3] = {0, 0, 0};
  void* z = nullptr;
  void* w = nullptr;
  double* ls = nullptr;
  void* work = nullptr;
  double* rec = nullptr;
  int64_t tag = 0;
};

using namespace iga;

extern "C" {

int iga_arnoldi_create(iga_arnoldi_t* out, int dtype, int64_t n, int m, iga_fused_t op, const void* packed,
                       const void* nurbs_w, void* op_work, iga_fd_t fd, void* fd_work, int ndim, const int64_t* box_n,
                       const int64_t* box_lo, const int64_t* box_len, void* z, void* w, double* ls, void* work,
                       double* host_rec, int64_t tag) {
  return guard([&] {
    IGA_CHECK_ARG(out && op && packed && op_work && z && w && ls && work && host_rec, "bad argument");
    IGA_CHECK_ARG(n >= 1 && m >= 1, "bad size");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    IGA_CHECK_ARG(ndim == 0 || ndim == 3, "the mask box of the fused operator has three directions");
    IGA_CHECK_ARG(ndim == 0 || (box_n && box_lo && box_len), "box arrays missing");
    IGA_CHECK_ARG(!fd || fd_work, "FD workspace missing");
    IGA_CHECK_ARG(tag >= 0 && (tag & 0xffffffffLL) == 0, "the tag's low 32 bits carry the entry index");
    int64_t total = 1;
    for (int d = 0; d < ndim; ++d) {
      IGA_CHECK_ARG(box_lo[d] >= 0 && box_len[d] >= 0 && box_lo[d] + box_len[d] <= box_n[d], "box outside the grid");
      total *= box_n[d];
    }
    IGA_CHECK_ARG(ndim == 0 || total == n, "box grid does not match the vector length");
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
    c->masked = ndim == 3;
    for (int d = 0; d < ndim; ++d) {
      c->box_lo[d] = box_lo[d];
      c->box_len[d] = box_len[d];
    }
    c->z = z;
    c->w = w;
    c->ls = ls;
    c->work = work;
    c->rec = host_rec;
    c->tag = tag;
    *out = c;
  });
}

int iga_arnoldi_start(iga_arnoldi_t c, const void* r, void* V, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(c && r && V, "bad argument");
    gmres_start_impl(c->dtype, c->n, c->m, r, V, c->ls, c->work, c->rec, c->tag, stream);
  });
}

int iga_arnoldi_step(iga_arnoldi_t c, int k, void* V, int64_t ldv, int preconditioned, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(c && V && ldv >= c->n, "bad argument");
    IGA_CHECK_ARG(k >= 1 && k <= c->m + 1, "step outside the Krylov dimension");
    IGA_CHECK_ARG(preconditioned >= 0 && preconditioned <= 2, "preconditioned is 0, 1 or 2");
    hipStream_t s = as_stream(stream);
    const size_t es = c->dtype == IGA_F64 ? 8 : 4;
    const void* v = (const char*)V + (size_t)(k - 1) * ldv * es;
    if (preconditioned == 0) {
      if (c->fd)
        IGA_CHECK(iga_fd_apply(c->fd, v, c->z, 1, c->fd_work, stream));
      else
        IGA_HIP(hipMemcpyAsync(c->z, v, (size_t)c->n * es, hipMemcpyDeviceToDevice, s));
    }
    if (preconditioned <= 1) {
      if (c->masked)
        fused_apply_masked(c->op, c->packed, c->z, c->w, c->nurbs_w, c->op_work, c->box_lo, c->box_len, stream);
      else
        IGA_CHECK(iga_fused_apply(c->op, c->packed, c->z, c->w, 1, c->nurbs_w, 1.0, 0, c->op_work, stream));
    }
    gmres_step_impl(c->dtype, c->n, k, c->m, V, ldv, c->w, c->ls, c->work, c->rec, c->tag, stream);
  });
}

int iga_arnoldi_destroy(iga_arnoldi_t c) {
  return guard([&] { delete c; });
}

}  // extern "C"
