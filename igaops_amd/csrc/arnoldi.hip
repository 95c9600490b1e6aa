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

#include <mutex>
#include <vector>

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
  // increment criterion (iga_arnoldi_set_increment): the basis of preconditioned
  // vectors, the Gram row of the step (device) and the solve's increment record
  void* Z = nullptr;
  int64_t ldz = 0;
  double* gram = nullptr;
  double* inc_rec = nullptr;
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
    // with the increment criterion z_{k-1} = P v_{k-1} is kept as row k - 1 of Z
    void* z = c->Z ? (char*)c->Z + (size_t)(k - 1) * c->ldz * es : c->z;
    IGA_CHECK_ARG(!c->Z || preconditioned == 0, "the increment criterion needs the step's own FD and operator");
    if (preconditioned == 0) {
      if (c->fd)
        IGA_CHECK(iga_fd_apply(c->fd, v, z, 1, c->fd_work, stream));
      else
        IGA_HIP(hipMemcpyAsync(z, v, (size_t)c->n * es, hipMemcpyDeviceToDevice, s));
    }
    if (preconditioned <= 1) {
      if (c->masked)
        fused_apply_masked(c->op, c->packed, z, c->w, c->nurbs_w, c->op_work, c->box_lo, c->box_len, stream);
      else
        IGA_CHECK(iga_fused_apply(c->op, c->packed, z, c->w, 1, c->nurbs_w, 1.0, 0, c->op_work, stream));
    }
    if (c->Z) {
      // Gram row z_{k-1} . z_0..z_{k-1}, published with the step's increment record entry
      dot_many_impl(c->dtype, c->n, k, c->Z, c->ldz, z, c->gram, c->work, stream);
      GmIncrement inc;
      inc.gram = c->gram;
      inc.rec = c->inc_rec;
      inc.tag = c->tag;
      gmres_step_impl(c->dtype, c->n, k, c->m, V, ldv, c->w, c->ls, c->work, c->rec, c->tag, stream, &inc);
      return;
    }
    gmres_step_impl(c->dtype, c->n, k, c->m, V, ldv, c->w, c->ls, c->work, c->rec, c->tag, stream);
  });
}

int iga_arnoldi_destroy(iga_arnoldi_t c) {
  return guard([&] { delete c; });
}

int64_t iga_arnoldi_increment_doubles(int m) { return m >= 1 ? gmres_increment_doubles(m) : 0; }

int iga_arnoldi_set_increment(iga_arnoldi_t c, void* Z, int64_t ldz, double* gram, double* inc_rec) {
  return guard([&] {
    IGA_CHECK_ARG(c, "bad argument");
    IGA_CHECK_ARG(!Z || (gram && inc_rec && ldz >= c->n), "the increment criterion needs Z, the Gram row and the "
                                                         "record");
    c->Z = Z;
    c->ldz = ldz;
    c->gram = gram;
    c->inc_rec = inc_rec;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One Arnoldi step of several solves of one shape (same FD plan, operator plan, mask
// box, vector type and length) as one host call: their vectors V_i[k_i - 1] stacked one
// after the other in X, one multi-right-hand-side FD launch X -> Y, one grouped operator
// launch Y -> w_i (each solve's coefficients, outputs masked), then each solve's
// orthogonalisation on its own stream.  Events order the batch after each member's
// previous work and each member's orthogonalisation after the batch.
// ---------------------------------------------------------------------------

namespace {

constexpr int kStepGroupMax = 16;
static_assert(kStepGroupMax <= kGmGroupMax, "a group step's orthogonalisations fit one grouped launch");

template <typename T>
struct StackArgs {
  const T* src[kStepGroupMax];
};

// X[i * n + row] = src_i[row]: the members' vectors one after the other (grid row i)
template <typename T>
__global__ void __launch_bounds__(256) stack_rows_kernel(int64_t n, StackArgs<T> a, T* __restrict__ X) {
  const T* src = a.src[blockIdx.y];
  T* dst = X + (int64_t)blockIdx.y * n;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    dst[r] = src[r];
}

// events for ordering the streams of a group step, per device; an event may be
// recorded again once the waits on its previous record have been enqueued
std::mutex g_ev_mu;
std::vector<std::vector<hipEvent_t>> g_events;

hipEvent_t* group_events(int count) {
  int dev = 0;
  IGA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_ev_mu);
  if ((int)g_events.size() <= dev) g_events.resize(dev + 1);
  auto& pool = g_events[dev];
  while ((int)pool.size() < count) {
    hipEvent_t e;
    IGA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    pool.push_back(e);
  }
  return pool.data();
}

}  // namespace

extern "C" {

int64_t iga_arnoldi_group_workspace_bytes(iga_arnoldi_t c, int g) {
  if (!c || g < 1) return 0;
  const int64_t es = c->dtype == IGA_F64 ? 8 : 4;
  return 2 * c->n * g * es + 256;
}

int iga_arnoldi_step_group(int g, const iga_arnoldi_t* ctxs, const int* ks, void* const* Vs, const int64_t* ldvs,
                           void* const* streams, void* xy, void* fd_work, void* op_work, void* batch_stream) {
  return guard([&] {
    IGA_CHECK_ARG(g >= 1 && g <= kStepGroupMax, "a group step takes 1 to 16 solves");
    IGA_CHECK_ARG(ctxs && ks && Vs && ldvs && streams && xy && op_work, "NULL argument");
    const iga_arnoldi_s* c0 = ctxs[0];
    IGA_CHECK_ARG(c0 && c0->fd && fd_work, "a group step needs the FD preconditioner and its workspace");
    for (int i = 0; i < g; ++i) {
      const iga_arnoldi_s* c = ctxs[i];
      IGA_CHECK_ARG(c && Vs[i] && ldvs[i] >= c0->n, "bad member");
      IGA_CHECK_ARG(ks[i] >= 1 && ks[i] <= c->m + 1, "step outside the Krylov dimension");
      IGA_CHECK_ARG(!c->Z, "the increment criterion is not taken by group steps");
      IGA_CHECK_ARG(c->dtype == c0->dtype && c->n == c0->n && c->fd == c0->fd && c->op == c0->op &&
                        c->masked == c0->masked && c->nurbs_w == c0->nurbs_w,
                    "the members of a group step share the FD, the operator plan, the mask and the weights");
      for (int d = 0; d < 3; ++d)
        IGA_CHECK_ARG(c->box_lo[d] == c0->box_lo[d] && c->box_len[d] == c0->box_len[d], "members share the box");
    }
    hipStream_t bs = as_stream(batch_stream);
    hipEvent_t* ev = group_events(g + 1);
    for (int i = 0; i < g; ++i) {
      hipStream_t si = as_stream(streams[i]);
      if (si == bs) continue;
      IGA_HIP(hipEventRecord(ev[i], si));
      IGA_HIP(hipStreamWaitEvent(bs, ev[i], 0));
    }
    const int64_t n = c0->n;
    const size_t es = c0->dtype == IGA_F64 ? 8 : 4;
    void* X = xy;
    void* Y = (char*)xy + ((n * g * (int64_t)es + 255) & ~int64_t(255));
    const void* packed[kStepGroupMax];
    void* outs[kStepGroupMax];
    const dim3 sgrid(grid_for(n, 256, 4096), (unsigned)g);
    if (c0->dtype == IGA_F64) {
      StackArgs<double> a{};
      for (int i = 0; i < g; ++i) a.src[i] = (const double*)Vs[i] + (int64_t)(ks[i] - 1) * ldvs[i];
      stack_rows_kernel<double><<<sgrid, 256, 0, bs>>>(n, a, (double*)X);
    } else {
      StackArgs<float> a{};
      for (int i = 0; i < g; ++i) a.src[i] = (const float*)Vs[i] + (int64_t)(ks[i] - 1) * ldvs[i];
      stack_rows_kernel<float><<<sgrid, 256, 0, bs>>>(n, a, (float*)X);
    }
    IGA_LAUNCH_CHECK();
    // member-major throughout: every kernel reads and writes each member's vector contiguously
    IGA_CHECK(iga_fd_apply_rhs_major(c0->fd, X, Y, g, fd_work, batch_stream));
    for (int i = 0; i < g; ++i) {
      packed[i] = ctxs[i]->packed;
      outs[i] = ctxs[i]->w;
    }
    IGA_CHECK(iga_fused_apply_group(c0->op, g, packed, Y, 1, n, outs, c0->nurbs_w, op_work, c0->masked ? 3 : 0,
                                    c0->box_lo, c0->box_len, batch_stream));
    // the members' orthogonalisations: one launch per kernel of the step for all of
    // them on the batch stream, each member's stream ordered after it; members whose
    // basis layouts differ take their own steps on their own streams instead
    GmGroupMember mem[kStepGroupMax];
    for (int i = 0; i < g; ++i) {
      const iga_arnoldi_s* c = ctxs[i];
      mem[i].k = ks[i];
      mem[i].m = c->m;
      mem[i].V = Vs[i];
      mem[i].ldv = ldvs[i];
      mem[i].w = c->w;
      mem[i].ls = c->ls;
      mem[i].work = c->work;
      mem[i].rec = c->rec;
      mem[i].tag = c->tag;
    }
    const bool grouped = gmres_step_group_impl(c0->dtype, n, g, mem, batch_stream);
    IGA_HIP(hipEventRecord(ev[g], bs));
    for (int i = 0; i < g; ++i) {
      const iga_arnoldi_s* c = ctxs[i];
      hipStream_t si = as_stream(streams[i]);
      if (si != bs) IGA_HIP(hipStreamWaitEvent(si, ev[g], 0));
      if (!grouped)
        gmres_step_impl(c->dtype, c->n, ks[i], c->m, Vs[i], ldvs[i], c->w, c->ls, c->work, c->rec, c->tag,
                        streams[i]);
    }
  });
}

}  // extern "C"
