// Functions shared between the library's source files; not part of the C ABI.
// They throw iga::Error (callers run them inside guard()).
#pragma once

#include <cstdint>

#include "common.hpp"

namespace iga {

// The increment criterion's record of a GMRES step: gram holds z_j . z_0..z_j (device),
// rec the host-visible record of the solve (rec == nullptr: none), tag its tag.
struct GmIncrement {
  const double* gram = nullptr;
  double* rec = nullptr;
  int64_t tag = 0;
};

// doubles of the increment record of a Krylov dimension m (krylov.hip)
int64_t gmres_increment_doubles(int m);

// iga_gmres_start / iga_gmres_step with the hist entries also published to
// host-visible memory `rec` under `tag` (krylov.hip; rec == nullptr: not published),
// and with `inc` the increment record entry of the step.
void gmres_start_impl(int dtype, int64_t n, int m, const void* r, void* V0, double* ls, void* work, double* rec,
                      int64_t tag, void* stream);
void gmres_step_impl(int dtype, int64_t n, int k, int m, void* V, int64_t ldv, void* w, double* ls, void* work,
                     double* rec, int64_t tag, void* stream, const GmIncrement* inc = nullptr);

// One GMRES orthogonalisation step (gmres_step_impl without the increment record) of
// each of up to kGmGroupMax solves of one vector type and length, as one launch per
// kernel of the step for all of them; the results are those of gmres_step_impl bit for
// bit.  Returns false, launching nothing, when the solves' basis layouts differ (the
// vector path of the many-dot pass must be the same for all).
constexpr int kGmGroupMax = 16;
struct GmGroupMember {
  int k = 0, m = 0;
  void* V = nullptr;
  int64_t ldv = 0;
  void* w = nullptr;
  double* ls = nullptr;
  void* work = nullptr;
  double* rec = nullptr;
  int64_t tag = 0;
};
bool gmres_step_group_impl(int dtype, int64_t n, int g, const GmGroupMember* mem, void* stream);

// iga_dot_many (krylov.hip): out[i] = V[i] . x for i < k
void dot_many_impl(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* x, double* out, void* work,
                   void* stream);

// iga_fused_apply whose output is zeroed outside the box [lo, lo + len) of the DOF grid
// (the Dirichlet mask, applied by the gather that writes y; fused.hip).
void fused_apply_masked(iga_fused_t f, const void* packed, const void* x, void* y, const void* nurbs_w, void* work,
                        const int64_t* box_lo, const int64_t* box_len, void* stream);

}  // namespace iga
