// Fused sum-factorised 3-D operator  y = sum_{a,b} T_a^T diag(D_ab) I_b x
// (trial interpolation I_b with the banded basis, point-wise 4x4 coefficient
// contraction D at the quadrature points, test contraction T_a with the banded
// weighted-quadrature weights), in one pass over the coefficient stream.
//
// Decomposition (all indices direction-1 fastest):
//   * a work-group owns a TQ1 x TQ2 tile of (q1, q2) quadrature columns and a
//     slab of q3 planes; thread (l1, l2) owns one column.
//   * sweeping q3 upwards, each thread keeps two register rings of P = p+1
//     DOF planes c3: the trial partials T2(q1,q2,c3) (3 fields) and the test
//     partials S2(q1,q2,c3) (3 fields).  A trial plane is built from the DOF
//     sub-block x(c1 in band(tile), c2 in band(tile), c3) staged in LDS; a test
//     plane leaves the ring when no later q3 touches it and is contracted in
//     directions 2 and 1 through LDS into a partial DOF block of the tile.
//   * partial blocks of neighbouring tiles / slabs overlap by the band halo; a
//     second kernel adds them in a fixed order (deterministic result).
// Coefficients are stored per (tile, q3, field) as contiguous runs over the
// tile's columns, so every load of the stream is a coalesced full-wave read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace iga {

// FORM_KM7X: the KM7 fields with the trial value taken from a second input x1 (the
// gradient from x): the space-time operator's stiffness on the time-interpolated
// vector and mass on the time-differentiated one in one sweep.
// FORM_GV4 / FORM_UG4: the first-order terms, D_i0 (test derivative i, trial value) /
// D_0j (test value, trial derivative j) plus D_00 (the NURBS correction, zero for
// B-splines): four fields per point.
enum FusedForm {
  FORM_M1 = 0, FORM_K6 = 1, FORM_KM7 = 2, FORM_SYM10 = 3, FORM_GEN16 = 4, FORM_KM7X = 5, FORM_GV4 = 6, FORM_UG4 = 7
};

__host__ __device__ constexpr int form_ncoef(int f) {
  return f == FORM_M1 ? 1
         : f == FORM_K6 ? 6
         : (f == FORM_KM7 || f == FORM_KM7X) ? 7
         : f == FORM_SYM10 ? 10
         : (f == FORM_GV4 || f == FORM_UG4) ? 4
                                            : 16;
}

constexpr int kTQ1 = 16, kTQ2 = 16, kThreads = kTQ1 * kTQ2;
// Max DOF footprint of a 16-point tile in directions 1, 2 (the plane products are
// 16 x 16): 16 points of rules with >= 3 points per element (WQ type "2", Gauss with
// p >= 2) span at most 7 elements, so at most p + 7 <= 14 DOFs for p <= 7.
constexpr int kDMax = 16;
constexpr int kKMax = 4;   // max tiles/slabs whose footprint contains one DOF index
constexpr int kSlabMax = 128;  // max q3 planes of one slab (direction-3 table staged in LDS)
constexpr int kGroupMax = 16;  // members of a grouped apply (coefficient sets of one launch)

template <typename T>
struct DirTab {
  const int32_t* first;  // [nq] band start
  const T* tab;          // [4][nq][P]: B0, B1, W0, W1 (W are the transposed test weights)
};

template <typename T>
struct FusedArgs {
  int nc[3], nq[3];
  int nt1, nt2, nslab, nrhs, nlogical;
  const int32_t *t1_q0, *t1_c0, *t1_nc;  // [nt1+1], [nt1], [nt1]
  const int32_t *t2_q0, *t2_c0, *t2_nc;
  const int32_t *s_q0, *s_c0, *s_nc;  // [nslab+1], [nslab], [nslab]
  const int64_t* coef_base;           // [nt1*nt2]
  const int64_t* part_base;           // [nt1*nt2*nslab]
  int64_t part_stride;                // elements per right-hand side
  // right-hand side k reads x[dof * x_dstride + k * x_rstride] and the coefficients
  // at coef + k * coef_rstride (0: one coefficient set shared by all of them)
  int64_t x_dstride, x_rstride, coef_rstride;
  const T* coef;
  // grouped apply (ngroup > 0): right-hand side k reads its own coefficient set gcoef[k]
  int ngroup;
  const T* gcoef[kGroupMax];
  const T* x;
  const T* x1;  // FORM_KM7X: the input of the value (mass) term, strides of x
  const T* w;   // NURBS weights or nullptr
  T* partial;
  DirTab<T> d[3];
};

// logical block id: the blocks dispatched to one XCD (round-robin over 8) get a
// contiguous logical range, so the right-hand sides of one tile-slab, adjacent
// in logical order, share that XCD's L2 for the coefficient stream.
__device__ inline int xcd_logical(int b, int nblk) {
  const int per = nblk / 8;
  return (b % 8) * per + b / 8;
}

template <typename T, int P, int FORM>
__device__ inline void point_contract(const T* __restrict__ c, int64_t stride, const T u[4], T f[4]) {
  if constexpr (FORM == FORM_M1) {
    f[0] = c[0] * u[0];
    f[1] = f[2] = f[3] = T(0);
  } else if constexpr (FORM == FORM_K6) {
    const T c11 = c[0], c22 = c[stride], c33 = c[2 * stride];
    const T c12 = c[3 * stride], c13 = c[4 * stride], c23 = c[5 * stride];
    f[0] = T(0);
    f[1] = c11 * u[1] + c12 * u[2] + c13 * u[3];
    f[2] = c12 * u[1] + c22 * u[2] + c23 * u[3];
    f[3] = c13 * u[1] + c23 * u[2] + c33 * u[3];
  } else if constexpr (FORM == FORM_KM7 || FORM == FORM_KM7X) {
    const T m = c[0], c11 = c[stride], c22 = c[2 * stride], c33 = c[3 * stride];
    const T c12 = c[4 * stride], c13 = c[5 * stride], c23 = c[6 * stride];
    f[0] = m * u[0];
    f[1] = c11 * u[1] + c12 * u[2] + c13 * u[3];
    f[2] = c12 * u[1] + c22 * u[2] + c23 * u[3];
    f[3] = c13 * u[1] + c23 * u[2] + c33 * u[3];
  } else if constexpr (FORM == FORM_GV4) {
    const T d00 = c[0], d10 = c[stride], d20 = c[2 * stride], d30 = c[3 * stride];
    f[0] = d00 * u[0];
    f[1] = d10 * u[0];
    f[2] = d20 * u[0];
    f[3] = d30 * u[0];
  } else if constexpr (FORM == FORM_UG4) {
    const T d00 = c[0], d01 = c[stride], d02 = c[2 * stride], d03 = c[3 * stride];
    f[0] = d00 * u[0] + d01 * u[1] + d02 * u[2] + d03 * u[3];
    f[1] = f[2] = f[3] = T(0);
  } else if constexpr (FORM == FORM_SYM10) {
    const T d00 = c[0], d11 = c[stride], d22 = c[2 * stride], d33 = c[3 * stride];
    const T d01 = c[4 * stride], d02 = c[5 * stride], d03 = c[6 * stride];
    const T d12 = c[7 * stride], d13 = c[8 * stride], d23 = c[9 * stride];
    f[0] = d00 * u[0] + d01 * u[1] + d02 * u[2] + d03 * u[3];
    f[1] = d01 * u[0] + d11 * u[1] + d12 * u[2] + d13 * u[3];
    f[2] = d02 * u[0] + d12 * u[1] + d22 * u[2] + d23 * u[3];
    f[3] = d03 * u[0] + d13 * u[1] + d23 * u[2] + d33 * u[3];
  } else {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      T acc = T(0);
#pragma unroll
      for (int b = 0; b < 4; ++b) acc += c[(a * 4 + b) * stride] * u[b];
      f[a] = acc;
    }
  }
}

// 16x16 products on the matrix cores for the plane contractions.  Lane layout of the
// 16x16x4 MFMA (measured on gfx950): A operand A[l % 16][k0 + l / 16], B operand
// B[k0 + l / 16][l % 16]; result register r of lane l holds row (l / 16) + 4r for FP64
// and 4 (l / 16) + r for FP32, column l % 16.
template <typename T>
struct PlaneMma;
template <>
struct PlaneMma<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  __device__ static V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct PlaneMma<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  __device__ static V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// acc += A B over K = 16: A row-major [16][lda] (rows M, columns K), B row-major
// [16][ldb] (rows K, columns N); four dependent MFMAs.
template <typename T>
__device__ inline typename PlaneMma<T>::V mm16(const T* A, int lda, const T* B, int ldb, typename PlaneMma<T>::V acc,
                                               int lane) {
  const int i = lane & 15, g = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) acc = PlaneMma<T>::mma(A[i * lda + k0 + g], B[(k0 + g) * ldb + i], acc);
  return acc;
}

template <typename T>
__device__ inline void st16(T* C, int ldc, typename PlaneMma<T>::V acc, int lane) {
#pragma unroll
  for (int r = 0; r < 4; ++r) C[PlaneMma<T>::row(lane, r) * ldc + (lane & 15)] = acc[r];
}

// NR right-hand sides per work-group (1 or 2; 2 only for forms with one input and
// shared coefficients): the coefficient loads, the band tables and the barriers of the
// plane phases serve NR right-hand sides, each with its own register rings and LDS
// planes.
template <typename T, int P, int FORM, int NR, int PD>
__device__ __attribute__((always_inline)) inline void fused_apply_body(const FusedArgs<T>& a) {
  using MM = PlaneMma<T>;
  using V = typename MM::V;
  // PD: prefetch depth (q3 planes) of the coefficient stream
  constexpr bool DUAL = FORM == FORM_KM7X;
  static_assert(NR == 1 || (NR == 2 && !DUAL), "two right-hand sides per work-group for one-input forms only");
  constexpr int NT1 = DUAL ? 3 : 2;  // direction-1 trial partials: B0 x, B1 x (, B0 x1)
  constexpr int NT2 = DUAL ? 4 : 3;  // trial plane fields
  constexpr int AP = 17;             // pitch of arrays read as the A operand (rows over lanes)
  // The footprint block of a DOF plane in LDS, zero outside D1 x D2 (set once)
  __shared__ T xs[NR][kTQ2][AP];
  __shared__ T xs1[DUAL ? kTQ2 : 1][AP];
  // dense band matrices of the tile (zero outside the band and past the tile):
  //   bd1[f][c1][l1] = B_f of direction 1 at point l1, DOF c1   (B operand, K = c1)
  //   bd2[f][l2][c2] = B_f of direction 2                        (A operand, K = c2)
  //   wd2[h][c2][m]  = W_h of direction 2 at point m, DOF c2      (A operand, K = m)
  //   wd1[h][l][c1]  = W_h of direction 1                        (B operand, K = l)
  __shared__ T bd1[2][16][16];
  __shared__ T bd2[2][16][AP];
  __shared__ T wd2[2][16][AP];
  __shared__ T wd1[2][16][16];
  __shared__ T s2s[NR][NT2 > 3 ? NT2 : 3][16][16];  // test values [m][l], then the trial plane [l2][l1]
  __shared__ T t1s[NR][NT1][16][16];                // direction-1 trial partials [c2][l1]
  __shared__ T r1s[NR][2][16][AP];                  // direction-2 test partials [c2][l]
  __shared__ T b3s[kSlabMax][4][P];             // direction-3 band values of the slab's planes

  const int L = xcd_logical(blockIdx.x, gridDim.x);
  if (L >= a.nlogical) return;  // padding blocks of the XCD remap (whole block exits)
  int rem = L;
  const int nrb = a.nrhs / NR;  // work-groups per tile-slab (the host ensures NR | nrhs)
  const int k = (rem % nrb) * NR;  // first right-hand side of this work-group
  rem /= nrb;
  const int s = rem % a.nslab;
  rem /= a.nslab;
  const int t1 = rem % a.nt1;
  const int t2 = rem / a.nt1;
  const int tile = t2 * a.nt1 + t1;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l1 = tid % kTQ1, l2 = tid / kTQ1;
  const int q1s = a.t1_q0[t1], q1n = a.t1_q0[t1 + 1] - q1s;
  const int q2s = a.t2_q0[t2], q2n = a.t2_q0[t2 + 1] - q2s;
  const int c1lo = a.t1_c0[t1], D1 = a.t1_nc[t1];
  const int c2lo = a.t2_c0[t2], D2 = a.t2_nc[t2];
  const int q3s = a.s_q0[s], q3e = a.s_q0[s + 1];
  const int c3lo = a.s_c0[s];
  const bool active = (l1 < q1n) && (l2 < q2n);
  const int n1 = a.nc[0], n2 = a.nc[1];

  // dense band matrices of this tile: thread (row r = tid / 16, column c = tid % 16)
  {
    const int r = tid >> 4, c = tid & 15;
    // direction 1, point c, DOF r (bd1) / point r, DOF c (wd1)
    {
      const int l = c, cd = r;
      const int f = l < q1n ? a.d[0].first[q1s + l] - c1lo : -P;
      const int j = cd - f;
      const bool in = l < q1n && j >= 0 && j < P;
      const int64_t o = (int64_t)(q1s + (l < q1n ? l : 0)) * P + (in ? j : 0);
      bd1[0][cd][l] = in ? a.d[0].tab[o] : T(0);
      bd1[1][cd][l] = in ? a.d[0].tab[(int64_t)a.nq[0] * P + o] : T(0);
    }
    {
      const int l = r, cd = c;
      const int f = l < q1n ? a.d[0].first[q1s + l] - c1lo : -P;
      const int j = cd - f;
      const bool in = l < q1n && j >= 0 && j < P;
      const int64_t o = (int64_t)(q1s + (l < q1n ? l : 0)) * P + (in ? j : 0);
      wd1[0][l][cd] = in ? a.d[0].tab[2 * (int64_t)a.nq[0] * P + o] : T(0);
      wd1[1][l][cd] = in ? a.d[0].tab[3 * (int64_t)a.nq[0] * P + o] : T(0);
    }
    // direction 2, point r, DOF c (bd2) / point c, DOF r (wd2)
    {
      const int m = r, cd = c;
      const int f = m < q2n ? a.d[1].first[q2s + m] - c2lo : -P;
      const int j = cd - f;
      const bool in = m < q2n && j >= 0 && j < P;
      const int64_t o = (int64_t)(q2s + (m < q2n ? m : 0)) * P + (in ? j : 0);
      bd2[0][m][cd] = in ? a.d[1].tab[o] : T(0);
      bd2[1][m][cd] = in ? a.d[1].tab[(int64_t)a.nq[1] * P + o] : T(0);
    }
    {
      const int m = c, cd = r;
      const int f = m < q2n ? a.d[1].first[q2s + m] - c2lo : -P;
      const int j = cd - f;
      const bool in = m < q2n && j >= 0 && j < P;
      const int64_t o = (int64_t)(q2s + (m < q2n ? m : 0)) * P + (in ? j : 0);
      wd2[0][cd][m] = in ? a.d[1].tab[2 * (int64_t)a.nq[1] * P + o] : T(0);
      wd2[1][cd][m] = in ? a.d[1].tab[3 * (int64_t)a.nq[1] * P + o] : T(0);
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) xs[q][r][c] = T(0);
    if constexpr (DUAL) xs1[r][c] = T(0);
  }
  for (int i = tid; i < (q3e - q3s) * 4 * P; i += kThreads) {
    const int lq = i / (4 * P), m = (i / P) % 4, j = i % P;
    b3s[lq][m][j] = a.d[2].tab[((int64_t)m * a.nq[2] + q3s + lq) * P + j];
  }
  __syncthreads();

  T tr[NR][NT2][P];  // trial partials of the window planes: (val,val), (d1,val), (val,d2) [, (val,val) of x1]
  T te[NR][3][P];    // test partials: (val,val), (d1,val), (val,d2) in directions (1,2)

  // x block of one DOF plane, held in registers between its (early) load and its
  // use, so that the load latency overlaps the work of the current step.
  // Loads are unconditional from clamped (always valid) addresses so that the
  // compiler does not wait on them at the point of issue; NURBS weights are kept
  // apart and applied when the block is used.
  constexpr int XR = (kDMax * kDMax + kThreads - 1) / kThreads;
  const T* xk[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) xk[q] = a.x + (int64_t)(k + q) * a.x_rstride;
  const T* xk1 = DUAL ? a.x1 + (int64_t)k * a.x_rstride : xk[0];
  const int64_t xds = a.x_dstride;
  const bool nurbs = a.w != nullptr;
  // Per-thread offsets of its x-block elements within a DOF plane and their LDS
  // slots, computed once (no division in the sweep).  Lanes without an element
  // load the block's first entry (a valid address) and do not store it.
  // The same (i2, i1) split of the footprint serves the flush of test planes.
  int64_t xoff[XR];
  int xij[XR];  // (i2 << 8) | i1, or -1 past the footprint
#pragma unroll
  for (int e = 0; e < XR; ++e) {
    const int i = tid + e * kThreads;
    const bool ok = i < D1 * D2;
    const int ii = ok ? i : 0;
    const int i2 = ii / D1, i1 = ii % D1;
    xoff[e] = (int64_t)(c1lo + i1) + (int64_t)n1 * (c2lo + i2);
    xij[e] = ok ? (i2 << 8) | i1 : -1;
  }
  const int64_t plane = (int64_t)n1 * n2;
  // the second input's registers exist only in the two-input form
  struct XBlock1 {
    T x[NR][XR], w[XR];
  };
  struct XBlock2 {
    T x[NR][XR], w[XR], x1[XR];
  };
  using XBlock = std::conditional_t<DUAL, XBlock2, XBlock1>;
  struct Trial {
    T v[NR][NT2];
  };
  auto load_x = [&](int c3) __attribute__((always_inline)) {
    XBlock b;
#pragma unroll
    for (int e = 0; e < XR; ++e) {
      const int64_t dof = xoff[e] + plane * c3;
#pragma unroll
      for (int q = 0; q < NR; ++q) b.x[q][e] = xk[q][dof * xds];
      if constexpr (DUAL) b.x1[e] = xk1[dof * xds];
      b.w[e] = nurbs ? a.w[dof] : T(1);
    }
    return b;
  };

  // One LDS phase of the sweep.  It may (a) contract the oldest test plane (ring
  // slot 0, DOF plane c3f) in directions 2 and 1 into the tile's partial block and
  // (b) build the trial partials of a new DOF plane from an x block held in
  // registers.  Every contraction is a 16x16 product on the matrix cores:
  //   stage 1: T1_f = X Bd1_f (per trial field f), R1_0 = Wd2_0 S_0 + Wd2_1 S_2,
  //            R1_1 = Wd2_0 S_1
  //   stage 2: T2 = Bd2_a T1_b (the trial plane fields), partial = R1_0 Wd1_0 + R1_1 Wd1_1
  // with the products spread over the four waves.
  T* part[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q)
    part[q] = a.partial + (int64_t)(k + q) * a.part_stride + a.part_base[(int64_t)tile * a.nslab + s];
  // the flush of right-hand side q's oldest test plane: R1 Wd1 into its partial block
  auto flush_block = [&](int q, int c3f) __attribute__((always_inline)) {
    V acc = mm16<T>(&r1s[q][0][0][0], AP, &wd1[0][0][0], 16, V{0, 0, 0, 0}, lane);
    acc = mm16<T>(&r1s[q][1][0][0], AP, &wd1[1][0][0], 16, acc, lane);
    T* blk = part[q] + (int64_t)(c3f - c3lo) * D2 * D1;
    const int c1 = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c2 = MM::row(lane, r);
      if (c1 < D1 && c2 < D2) blk[c2 * D1 + c1] = acc[r];
    }
  };
  // trial plane field g of right-hand side q: fields (B0, T1_0), (B0, T1_1), (B1, T1_0)
  // [, (B0, T1_2)]
  auto field = [&](int q, int g) __attribute__((always_inline)) {
    const int fa = g == 2 ? 1 : 0, fb = g == 1 ? 1 : g == 3 ? 2 : 0;
    V acc = mm16<T>(&bd2[fa][0][0], AP, &t1s[q][fb][0][0], 16, V{0, 0, 0, 0}, lane);
    st16<T>(&s2s[q][g][0][0], 16, acc, lane);
  };
  auto plane_phase = [&](bool do_flush, int c3f, bool do_trial, XBlock xb) __attribute__((always_inline)) {
    Trial out{};
    __syncthreads();  // LDS buffers are reused
    if (do_flush) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        s2s[q][0][l2][l1] = active ? te[q][0][0] : T(0);
        s2s[q][1][l2][l1] = active ? te[q][1][0] : T(0);
        s2s[q][2][l2][l1] = active ? te[q][2][0] : T(0);
      }
    }
    if (do_trial) {
#pragma unroll
      for (int e = 0; e < XR; ++e) {
        if (xij[e] >= 0) {
          const int i2 = xij[e] >> 8, i1 = xij[e] & 255;
#pragma unroll
          for (int q = 0; q < NR; ++q) xs[q][i2][i1] = nurbs ? xb.x[q][e] * xb.w[e] : xb.x[q][e];
          if constexpr (DUAL) xs1[i2][i1] = nurbs ? xb.x1[e] * xb.w[e] : xb.x1[e];
        }
      }
    }
    __syncthreads();
    // stage 1: trial T1_f = X Bd1_f; test R1_0 = Wd2_0 S_0 + Wd2_1 S_2, R1_1 = Wd2_0 S_1
    if constexpr (NR == 1) {
      if (do_trial) {
        if (wave < NT1) {
          const T* X = (DUAL && wave == 2) ? &xs1[0][0] : &xs[0][0][0];
          const int f = wave == 1 ? 1 : 0;
          V acc = mm16<T>(X, AP, &bd1[f][0][0], 16, V{0, 0, 0, 0}, lane);
          st16<T>(&t1s[0][wave][0][0], 16, acc, lane);
        }
      }
      if (do_flush) {
        if (wave == 3) {
          V acc = mm16<T>(&wd2[0][0][0], AP, &s2s[0][0][0][0], 16, V{0, 0, 0, 0}, lane);
          acc = mm16<T>(&wd2[1][0][0], AP, &s2s[0][2][0][0], 16, acc, lane);
          st16<T>(&r1s[0][0][0][0], AP, acc, lane);
        } else if (wave == 2) {
          V acc = mm16<T>(&wd2[0][0][0], AP, &s2s[0][1][0][0], 16, V{0, 0, 0, 0}, lane);
          st16<T>(&r1s[0][1][0][0], AP, acc, lane);
        }
      }
    } else {
      // wave w: trial (q, f) = (w / 2, w % 2); test R1_0 of q = w (waves 0, 1) or R1_1 of
      // q = w - 2 (waves 2, 3)
      if (do_trial) {
        const int q = wave >> 1, f = wave & 1;
        V acc = mm16<T>(&xs[q][0][0], AP, &bd1[f][0][0], 16, V{0, 0, 0, 0}, lane);
        st16<T>(&t1s[q][f][0][0], 16, acc, lane);
      }
      if (do_flush) {
        if (wave < 2) {
          V acc = mm16<T>(&wd2[0][0][0], AP, &s2s[wave][0][0][0], 16, V{0, 0, 0, 0}, lane);
          acc = mm16<T>(&wd2[1][0][0], AP, &s2s[wave][2][0][0], 16, acc, lane);
          st16<T>(&r1s[wave][0][0][0], AP, acc, lane);
        } else {
          V acc = mm16<T>(&wd2[0][0][0], AP, &s2s[wave - 2][1][0][0], 16, V{0, 0, 0, 0}, lane);
          st16<T>(&r1s[wave - 2][1][0][0], AP, acc, lane);
        }
      }
    }
    __syncthreads();
    // stage 2: the trial plane fields T2 = Bd2 T1, the flush R1 Wd1 into the partial block
    if constexpr (NR == 1) {
      if (do_trial) {
        if (wave < 3) field(0, wave);
        if constexpr (DUAL) {
          if (wave == 0) field(0, 3);
        }
      }
      if (do_flush && wave == 3) flush_block(0, c3f);
    } else {
      // waves 0, 1: field 0 / 1 of both right-hand sides; waves 2, 3: field 2 of one and
      // the flush of the other
      if (do_trial) {
        if (wave < 2) {
          field(0, wave);
          field(1, wave);
        } else {
          field(wave - 2, 2);
        }
      }
      if (do_flush && wave >= 2) flush_block(3 - wave, c3f);
    }
    if (do_trial) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int g = 0; g < NT2; ++g) out.v[q][g] = active ? s2s[q][g][l2][l1] : T(0);
    }
    return out;
  };

  const int32_t* first3 = a.d[2].first;
  const int c3hi = c3lo + a.s_nc[s] - 1;  // last DOF plane this slab touches
  int cur = first3[q3s];
  {
    // all P window planes: issue every load first, then build them one by one
    XBlock xw[P];
#pragma unroll
    for (int j = 0; j < P; ++j) xw[j] = load_x(cur + j);
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const Trial t = plane_phase(false, 0, true, xw[j]);
#pragma unroll
      for (int q = 0; q < NR; ++q) {
#pragma unroll
        for (int g = 0; g < NT2; ++g) tr[q][g][j] = t.v[q][g];
        te[q][0][j] = te[q][1][j] = te[q][2][j] = T(0);
      }
    }
  }
  // x block of the next plane to enter the trial window
  XBlock xnext = load_x(cur + P <= c3hi ? cur + P : cur);

  const int npts = q1n * q2n;
  // inactive lanes read the tile's first point (a valid address) and discard it
  const T* coefk = a.ngroup > 0 ? a.gcoef[k] : a.coef + (int64_t)k * a.coef_rstride;
  const T* cptr = coefk + a.coef_base[tile] + (active ? (l1 + q1n * l2) : 0);
  constexpr int NCF = form_ncoef(FORM);
  const int64_t step_stride = (int64_t)NCF * npts;
  // Coefficients of plane q3 are loaded PD planes ahead into a ring of buffers;
  // the loop is unrolled PD times so that buffer b is always addressed statically
  // and no register with a load in flight is ever copied.
  T cbuf[PD][NCF];
  int fbuf[PD];
  // The coefficient stream is loaded non-temporally (it is read once per launch when
  // there is one right-hand side): 8% faster on the p=4 n=64 sweep, 5% slower when 16
  // right-hand sides share one stream through the L2.  A uniform run-time choice
  // between the two load forms lost the gain (the branch serialises the loads).
  auto load_step = [&](int q3, T (&c)[NCF], int& fb) __attribute__((always_inline)) {
    const T* p = cptr + (int64_t)q3 * step_stride;
#pragma unroll
    for (int f = 0; f < NCF; ++f) c[f] = __builtin_nontemporal_load(p + f * npts);
    fb = first3[q3];
  };
#pragma unroll
  for (int b = 0; b < PD; ++b)
    if (q3s + b < q3e) load_step(q3s + b, cbuf[b], fbuf[b]);

  // the oldest test plane leaves the window, the next trial plane enters it
  auto advance = [&]() __attribute__((always_inline)) {
    const Trial t = plane_phase(true, cur, true, xnext);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
      for (int j = 0; j < P - 1; ++j) {
#pragma unroll
        for (int g = 0; g < 3; ++g) te[q][g][j] = te[q][g][j + 1];
#pragma unroll
        for (int g = 0; g < NT2; ++g) tr[q][g][j] = tr[q][g][j + 1];
      }
      te[q][0][P - 1] = te[q][1][P - 1] = te[q][2][P - 1] = T(0);
#pragma unroll
      for (int g = 0; g < NT2; ++g) tr[q][g][P - 1] = t.v[q][g];
    }
    ++cur;
    // unconditional (clamped plane) so that the load queue has the same shape on every path
    xnext = load_x(cur + P <= c3hi ? cur + P : c3hi);
  };
  auto step = [&](int q3, T (&c)[NCF], int& fb) __attribute__((always_inline)) {
    const int f = fb;
    // Band starts grow by at most one plane per q3 point for the rules in use, so the
    // first advance is peeled: between two advances of that path at least one
    // coefficient refill is issued, and the wait for the x block leaves the newest
    // coefficient loads in flight.  Further advances of the same step (the loop) are
    // rare and may drain the load queue.
    if (cur < f) {
      advance();
      while (cur < f) advance();
    }
    if (active) {
      const T (&tb)[4][P] = b3s[q3 - q3s];
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        T u[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int j = 0; j < P; ++j) {
          u[0] += tb[0][j] * tr[q][DUAL ? NT2 - 1 : 0][j];  // the value: from x1 for FORM_KM7X
          u[1] += tb[0][j] * tr[q][1][j];
          u[2] += tb[0][j] * tr[q][2][j];
          u[3] += tb[1][j] * tr[q][0][j];
        }
        T fv[4];
        point_contract<T, P, FORM>(c, 1, u, fv);
#pragma unroll
        for (int j = 0; j < P; ++j) {
          te[q][0][j] += tb[2][j] * fv[0] + tb[3][j] * fv[3];
          te[q][1][j] += tb[2][j] * fv[1];
          te[q][2][j] += tb[2][j] * fv[2];
        }
      }
    }
    // refill this buffer with the plane PD steps ahead (past the slab end: a valid plane,
    // unused, so that every path issues the same loads)
    load_step(q3 + PD < q3e ? q3 + PD : q3e - 1, c, fb);
  };
  for (int q3 = q3s; q3 < q3e; q3 += PD) {
#pragma unroll
    for (int b = 0; b < PD; ++b)
      if (q3 + b < q3e) step(q3 + b, cbuf[b], fbuf[b]);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    plane_phase(true, cur, false, xnext);
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int jj = 0; jj < P - 1; ++jj) {
#pragma unroll
        for (int g = 0; g < 3; ++g) te[q][g][jj] = te[q][g][jj + 1];
      }
    ++cur;
  }
}

// Coefficient prefetch depth (q3 planes) of the FP64 variant without an occupancy target,
// which has registers to spare at two waves per SIMD.
#ifndef IGA_FUSED_WIDE_PD_F64
#define IGA_FUSED_WIDE_PD_F64 2
#endif
constexpr int kWidePrefetchF64 = IGA_FUSED_WIDE_PD_F64;

// Coefficient prefetch depth (q3 planes) of the FP32 sweep: two for the ten-field NURBS
// form up to degree 5 (p=4 n=64: 78.3 -> 75.3 us, p=5: 88.0 -> 78.0 us against three),
// three otherwise (the unit cube's seven fields at p=4 and the NURBS form at p=6 were
// 1.5-4% slower with two).
template <int P, int FORM>
constexpr int prefetch_f32() {
  return FORM == FORM_SYM10 && P <= 6 ? 2 : 3;
}

// Two launch variants of the same sweep: with an occupancy target of three waves
// per SIMD for FP64 and four for FP32 (used when the band is narrow enough to fit
// without spilling) and without (wider bands, one wave per SIMD fewer).
template <typename T, int P, int FORM, int NR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3, 8)))
fused_apply_kernel(FusedArgs<T> a) {
  fused_apply_body<T, P, FORM, NR, sizeof(T) == 8 ? 2 : prefetch_f32<P, FORM>()>(a);
}

template <typename T, int P, int FORM, int NR>
__global__ void __launch_bounds__(kThreads) fused_apply_kernel_wide(FusedArgs<T> a) {
  fused_apply_body<T, P, FORM, NR, sizeof(T) == 8 ? kWidePrefetchF64 : prefetch_f32<P, FORM>()>(a);
}

// forms with a two-right-hand-side variant (the solver and benchmark forms)
__host__ __device__ constexpr bool form_has_nr2(int f) { return f == FORM_K6 || f == FORM_KM7 || f == FORM_SYM10; }

// y[a*r + k] = alpha * wscale[a] * sum of the partial blocks covering DOF a (+ y)
template <typename T>
struct GatherArgs {
  int nc[3], nt1, nt2, nslab, nrhs;
  // [nc_d], [nc_d][kKMax] x2; in directions 1 and 2 a tile slot holds
  // t | (DOF count of tile t) << 16, in direction 3 the slab index
  const int32_t *l1_cnt, *l1_t, *l1_off;
  const int32_t *l2_cnt, *l2_t, *l2_off;
  const int32_t *l3_cnt, *l3_t, *l3_off;
  int64_t part_blk;     // elements of every (tile, slab) partial block: block b starts at b * part_blk
  int64_t part_stride;
  const T* partial;
  const T* w;
  T* y;
  int64_t y_dstride, y_rstride;  // y[dof * y_dstride + k * y_rstride]
  T alpha;
  int accumulate;
  int box;               // zero y outside the DOF box [blo, bhi) (the Dirichlet mask)
  int blo[3], bhi[3];
  // grouped apply (ngroup > 0): right-hand side k = blockIdx.y is written to gy[k][dof]
  int ngroup;
  T* gy[kGroupMax];
};

template <typename T>
__global__ void __launch_bounds__(256) fused_gather_kernel(GatherArgs<T> g) {
  const int64_t ndof = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
  // grouped: one grid row per right-hand side, DOF fastest
  const int64_t total = g.ngroup > 0 ? ndof : ndof * g.nrhs;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    // right-hand side fastest for interleaved y, DOF fastest for sliced y (coalesced stores)
    const bool interleaved = g.y_rstride == 1;
    const int k = g.ngroup > 0 ? (int)blockIdx.y : interleaved ? (int)(idx % g.nrhs) : (int)(idx / ndof);
    const int64_t dof = g.ngroup > 0 ? idx : interleaved ? idx / g.nrhs : idx % ndof;
    T* yk = g.ngroup > 0 ? g.gy[k] : g.y;
    const int64_t yi = g.ngroup > 0 ? dof : dof * g.y_dstride + (int64_t)k * g.y_rstride;
    const int c1 = (int)(dof % g.nc[0]);
    const int c2 = (int)((dof / g.nc[0]) % g.nc[1]);
    const int c3 = (int)(dof / ((int64_t)g.nc[0] * g.nc[1]));
    const T* pk = g.partial + (int64_t)k * g.part_stride;
    // All index loads first, then all partial loads, every load unconditional (the
    // tables hold kKMax slots per DOF, zero past the count): the dependent chain is
    // three loads deep, and no load waits for the one before it.
    const int n1 = g.l1_cnt[c1], n2 = g.l2_cnt[c2], n3 = g.l3_cnt[c3];
    int t1v[kKMax], o1v[kKMax], D1v[kKMax], t2v[kKMax], o2v[kKMax], D2v[kKMax], sv[kKMax], o3v[kKMax];
#pragma unroll
    for (int i = 0; i < kKMax; ++i) {
      const int a1 = g.l1_t[c1 * kKMax + i], b1 = g.l1_off[c1 * kKMax + i];
      const int a2 = g.l2_t[c2 * kKMax + i], b2 = g.l2_off[c2 * kKMax + i];
      const int a3 = g.l3_t[c3 * kKMax + i], b3 = g.l3_off[c3 * kKMax + i];
      t1v[i] = i < n1 ? (a1 & 0xffff) : 0;
      D1v[i] = i < n1 ? (a1 >> 16) : 0;
      o1v[i] = i < n1 ? b1 : 0;
      t2v[i] = i < n2 ? (a2 & 0xffff) : 0;
      D2v[i] = i < n2 ? (a2 >> 16) : 0;
      o2v[i] = i < n2 ? b2 : 0;
      sv[i] = i < n3 ? a3 : 0;
      o3v[i] = i < n3 ? b3 : 0;
    }
    T acc = T(0);
    if (n1 <= 2 && n2 <= 2 && n3 <= 2) {
      // up to two covering tiles per direction (the common case): the eight
      // candidate contributions, absent ones read at slot 0 and skipped in the sum
      int64_t addr[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int i1 = c & 1, i2 = (c >> 1) & 1, i3 = c >> 2;
        const int64_t base = (((int64_t)t2v[i2] * g.nt1 + t1v[i1]) * g.nslab + sv[i3]) * g.part_blk;
        addr[c] = base + ((int64_t)o3v[i3] * D2v[i2] + o2v[i2]) * D1v[i1] + o1v[i1];
      }
      T val[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) val[c] = pk[addr[c]];
      // the order of the general loop below: i3, then i2, then i1
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int i1 = c & 1, i2 = (c >> 1) & 1, i3 = c >> 2;
        if (i1 < n1 && i2 < n2 && i3 < n3) acc += val[c];
      }
    } else {
      for (int i3 = 0; i3 < n3; ++i3) {
        for (int i2 = 0; i2 < n2; ++i2) {
          for (int i1 = 0; i1 < n1; ++i1) {
            const int64_t base = (((int64_t)t2v[i2] * g.nt1 + t1v[i1]) * g.nslab + sv[i3]) * g.part_blk;
            acc += pk[base + ((int64_t)o3v[i3] * D2v[i2] + o2v[i2]) * D1v[i1] + o1v[i1]];
          }
        }
      }
    }
    T v = g.alpha * acc;
    if (g.w) v *= g.w[dof];
    if (g.accumulate) v += yk[yi];
    if (g.box && (c1 < g.blo[0] || c1 >= g.bhi[0] || c2 < g.blo[1] || c2 >= g.bhi[1] || c3 < g.blo[2] ||
                  c3 >= g.bhi[2]))
      v = T(0);
    yk[yi] = v;
  }
}

// Packs reference-layout coefficients into the fused layout of one FORM.
//   C: (3,3,nq) with C[i][j] multiplying trial d_j and test d_i; m: (nq);
//   W, dW: NURBS weight function and its gradient at the points (nullptr for
//   B-splines).  NURBS folding (operators/nurbs_ops.py:343-390):
//   D_ij = C_ij / W^2, D_00 = z^T C z + m / W^2, D_0j = -(C^T z)_j / W,
//   D_i0 = -(C z)_i / W with z = dW / W^2.
// Coefficients of a 2-D cross-section map extruded along xi_3, generated at the
// points instead of read from C and m (q = plane + n_plane q3, plane = q1 + nq1 q2):
//   det = dz[q3] det2[plane],  kappa = exp(sigma Re sum_k Ez[q3][k] H[plane][k])
//   C_ab = kappa det inv_ab[plane] (a, b < 2, symmetric),  C_22 = kappa det / dz^2,
//   m = det (with_mass) or 0.
// plane holds det2, inv_00, inv_11, inv_01 (4 x n_plane); H (n_plane x modes) and
// Ez (n3 x modes) are complex (re, im interleaved); H == nullptr means kappa = 1.
// With this source, W (n_plane) and dW (2 x n_plane: d/dxi_1, d/dxi_2) are fields of the
// cross-section.
struct ExtrudedSource {
  const double* plane = nullptr;
  const double* dz = nullptr;
  const double* H = nullptr;
  const double* Ez = nullptr;
  int modes = 0;
  int with_mass = 0;
  double sigma = 0.0;
};

template <typename T>
struct PackArgs {
  int nq[3], nt1, form, ncoef;
  int tq1, tq2;
  const int32_t *t1_q0, *t2_q0;
  const int64_t* coef_base;
  const double* C;
  const double* m;
  const double* W;
  const double* dW;
  int64_t comp_stride;  // elements between the components of C and of dW
  ExtrudedSource ex;    // used instead of C and m when ex.plane is set
  T* out;
};

template <typename T>
__global__ void fused_pack_kernel(PackArgs<T> p) {
  const int64_t nqt = (int64_t)p.nq[0] * p.nq[1] * p.nq[2];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nqt; q += (int64_t)gridDim.x * blockDim.x) {
    const int q1 = (int)(q % p.nq[0]);
    const int q2 = (int)((q / p.nq[0]) % p.nq[1]);
    const int64_t q3 = q / ((int64_t)p.nq[0] * p.nq[1]);
    const int t1 = q1 / p.tq1, t2 = q2 / p.tq2;
    const int l1 = q1 - p.t1_q0[t1], l2 = q2 - p.t2_q0[t2];
    const int q1n = p.t1_q0[t1 + 1] - p.t1_q0[t1];
    const int q2n = p.t2_q0[t2 + 1] - p.t2_q0[t2];
    const int64_t npts = (int64_t)q1n * q2n;
    T* dst = p.out + p.coef_base[t2 * p.nt1 + t1] + q3 * p.ncoef * npts + (l1 + (int64_t)q1n * l2);
    double D[4][4] = {{0}};
    double c[3][3] = {{0}};
    double mm = 0.0;
    // NURBS weight function and its gradient at this point (rational: p.W set)
    double wq = 1.0, dwq[3] = {0.0, 0.0, 0.0};
    if (p.ex.plane) {
      const int64_t nplane = (int64_t)p.nq[0] * p.nq[1];
      const int64_t pl = q1 + (int64_t)p.nq[0] * q2;
      if (p.W) {
        // cross-section fields: the control weights repeat along xi_3, so W does not
        // depend on xi_3 and its xi_3 derivative is zero
        wq = p.W[pl];
        dwq[0] = p.dW[pl];
        dwq[1] = p.dW[nplane + pl];
      }
      const double dz = p.ex.dz[q3];
      const double det = dz * p.ex.plane[pl];
      double kap = 1.0;
      if (p.ex.H) {
        const double* h = p.ex.H + 2 * pl * p.ex.modes;
        const double* e = p.ex.Ez + 2 * q3 * p.ex.modes;
        double g = 0.0;
        for (int k = 0; k < p.ex.modes; ++k) g += h[2 * k] * e[2 * k] - h[2 * k + 1] * e[2 * k + 1];
        kap = exp(p.ex.sigma * g);
      }
      const double kd = kap * det;
      c[0][0] = kd * p.ex.plane[nplane + pl];
      c[1][1] = kd * p.ex.plane[2 * nplane + pl];
      c[0][1] = c[1][0] = kd * p.ex.plane[3 * nplane + pl];
      c[2][2] = kd / (dz * dz);
      mm = p.ex.with_mass ? det : 0.0;
    } else if (p.form == FORM_GV4 || p.form == FORM_UG4) {
      // C holds the first-order coefficients c (3, nq)
      double cv[3], cz = 0.0;
      for (int i = 0; i < 3; ++i) cv[i] = p.C[(int64_t)i * p.comp_stride + q];
      double wi2 = 1.0;
      if (p.W) {
        // rational basis (operators/nurbs_ops.py:392-430): c W~^2 on the B-spline
        // first-order term, -(c . z) W~ on the mass term, z = grad W W~^2
        const double wi = 1.0 / p.W[q];
        wi2 = wi * wi;
        for (int i = 0; i < 3; ++i) cz += cv[i] * p.dW[(int64_t)i * p.comp_stride + q] * wi2;
        cz *= -wi;
      }
      // both forms hold (D_00, c_1, c_2, c_3): the test derivative (GV4) or the trial
      // derivative (UG4) of c_i is the kernel's choice
      dst[0] = (T)cz;
      for (int i = 0; i < 3; ++i) dst[(i + 1) * npts] = (T)(cv[i] * wi2);
      continue;
    } else {
      if (p.C)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) c[i][j] = p.C[(int64_t)(i * 3 + j) * p.comp_stride + q];
      mm = p.m ? p.m[q] : 0.0;
      if (p.W) {
        wq = p.W[q];
        for (int i = 0; i < 3; ++i) dwq[i] = p.dW[(int64_t)i * p.comp_stride + q];
      }
    }
    if (p.W) {
      const double wi = 1.0 / wq;
      double z[3];
      for (int i = 0; i < 3; ++i) z[i] = dwq[i] * wi * wi;
      double cz[3] = {0, 0, 0}, ctz[3] = {0, 0, 0}, zcz = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          cz[i] += c[i][j] * z[j];
          ctz[j] += c[i][j] * z[i];
          zcz += z[i] * c[i][j] * z[j];
        }
      D[0][0] = zcz + mm * wi * wi;
      for (int j = 0; j < 3; ++j) {
        D[0][j + 1] = -ctz[j] * wi;
        D[j + 1][0] = -cz[j] * wi;
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i + 1][j + 1] = c[i][j] * wi * wi;
    } else {
      D[0][0] = mm;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i + 1][j + 1] = c[i][j];
    }
    const int64_t st = npts;
    switch (p.form) {
      case FORM_M1:
        dst[0] = (T)D[0][0];
        break;
      case FORM_K6:
        dst[0] = (T)D[1][1]; dst[st] = (T)D[2][2]; dst[2 * st] = (T)D[3][3];
        dst[3 * st] = (T)D[1][2]; dst[4 * st] = (T)D[1][3]; dst[5 * st] = (T)D[2][3];
        break;
      case FORM_KM7:
      case FORM_KM7X:
        dst[0] = (T)D[0][0]; dst[st] = (T)D[1][1]; dst[2 * st] = (T)D[2][2]; dst[3 * st] = (T)D[3][3];
        dst[4 * st] = (T)D[1][2]; dst[5 * st] = (T)D[1][3]; dst[6 * st] = (T)D[2][3];
        break;
      case FORM_SYM10:
        dst[0] = (T)D[0][0]; dst[st] = (T)D[1][1]; dst[2 * st] = (T)D[2][2]; dst[3 * st] = (T)D[3][3];
        dst[4 * st] = (T)D[0][1]; dst[5 * st] = (T)D[0][2]; dst[6 * st] = (T)D[0][3];
        dst[7 * st] = (T)D[1][2]; dst[8 * st] = (T)D[1][3]; dst[9 * st] = (T)D[2][3];
        break;
      default:
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) dst[(a * 4 + b) * st] = (T)D[a][b];
    }
  }
}

// H[plane][c] = sum_{a,b} coef[a][b][c] e^{i pi (a+1) x} e^{i pi (b+1) y} for the
// scaled cross-section coordinates (x, y) = xy[plane], xy[n_plane + plane]: the
// cross-section factor of the separable Fourier field of the lognormal
// conductivity (modes <= kMaxModes wave numbers per axis, coef complex modes^3 with
// c fastest, staged in LDS).  One thread per cross-section point.
constexpr int kMaxModes = 16;

template <int M>
__global__ void lognormal_plane_modes_kernel(int64_t nplane, const double* __restrict__ xy,
                                             const double* __restrict__ coef, double* __restrict__ H) {
  __shared__ double cs[2 * M * M * M];
  constexpr int ncoef = 2 * M * M * M;
  for (int i = threadIdx.x; i < ncoef; i += blockDim.x) cs[i] = coef[i];
  __syncthreads();
  for (int64_t pl = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pl < nplane;
       pl += (int64_t)gridDim.x * blockDim.x) {
    double ex[M][2], ey[M][2];
    const double x = xy[pl], y = xy[nplane + pl];
    for (int k = 0; k < M; ++k) {
      sincospi((k + 1) * x, &ex[k][1], &ex[k][0]);
      sincospi((k + 1) * y, &ey[k][1], &ey[k][0]);
    }
    double h[M][2];
    for (int c = 0; c < M; ++c) h[c][0] = h[c][1] = 0.0;
    for (int a = 0; a < M; ++a)
      for (int b = 0; b < M; ++b) {
        const double pr = ex[a][0] * ey[b][0] - ex[a][1] * ey[b][1];
        const double pi = ex[a][0] * ey[b][1] + ex[a][1] * ey[b][0];
        const double* cab = cs + 2 * (a * M + b) * M;
        for (int c = 0; c < M; ++c) {
          h[c][0] += cab[2 * c] * pr - cab[2 * c + 1] * pi;
          h[c][1] += cab[2 * c] * pi + cab[2 * c + 1] * pr;
        }
      }
    double* out = H + 2 * pl * M;
    for (int c = 0; c < M; ++c) {
      out[2 * c] = h[c][0];
      out[2 * c + 1] = h[c][1];
    }
  }
}

// Dispatch of the wave-number count to its kernel instance.
template <int M>
void launch_plane_modes(int modes, int64_t nplane, const double* xy, const double* coef, double* H, hipStream_t s) {
  if constexpr (M > kMaxModes) {
    throw Error(IGA_ERR_INVALID, "modes must be 1..16");
  } else {
    if (modes != M) return launch_plane_modes<M + 1>(modes, nplane, xy, coef, H, s);
    lognormal_plane_modes_kernel<M><<<grid_for(nplane, 128), 128, 0, s>>>(nplane, xy, coef, H);
    IGA_LAUNCH_CHECK();
  }
}

}  // namespace iga

using namespace iga;

struct iga_fused_s {
  int dtype = IGA_F64, form = FORM_KM7, P = 0;
  int nc[3] = {0, 0, 0}, nq[3] = {0, 0, 0};
  int nt1 = 0, nt2 = 0, nslab = 0;
  int64_t coef_elems = 0, part_stride = 0, part_blk = 0;
  int rhs_block = 1;  // right-hand sides per work-group when several share the coefficients
  // device tables
  int32_t* d_first[3] = {nullptr, nullptr, nullptr};
  void* d_tab[3] = {nullptr, nullptr, nullptr};  // T[4][nq][P]
  int32_t *d_t1_q0 = nullptr, *d_t1_c0 = nullptr, *d_t1_nc = nullptr;
  int32_t *d_t2_q0 = nullptr, *d_t2_c0 = nullptr, *d_t2_nc = nullptr;
  int32_t *d_s_q0 = nullptr, *d_s_c0 = nullptr, *d_s_nc = nullptr;
  int64_t *d_coef_base = nullptr, *d_part_base = nullptr;
  int32_t* d_list[3][3] = {{nullptr}};  // [dir][cnt, t, off]
  std::vector<void*> owned;
};

namespace iga {

// All device tables of a plan go up in one allocation and one copy: the host
// vectors are staged with 256-byte alignment and each destination pointer is set
// once the block exists (plan creation then costs one synchronous transfer).
struct Staging {
  std::vector<char> host;
  std::vector<std::function<void(char*)>> fix;
  template <typename Ptr, typename V>
  void put(Ptr* dst, const std::vector<V>& h) {
    const size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + std::max<size_t>(h.size(), 1) * sizeof(V));
    if (!h.empty()) std::memcpy(host.data() + off, h.data(), h.size() * sizeof(V));
    fix.emplace_back([dst, off](char* base) { *dst = reinterpret_cast<Ptr>(base + off); });
  }
  void commit(iga_fused_s* f) {
    char* d = nullptr;
    IGA_HIP(hipMalloc(&d, std::max<size_t>(host.size(), 256)));
    f->owned.push_back(d);
    IGA_HIP(hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
    for (auto& set : fix) set(d);
  }
};

static void free_fused(iga_fused_s* f) {
  for (void* p : f->owned) (void)hipFree(p);
  f->owned.clear();
}

// Tiling of one direction into runs of `tq` points, with the DOF footprint of each run.
struct DirTiling {
  std::vector<int32_t> q0, c0, nc;                  // q0 has ntile+1 entries
  std::vector<int32_t> cnt, t, off;                 // per DOF: covering tiles (kKMax slots)
};

static DirTiling tile_direction(const int32_t* first, int nq, int ncdof, int P, int tq, int dmax) {
  DirTiling d;
  for (int q = 0; q < nq; q += tq) d.q0.push_back(q);
  d.q0.push_back(nq);
  const int nt = (int)d.q0.size() - 1;
  d.cnt.assign(ncdof, 0);
  d.t.assign((size_t)ncdof * kKMax, 0);
  d.off.assign((size_t)ncdof * kKMax, 0);
  for (int t = 0; t < nt; ++t) {
    int lo = first[d.q0[t]], hi = first[d.q0[t]];
    for (int q = d.q0[t]; q < d.q0[t + 1]; ++q) {
      lo = std::min(lo, (int)first[q]);
      hi = std::max(hi, (int)first[q]);
    }
    const int n = hi + P - lo;
    IGA_CHECK_ARG(dmax <= 0 || n <= dmax, "tile footprint exceeds the kernel's LDS block");
    d.c0.push_back(lo);
    d.nc.push_back(n);
    for (int c = lo; c < lo + n; ++c) {
      IGA_CHECK_ARG(d.cnt[c] < kKMax, "a DOF is covered by too many tiles");
      d.t[(size_t)c * kKMax + d.cnt[c]] = t;
      d.off[(size_t)c * kKMax + d.cnt[c]] = c - lo;
      d.cnt[c]++;
    }
  }
  for (int c = 0; c < ncdof; ++c) IGA_CHECK_ARG(d.cnt[c] > 0, "a DOF is not covered by any tile");
  return d;
}

// The occupancy-target variant (three waves per SIMD for FP64, four for FP32) is used
// unless the register budget it imposes makes the compiler spill more than
// kSpillTolerance bytes per lane (private scratch), which costs more than the
// occupancy gains; decided once per instance from the code object's attributes.
constexpr size_t kSpillTolerance = 16;
template <typename T, int P, int FORM, int NR>
static bool use_wide() {
  static std::atomic<int> wide{-1};
  int w = wide.load(std::memory_order_relaxed);
  if (w < 0) {
    hipFuncAttributes fa{};
    w = (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&fused_apply_kernel<T, P, FORM, NR>)) ==
             hipSuccess &&
         fa.localSizeBytes > kSpillTolerance)
            ? 1
            : 0;
    wide.store(w, std::memory_order_relaxed);
  }
  return w != 0;
}

template <typename T, int P, int FORM, int NR>
static void launch_apply_pfn(const FusedArgs<T>& a, int nblk, hipStream_t s) {
  if (use_wide<T, P, FORM, NR>())
    fused_apply_kernel_wide<T, P, FORM, NR><<<nblk, kThreads, 0, s>>>(a);
  else
    fused_apply_kernel<T, P, FORM, NR><<<nblk, kThreads, 0, s>>>(a);
}

// nr right-hand sides per work-group (2 only where the form has that variant)
template <typename T, int P, int FORM>
static void launch_apply_pf(const FusedArgs<T>& a, int nblk, int nr, hipStream_t s) {
  if constexpr (form_has_nr2(FORM)) {
    if (nr == 2) {
      launch_apply_pfn<T, P, FORM, 2>(a, nblk, s);
      return;
    }
  }
  launch_apply_pfn<T, P, FORM, 1>(a, nblk, s);
}

// resident work-groups per CU of the variant launch_apply_pf picks for one right-hand
// side per work-group (0 if unknown)
template <typename T, int P, int FORM>
static int resident_blocks_pf() {
  int nb = 0;
  const void* fn = use_wide<T, P, FORM, 1>()
                       ? reinterpret_cast<const void*>(&fused_apply_kernel_wide<T, P, FORM, 1>)
                       : reinterpret_cast<const void*>(&fused_apply_kernel<T, P, FORM, 1>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kThreads, 0) != hipSuccess) return 0;
  return nb;
}

template <typename T, int P>
static int resident_blocks_p(int form) {
  switch (form) {
    case FORM_M1: return resident_blocks_pf<T, P, FORM_M1>();
    case FORM_K6: return resident_blocks_pf<T, P, FORM_K6>();
    case FORM_KM7: return resident_blocks_pf<T, P, FORM_KM7>();
    case FORM_SYM10: return resident_blocks_pf<T, P, FORM_SYM10>();
    case FORM_KM7X: return resident_blocks_pf<T, P, FORM_KM7X>();
    case FORM_GV4: return resident_blocks_pf<T, P, FORM_GV4>();
    case FORM_UG4: return resident_blocks_pf<T, P, FORM_UG4>();
    default: return resident_blocks_pf<T, P, FORM_GEN16>();
  }
}

template <typename T>
static int resident_blocks(int P, int form) {
  switch (P) {
    case 3: return resident_blocks_p<T, 3>(form);
    case 4: return resident_blocks_p<T, 4>(form);
    case 5: return resident_blocks_p<T, 5>(form);
    case 6: return resident_blocks_p<T, 6>(form);
    case 7: return resident_blocks_p<T, 7>(form);
    default: return resident_blocks_p<T, 8>(form);
  }
}

template <typename T, int P>
static void launch_apply_p(const iga_fused_s* f, const FusedArgs<T>& a, int nblk, int nr, hipStream_t s) {
  switch (f->form) {
    case FORM_M1: launch_apply_pf<T, P, FORM_M1>(a, nblk, nr, s); break;
    case FORM_K6: launch_apply_pf<T, P, FORM_K6>(a, nblk, nr, s); break;
    case FORM_KM7: launch_apply_pf<T, P, FORM_KM7>(a, nblk, nr, s); break;
    case FORM_SYM10: launch_apply_pf<T, P, FORM_SYM10>(a, nblk, nr, s); break;
    case FORM_KM7X: launch_apply_pf<T, P, FORM_KM7X>(a, nblk, nr, s); break;
    case FORM_GV4: launch_apply_pf<T, P, FORM_GV4>(a, nblk, nr, s); break;
    case FORM_UG4: launch_apply_pf<T, P, FORM_UG4>(a, nblk, nr, s); break;
    default: launch_apply_pf<T, P, FORM_GEN16>(a, nblk, nr, s); break;
  }
}

template <typename T>
static void launch_apply(const iga_fused_s* f, const FusedArgs<T>& a, int nblk, int nr, hipStream_t s) {
  switch (f->P) {
    case 3: launch_apply_p<T, 3>(f, a, nblk, nr, s); break;
    case 4: launch_apply_p<T, 4>(f, a, nblk, nr, s); break;
    case 5: launch_apply_p<T, 5>(f, a, nblk, nr, s); break;
    case 6: launch_apply_p<T, 6>(f, a, nblk, nr, s); break;
    case 7: launch_apply_p<T, 7>(f, a, nblk, nr, s); break;
    case 8: launch_apply_p<T, 8>(f, a, nblk, nr, s); break;
    default: throw Error(IGA_ERR_UNSUPPORTED, "band width p+1 must be in 3..8");
  }
}

// Where right-hand side k lives: x[dof * xd + k * xr], y[dof * yd + k * yr], its
// coefficients at coef + k * cr.
struct Layout {
  int64_t xd, xr, yd, yr, cr;
  static Layout interleaved(int nrhs) { return Layout{nrhs, 1, nrhs, 1, 0}; }
};

// stages: bit 1 = main sweep (partial blocks), bit 2 = gather into y
template <typename T>
static void run_apply(const iga_fused_s* f, const void* coef, const void* x, void* y, int nrhs, const void* w,
                      double alpha, int accumulate, void* work, hipStream_t s, int stages = 3,
                      Layout lay = Layout{0, 0, 0, 0, 0}, const void* x1 = nullptr,
                      const int64_t* box_lo = nullptr, const int64_t* box_len = nullptr,
                      const void* const* gcoef = nullptr, void* const* gy = nullptr) {
  if (lay.xd == 0) lay = Layout::interleaved(nrhs);
  IGA_CHECK_ARG((f->form == FORM_KM7X) == (x1 != nullptr),
                "the two-input form takes a second input (iga_fused_apply_sliced2), the others do not");
  if (stages & 1) {
  FusedArgs<T> a{};
  for (int d = 0; d < 3; ++d) {
    a.nc[d] = f->nc[d];
    a.nq[d] = f->nq[d];
    a.d[d].first = f->d_first[d];
    a.d[d].tab = (const T*)f->d_tab[d];
  }
  a.nt1 = f->nt1;
  a.nt2 = f->nt2;
  a.nslab = f->nslab;
  a.nrhs = nrhs;
  a.t1_q0 = f->d_t1_q0; a.t1_c0 = f->d_t1_c0; a.t1_nc = f->d_t1_nc;
  a.t2_q0 = f->d_t2_q0; a.t2_c0 = f->d_t2_c0; a.t2_nc = f->d_t2_nc;
  a.s_q0 = f->d_s_q0; a.s_c0 = f->d_s_c0; a.s_nc = f->d_s_nc;
  a.coef_base = f->d_coef_base;
  a.part_base = f->d_part_base;
  a.part_stride = f->part_stride;
  a.x_dstride = lay.xd;
  a.x_rstride = lay.xr;
  a.coef_rstride = lay.cr;
  a.coef = (const T*)coef;
  if (gcoef) {
    IGA_CHECK_ARG(nrhs <= kGroupMax, "too many members for one grouped apply");
    a.ngroup = nrhs;
    for (int k = 0; k < nrhs; ++k) a.gcoef[k] = (const T*)gcoef[k];
  }
  a.x = (const T*)x;
  a.x1 = (const T*)x1;
  a.w = (const T*)w;
  a.partial = (T*)work;
  // two right-hand sides per work-group when they share one coefficient set (every
  // coefficient load, band table and plane-phase barrier then serves both)
  const int nr = (f->rhs_block == 2 && form_has_nr2(f->form) && nrhs % 2 == 0 && !gcoef && lay.cr == 0) ? 2 : 1;
  const int64_t nlog = (int64_t)f->nt1 * f->nt2 * f->nslab * (nrhs / nr);
  IGA_CHECK_ARG(nlog < (1LL << 30), "too many work-groups");
  a.nlogical = (int)nlog;
  const int nblk = (int)((nlog + 7) / 8 * 8);
  launch_apply<T>(f, a, nblk, nr, s);
  IGA_LAUNCH_CHECK();
  }
  if (!(stages & 2)) return;

  GatherArgs<T> g{};
  for (int d = 0; d < 3; ++d) g.nc[d] = f->nc[d];
  g.nt1 = f->nt1;
  g.nt2 = f->nt2;
  g.nslab = f->nslab;
  g.nrhs = nrhs;
  g.l1_cnt = f->d_list[0][0]; g.l1_t = f->d_list[0][1]; g.l1_off = f->d_list[0][2];
  g.l2_cnt = f->d_list[1][0]; g.l2_t = f->d_list[1][1]; g.l2_off = f->d_list[1][2];
  g.l3_cnt = f->d_list[2][0]; g.l3_t = f->d_list[2][1]; g.l3_off = f->d_list[2][2];
  g.part_blk = f->part_blk;
  g.part_stride = f->part_stride;
  g.partial = (const T*)work;
  g.w = (const T*)w;
  g.y = (T*)y;
  g.y_dstride = lay.yd;
  g.y_rstride = lay.yr;
  g.alpha = (T)alpha;
  g.accumulate = accumulate;
  if (box_lo) {
    g.box = 1;
    for (int d = 0; d < 3; ++d) {
      g.blo[d] = (int)box_lo[d];
      g.bhi[d] = (int)(box_lo[d] + box_len[d]);
    }
  }
  const int64_t ndof = (int64_t)f->nc[0] * f->nc[1] * f->nc[2];
  if (gy) {
    IGA_CHECK_ARG(nrhs <= kGroupMax, "too many members for one grouped apply");
    g.ngroup = nrhs;
    for (int k = 0; k < nrhs; ++k) g.gy[k] = (T*)gy[k];
    fused_gather_kernel<T><<<dim3(grid_for(ndof, 256), (unsigned)nrhs), 256, 0, s>>>(g);
  } else {
    fused_gather_kernel<T><<<grid_for(ndof * nrhs, 256), 256, 0, s>>>(g);
  }
  IGA_LAUNCH_CHECK();
}

template <typename T>
static void run_pack(const iga_fused_s* f, const double* C, const double* m, const double* W, const double* dW,
                     void* out, hipStream_t s, int64_t comp_stride = 0, const ExtrudedSource* ex = nullptr) {
  PackArgs<T> p{};
  for (int d = 0; d < 3; ++d) p.nq[d] = f->nq[d];
  p.nt1 = f->nt1;
  p.form = f->form;
  p.ncoef = form_ncoef(f->form);
  p.tq1 = kTQ1;
  p.tq2 = kTQ2;
  p.t1_q0 = f->d_t1_q0;
  p.t2_q0 = f->d_t2_q0;
  p.coef_base = f->d_coef_base;
  p.C = C;
  p.m = m;
  p.W = W;
  p.dW = dW;
  p.out = (T*)out;
  const int64_t nqt = (int64_t)f->nq[0] * f->nq[1] * f->nq[2];
  p.comp_stride = comp_stride > 0 ? comp_stride : nqt;
  if (ex) p.ex = *ex;
  fused_pack_kernel<T><<<grid_for(nqt, 256), 256, 0, s>>>(p);
  IGA_LAUNCH_CHECK();
}

// Scratch of iga_fused_apply_noalloc: the sweep's partial blocks live in this
// file-scope array instead of a caller workspace, so that the entry point allocates
// nothing.  Every GPU the library runs on holds its own copy (kNoAllocBytes, until
// the process ends).  Host code never passes the array's name to a kernel: it uses
// the device address hipGetSymbolAddress returns, looked up once per device with
// that device current and kept in g_scratch_ptr.
constexpr int64_t kNoAllocBytes = int64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

static void* scratch_of_current_device() {
  int dev = -1;
  IGA_HIP(hipGetDevice(&dev));
  IGA_CHECK_ARG(dev >= 0 && dev < kScratchDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    IGA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    IGA_CHECK_ARG(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace iga

extern "C" {

int iga_fused_create(iga_fused_t* out, const int32_t* nc, const int32_t* nq, int P, const int32_t* const* first,
                     const double* const* tabs, int form, int dtype, int nslab) {
  return guard([&] {
    IGA_CHECK_ARG(out && nc && nq && first && tabs, "NULL argument");
    IGA_CHECK_ARG(P >= 3 && P <= 8, "band width p+1 must be in 3..8");
    IGA_CHECK_ARG(form >= FORM_M1 && form <= FORM_UG4, "unknown coefficient form");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    for (int d = 0; d < 3; ++d) {
      IGA_CHECK_ARG(nc[d] >= P && nq[d] >= 1, "bad sizes");
      for (int q = 0; q < nq[d]; ++q) {
        IGA_CHECK_ARG(first[d][q] >= 0 && first[d][q] + P <= nc[d], "band leaves the DOF range");
        if (q > 0) IGA_CHECK_ARG(first[d][q] >= first[d][q - 1], "band starts must be non-decreasing");
      }
    }
    auto* f = new iga_fused_s();
    try {
      f->dtype = dtype;
      f->form = form;
      f->P = P;
      for (int d = 0; d < 3; ++d) {
        f->nc[d] = nc[d];
        f->nq[d] = nq[d];
      }
      DirTiling d1 = tile_direction(first[0], nq[0], nc[0], P, kTQ1, kDMax);
      DirTiling d2 = tile_direction(first[1], nq[1], nc[1], P, kTQ2, kDMax);
      f->nt1 = (int)d1.c0.size();
      f->nt2 = (int)d2.c0.size();
      // Slabs of q3 (one work-group per tile and slab with one right-hand side).  A
      // work-group costs about (its q3 planes + 6P) plane steps, the 6P being the
      // window's fill and drain; work-groups run in rounds of the device's resident
      // slots.  The count minimising rounds x cost is taken (the slab table lives in
      // LDS, which bounds a slab to kSlabMax planes).
      const int ns_min = (nq[2] + kSlabMax - 1) / kSlabMax;
      int ns = nslab;
      if (ns <= 0) {
        const int64_t tiles = (int64_t)f->nt1 * f->nt2;
        int dev = 0, ncu = 0;
        // FP64 work-groups count two per compute unit whatever the occupancy allows: a
        // third resident one adds no throughput (over p 2..5, nel 32..96, both
        // geometries, sweep + gather within 2.2% of the best slab count on average
        // with two, 4.4% and up to 31% slower with the occupancy's three)
        int per_cu = dtype == IGA_F64 ? resident_blocks<double>(P, form) : resident_blocks<float>(P, form);
        if (dtype == IGA_F64 && per_cu > 2) per_cu = 2;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
          ncu = 0;
        const int64_t slots = per_cu > 0 && ncu > 0 ? (int64_t)per_cu * ncu : 1024;
        double best = 0.0;
        // slabs of at least ~3P planes keep each DOF plane within a few slabs
        const int ns_max = std::max(ns_min, std::min(nq[2] / (3 * P), 4 * ns_min + 64));
        for (int c = ns_min; c <= ns_max; ++c) {
          const double rounds = (double)((tiles * c + slots - 1) / slots);
          const double cost = rounds * ((double)nq[2] / c + 6.0 * P);
          if (ns <= 0 || cost < best) {
            best = cost;
            ns = c;
          }
        }
      }
      ns = std::max(1, std::min(ns, nq[2]));
      ns = std::max(ns, ns_min);
      const int slab = (nq[2] + ns - 1) / ns;
      DirTiling d3 = tile_direction(first[2], nq[2], nc[2], P, slab, 0);
      f->nslab = (int)d3.c0.size();
      // packed coefficient offsets per tile
      const int ncoef = form_ncoef(form);
      // partial blocks all get the size of the largest one, so that the gather
      // computes a block's start instead of loading it (one dependent load less)
      const int64_t blk = (int64_t)*std::max_element(d1.nc.begin(), d1.nc.end()) *
                          *std::max_element(d2.nc.begin(), d2.nc.end()) *
                          *std::max_element(d3.nc.begin(), d3.nc.end());
      std::vector<int64_t> coef_base(f->nt1 * f->nt2), part_base((size_t)f->nt1 * f->nt2 * f->nslab);
      int64_t cb = 0;
      for (int t2 = 0; t2 < f->nt2; ++t2)
        for (int t1 = 0; t1 < f->nt1; ++t1) {
          const int64_t npts = (int64_t)(d1.q0[t1 + 1] - d1.q0[t1]) * (d2.q0[t2 + 1] - d2.q0[t2]);
          coef_base[t2 * f->nt1 + t1] = cb;
          cb += npts * ncoef * nq[2];
          for (int sl = 0; sl < f->nslab; ++sl) {
            const size_t b = ((size_t)t2 * f->nt1 + t1) * f->nslab + sl;
            part_base[b] = (int64_t)b * blk;
          }
        }
      f->coef_elems = cb;
      f->part_blk = blk;
      f->part_stride = (int64_t)part_base.size() * blk;
      // band tables in the apply precision
      Staging st;
      for (int d = 0; d < 3; ++d) {
        st.put(&f->d_first[d], std::vector<int32_t>(first[d], first[d] + nq[d]));
        const size_t nt = (size_t)4 * nq[d] * P;
        if (dtype == IGA_F64) {
          st.put(&f->d_tab[d], std::vector<double>(tabs[d], tabs[d] + nt));
        } else {
          std::vector<float> t32(nt);
          for (size_t i = 0; i < nt; ++i) t32[i] = (float)tabs[d][i];
          st.put(&f->d_tab[d], t32);
        }
      }
      st.put(&f->d_t1_q0, d1.q0); st.put(&f->d_t1_c0, d1.c0); st.put(&f->d_t1_nc, d1.nc);
      st.put(&f->d_t2_q0, d2.q0); st.put(&f->d_t2_c0, d2.c0); st.put(&f->d_t2_nc, d2.nc);
      st.put(&f->d_s_q0, d3.q0); st.put(&f->d_s_c0, d3.c0); st.put(&f->d_s_nc, d3.nc);
      st.put(&f->d_coef_base, coef_base);
      st.put(&f->d_part_base, part_base);
      const DirTiling* dt[3] = {&d1, &d2, &d3};
      for (int d = 0; d < 3; ++d) {
        st.put(&f->d_list[d][0], dt[d]->cnt);
        std::vector<int32_t> tt = dt[d]->t;
        if (d < 2) {
          // tile index and that tile's DOF count in one word
          for (int32_t& v : tt) {
            if (v < 0 || v > 0xffff || dt[d]->nc[v] > 0x7fff)
              throw Error(IGA_ERR_UNSUPPORTED, "too many tiles for the gather tables");
            v |= dt[d]->nc[v] << 16;
          }
        }
        st.put(&f->d_list[d][1], tt);
        st.put(&f->d_list[d][2], dt[d]->off);
      }
      st.commit(f);
    } catch (...) {
      free_fused(f);
      delete f;
      throw;
    }
    *out = f;
  });
}

int iga_fused_set_rhs_block(iga_fused_t f, int nr) {
  return guard([&] {
    IGA_CHECK_ARG(f && (nr == 1 || nr == 2), "right-hand sides per work-group: 1 or 2");
    f->rhs_block = nr;
  });
}

int iga_fused_destroy(iga_fused_t f) {
  return guard([&] {
    if (!f) return;
    free_fused(f);
    delete f;
  });
}

int iga_fused_info(iga_fused_t f, int64_t* coef_elems, int64_t* part_elems_per_rhs, int* ntiles1, int* ntiles2,
                   int* nslabs) {
  return guard([&] {
    IGA_CHECK_ARG(f, "NULL plan");
    if (coef_elems) *coef_elems = f->coef_elems;
    if (part_elems_per_rhs) *part_elems_per_rhs = f->part_stride;
    if (ntiles1) *ntiles1 = f->nt1;
    if (ntiles2) *ntiles2 = f->nt2;
    if (nslabs) *nslabs = f->nslab;
  });
}

int64_t iga_fused_workspace_bytes(iga_fused_t f, int nrhs) {
  if (!f || nrhs < 1) return 0;
  return f->part_stride * (int64_t)nrhs * (f->dtype == IGA_F64 ? 8 : 4);
}

int iga_fused_pack(iga_fused_t f, const double* C, const double* m, const double* W, const double* dW, void* packed,
                   void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed, "NULL argument");
    IGA_CHECK_ARG((W == nullptr) == (dW == nullptr), "W and dW go together");
    if (f->form != FORM_M1) IGA_CHECK_ARG(C, "stiffness forms need C");
    if (f->form == FORM_M1 || f->form == FORM_KM7 || f->form == FORM_KM7X) IGA_CHECK_ARG(m, "mass forms need m");
    if (f->dtype == IGA_F64)
      run_pack<double>(f, C, m, W, dW, packed, as_stream(stream));
    else
      run_pack<float>(f, C, m, W, dW, packed, as_stream(stream));
  });
}

int iga_fused_apply(iga_fused_t f, const void* packed, const void* x, void* y, int nrhs, const void* nurbs_w,
                    double alpha, int accumulate, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && x && y && work, "NULL argument");
    IGA_CHECK_ARG(nrhs >= 1, "nrhs must be >= 1");
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream));
    else
      run_apply<float>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream));
  });
}

int iga_fused_apply_group(iga_fused_t f, int nmember, const void* const* packed, const void* x, int64_t x_dstride,
                          int64_t x_mstride, void* const* y, const void* nurbs_w, void* work, int ndim,
                          const int64_t* box_lo, const int64_t* box_len, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && x && y && work, "NULL argument");
    IGA_CHECK_ARG(x_dstride >= 1 && x_mstride >= 0, "bad input strides");
    IGA_CHECK_ARG(nmember >= 1 && nmember <= kGroupMax, "a grouped apply takes 1 to 16 members");
    for (int k = 0; k < nmember; ++k) IGA_CHECK_ARG(packed[k] && y[k], "NULL coefficient set or output");
    IGA_CHECK_ARG(f->form != FORM_KM7X, "the two-input form is not grouped");
    IGA_CHECK_ARG(ndim == 0 || ndim == 3, "the mask box of the fused operator has three directions");
    IGA_CHECK_ARG(ndim == 0 || (box_lo && box_len), "box arrays missing");
    for (int d = 0; d < ndim; ++d)
      IGA_CHECK_ARG(box_lo[d] >= 0 && box_len[d] >= 0 && box_lo[d] + box_len[d] <= f->nc[d], "box outside the grid");
    const int64_t* lo = ndim == 3 ? box_lo : nullptr;
    const int64_t* len = ndim == 3 ? box_len : nullptr;
    // the output strides are unused: the gather writes member k to y[k][dof]
    const Layout lay{x_dstride, x_mstride, 1, 0, 0};
    // one coefficient set for every member (right-hand sides of one operator): the
    // sweep's shared-coefficient path
    bool shared = true;
    for (int k = 1; k < nmember; ++k) shared = shared && packed[k] == packed[0];
    const void* const* gco = shared ? nullptr : packed;
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed[0], x, y[0], nmember, nurbs_w, 1.0, 0, work, as_stream(stream), 3, lay, nullptr, lo,
                        len, gco, y);
    else
      run_apply<float>(f, packed[0], x, y[0], nmember, nurbs_w, 1.0, 0, work, as_stream(stream), 3, lay, nullptr, lo,
                       len, gco, y);
  });
}

int iga_fused_apply_noalloc(iga_fused_t f, const void* packed, const void* x, void* y, int nrhs,
                            const void* nurbs_w, double alpha, int accumulate, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && x && y, "NULL argument");
    IGA_CHECK_ARG(nrhs >= 1, "nrhs must be >= 1");
    const int64_t need = iga_fused_workspace_bytes(f, nrhs);
    if (need > kNoAllocBytes)
      throw Error(IGA_ERR_INVALID, "the partial blocks of this call need " + std::to_string(need) +
                                       " bytes, above the " + std::to_string(kNoAllocBytes) +
                                       "-byte built-in scratch of iga_fused_apply_noalloc (use iga_fused_apply "
                                       "with a workspace)");
    void* work = scratch_of_current_device();
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream));
    else
      run_apply<float>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream));
  });
}

int iga_fused_apply_stage(iga_fused_t f, int stage, const void* packed, const void* x, void* y, int nrhs,
                          const void* nurbs_w, double alpha, int accumulate, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && work && (stage == 1 || stage == 2), "bad argument");
    IGA_CHECK_ARG(stage == 2 || (packed && x), "main stage needs coefficients and x");
    IGA_CHECK_ARG(stage == 1 || y, "gather stage needs y");
    IGA_CHECK_ARG(nrhs >= 1, "nrhs must be >= 1");
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream), stage);
    else
      run_apply<float>(f, packed, x, y, nrhs, nurbs_w, alpha, accumulate, work, as_stream(stream), stage);
  });
}

int iga_fused_pack_extruded(iga_fused_t f, const double* plane, const double* dz, const double* H, const double* Ez,
                            int modes, double sigma, int with_mass, const double* W, const double* dW, void* packed,
                            void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && plane && dz, "NULL argument");
    IGA_CHECK_ARG((W == nullptr) == (dW == nullptr), "W and dW go together");
    IGA_CHECK_ARG(!H || (Ez && modes >= 1 && modes <= kMaxModes), "field modes: 1..16 with H and Ez");
    IGA_CHECK_ARG(f->form != FORM_M1 && f->form != FORM_GV4 && f->form != FORM_UG4,
                  "the generated coefficients are stiffness coefficients");
    if (f->form == FORM_KM7 || f->form == FORM_KM7X) IGA_CHECK_ARG(with_mass, "mass forms need the mass term");
    ExtrudedSource ex;
    ex.plane = plane;
    ex.dz = dz;
    ex.H = H;
    ex.Ez = Ez;
    ex.modes = H ? modes : 0;
    ex.sigma = sigma;
    ex.with_mass = with_mass;
    if (f->dtype == IGA_F64)
      run_pack<double>(f, nullptr, nullptr, W, dW, packed, as_stream(stream), 0, &ex);
    else
      run_pack<float>(f, nullptr, nullptr, W, dW, packed, as_stream(stream), 0, &ex);
  });
}

int iga_lognormal_plane_modes(int64_t nplane, int modes, const double* xy, const double* coef, double* H,
                              void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(xy && coef && H && nplane >= 1, "bad argument");
    IGA_CHECK_ARG(modes >= 1 && modes <= kMaxModes, "modes must be 1..16");
    launch_plane_modes<1>(modes, nplane, xy, coef, H, as_stream(stream));
  });
}

int iga_fused_pack_strided(iga_fused_t f, const double* C, const double* m, const double* W, const double* dW,
                           int64_t comp_stride, void* packed, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed, "NULL argument");
    IGA_CHECK_ARG((W == nullptr) == (dW == nullptr), "W and dW go together");
    if (f->form != FORM_M1) IGA_CHECK_ARG(C, "stiffness forms need C");
    if (f->form == FORM_M1 || f->form == FORM_KM7 || f->form == FORM_KM7X) IGA_CHECK_ARG(m, "mass forms need m");
    const int64_t nqt = (int64_t)f->nq[0] * f->nq[1] * f->nq[2];
    IGA_CHECK_ARG(comp_stride >= nqt, "component stride below the number of points");
    if (f->dtype == IGA_F64)
      run_pack<double>(f, C, m, W, dW, packed, as_stream(stream), comp_stride);
    else
      run_pack<float>(f, C, m, W, dW, packed, as_stream(stream), comp_stride);
  });
}

int iga_fused_apply_sliced(iga_fused_t f, const void* packed, int64_t coef_slice_stride, const void* x,
                           int64_t x_dstride, int64_t x_sstride, void* y, int64_t y_dstride, int64_t y_sstride,
                           int nslices, const void* nurbs_w, double alpha, int accumulate, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && x && y && work, "NULL argument");
    IGA_CHECK_ARG(nslices >= 1, "nslices must be >= 1");
    IGA_CHECK_ARG(x_dstride >= 1 && y_dstride >= 1 && x_sstride >= 0 && y_sstride >= 0 && coef_slice_stride >= 0,
                  "bad strides");
    const Layout lay{x_dstride, x_sstride, y_dstride, y_sstride, coef_slice_stride};
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed, x, y, nslices, nurbs_w, alpha, accumulate, work, as_stream(stream), 3, lay);
    else
      run_apply<float>(f, packed, x, y, nslices, nurbs_w, alpha, accumulate, work, as_stream(stream), 3, lay);
  });
}

int iga_fused_apply_sliced2(iga_fused_t f, const void* packed, int64_t coef_slice_stride, const void* x,
                            const void* x1, int64_t x_dstride, int64_t x_sstride, void* y, int64_t y_dstride,
                            int64_t y_sstride, int nslices, const void* nurbs_w, double alpha, int accumulate,
                            void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && packed && x && x1 && y && work, "NULL argument");
    IGA_CHECK_ARG(nslices >= 1, "nslices must be >= 1");
    IGA_CHECK_ARG(x_dstride >= 1 && y_dstride >= 1 && x_sstride >= 0 && y_sstride >= 0 && coef_slice_stride >= 0,
                  "bad strides");
    const Layout lay{x_dstride, x_sstride, y_dstride, y_sstride, coef_slice_stride};
    if (f->dtype == IGA_F64)
      run_apply<double>(f, packed, x, y, nslices, nurbs_w, alpha, accumulate, work, as_stream(stream), 3, lay, x1);
    else
      run_apply<float>(f, packed, x, y, nslices, nurbs_w, alpha, accumulate, work, as_stream(stream), 3, lay, x1);
  });
}

}  // extern "C"

namespace iga {

void fused_apply_masked(iga_fused_t f, const void* packed, const void* x, void* y, const void* nurbs_w, void* work,
                        const int64_t* box_lo, const int64_t* box_len, void* stream) {
  IGA_CHECK_ARG(f && packed && x && y && work && box_lo && box_len, "NULL argument");
  for (int d = 0; d < 3; ++d)
    IGA_CHECK_ARG(box_lo[d] >= 0 && box_len[d] >= 0 && box_lo[d] + box_len[d] <= f->nc[d], "box outside the grid");
  if (f->dtype == IGA_F64)
    run_apply<double>(f, packed, x, y, 1, nurbs_w, 1.0, 0, work, as_stream(stream), 3, Layout{0, 0, 0, 0, 0}, nullptr,
                      box_lo, box_len);
  else
    run_apply<float>(f, packed, x, y, 1, nurbs_w, 1.0, 0, work, as_stream(stream), 3, Layout{0, 0, 0, 0, 0}, nullptr,
                     box_lo, box_len);
}

}  // namespace iga
