// Fast-diagonalization apply  y[free] = U (Lambda^{-1} o U^T x[free]),  y = 0 elsewhere,
// with every Kronecker mode product run as an MFMA GEMM tile kernel.
//
// A mode product of the free-box tensor viewed as (outer, n, inner) is
//   Y(o, i, t) = sum_k A[i][k] X(o, k, t)          (A = U_d^T or U_d, n x n)
// which is one of two strided GEMMs C(m, j) = sum_k L(m, k) R(k, j):
//   inner > 1 :  L = A, R = X(o, :, :), C = Y(o, :, :)   (batched over o)
//   inner = 1 :  L = X (rows o), R = A^T, C = Y (rows o)
// A work-group computes a 64 x 64 tile of C with four 64-wide waves, each owning
// 2 x 2 blocks of the 16 x 16 MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32);
// K advances in steps of 16 staged through LDS.  The Lambda^{-1} scaling is fused
// into the epilogue of the last forward product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "common.hpp"

namespace iga {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc;
// Result layouts differ between the two instructions (measured on gfx950 with an
// in-bounds probe): register r of lane l holds row (l/16) + 4r for the FP64 MFMA and
// row 4(l/16) + r for the FP32 one; the column is l % 16 in both.
template <>
struct Acc<double> {
  using V = v4d;
  __device__ static V mfma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Acc<float> {
  using V = v4f;
  __device__ static V mfma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct GemmDesc {
  int M, N, K;                  // C is M x N, inner dimension K
  int64_t lm, lk, rk, rj;       // element strides of L(m,k) and R(k,j)
  int64_t cm, cj;               // strides of C(m,j)
  int64_t l_batch, r_batch, c_batch;  // per-batch offsets (blockIdx.z)
  int64_t scale_m, scale_j, scale_batch;  // index of the epilogue scale for C(m,j)
};

constexpr int kTile = 64, kKStep = 16;

// C = L R (+ epilogue scale) on TILE x TILE output tiles (64 or 32): four waves in
// 2 x 2, each owning (TILE/32) x (TILE/32) blocks of the 16x16 MFMA.  Lane layout
// of the 16x16x4 MFMA (measured on gfx950): A operand: lane l -> A[l % 16][l / 16];
// B operand: lane l -> B[l / 16][l % 16]; D register r of lane l ->
// D[Acc<T>::row(l, r)][l % 16].  The small tile serves the short FD factors
// (n <= 96): more work-groups and less padding for n just above a multiple of 64.
template <typename T, int TILE>
__global__ void __launch_bounds__(256) mfma_gemm_kernel(GemmDesc g, const T* __restrict__ L, const T* __restrict__ R,
                                                        T* __restrict__ Cout, const T* __restrict__ scale) {
  using V = typename Acc<T>::V;
  constexpr int NB = TILE / 32;            // 16x16 blocks per wave and direction
  constexpr int NE = TILE * kKStep / 256;  // tile elements each thread stages per K step
  __shared__ T ls[TILE][kKStep + 1];
  __shared__ T rs[kKStep][TILE + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * TILE, j0 = blockIdx.y * TILE;
  const int64_t bz = blockIdx.z;
  const T* Lb = L + bz * g.l_batch;
  const T* Rb = R + bz * g.r_batch;
  T* Cb = Cout + bz * g.c_batch;
  const int wm = (wave >> 1) * (TILE / 2), wn = (wave & 1) * (TILE / 2);
  V acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = V{0, 0, 0, 0};

  // Each thread stages NE elements of the L tile (TILE x 16) and NE of the R tile
  // (16 x TILE) per K step.  Their loads are issued kPre steps ahead into a
  // register ring, from clamped (always valid) addresses on every path, and zeroed
  // outside the matrix when written to LDS; so the K loop waits on one load
  // latency, not on one per step.
  constexpr int kPre = 4;
  const int nsteps = (g.K + kKStep - 1) / kKStep;
  int l_mm[NE], l_kk[NE], r_kk[NE], r_jj[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int idx = tid + e * 256;
    l_mm[e] = idx / kKStep;
    l_kk[e] = idx % kKStep;
    if (g.rj == 1) {
      r_kk[e] = idx / TILE;
      r_jj[e] = idx % TILE;
    } else {
      r_jj[e] = idx / kKStep;
      r_kk[e] = idx % kKStep;
    }
  }
  T lreg[kPre][NE], rreg[kPre][NE];
  auto load_step = [&](int s, T (&lr)[NE], T (&rr)[NE]) __attribute__((always_inline)) {
    const int k0 = (s < nsteps ? s : nsteps - 1) * kKStep;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int m = min(m0 + l_mm[e], g.M - 1), k = min(k0 + l_kk[e], g.K - 1);
      lr[e] = Lb[(int64_t)m * g.lm + (int64_t)k * g.lk];
      const int kr = min(k0 + r_kk[e], g.K - 1), j = min(j0 + r_jj[e], g.N - 1);
      rr[e] = Rb[(int64_t)kr * g.rk + (int64_t)j * g.rj];
    }
  };
#pragma unroll
  for (int b = 0; b < kPre; ++b) load_step(b, lreg[b], rreg[b]);

  for (int s0 = 0; s0 < nsteps; s0 += kPre) {
#pragma unroll
    for (int b = 0; b < kPre; ++b) {
      const int st = s0 + b;
      if (st < nsteps) {
        const int k0 = st * kKStep;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const bool lok = (m0 + l_mm[e] < g.M) && (k0 + l_kk[e] < g.K);
          ls[l_mm[e]][l_kk[e]] = lok ? lreg[b][e] : T(0);
          const bool rok = (k0 + r_kk[e] < g.K) && (j0 + r_jj[e] < g.N);
          rs[r_kk[e]][r_jj[e]] = rok ? rreg[b][e] : T(0);
        }
        __syncthreads();
        load_step(st + kPre, lreg[b], rreg[b]);
#pragma unroll
        for (int k4 = 0; k4 < kKStep; k4 += 4) {
          T av[NB], bv[NB];
#pragma unroll
          for (int a = 0; a < NB; ++a) av[a] = ls[wm + a * 16 + (lane & 15)][k4 + (lane >> 4)];
#pragma unroll
          for (int c = 0; c < NB; ++c) bv[c] = rs[k4 + (lane >> 4)][wn + c * 16 + (lane & 15)];
#pragma unroll
          for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int c = 0; c < NB; ++c) acc[a][c] = Acc<T>::mfma(av[a], bv[c], acc[a][c]);
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + Acc<T>::row(lane, r);
        const int j = j0 + wn + b * 16 + (lane & 15);
        if (m < g.M && j < g.N) {
          T v = acc[a][b][r];
          if (scale) v *= scale[bz * g.scale_batch + m * g.scale_m + j * g.scale_j];
          Cb[m * g.cm + j * g.cj] = v;
        }
      }
}

}  // namespace iga

using namespace iga;

struct iga_fd_s {
  int dtype = IGA_F64, ndim = 0;
  int64_t n[4] = {0, 0, 0, 0};        // full sizes per direction
  int64_t lo[4] = {0, 0, 0, 0};       // free box start
  int64_t len[4] = {0, 0, 0, 0};      // free box length (= size of U_d)
  void* d_Ut[4] = {nullptr, nullptr, nullptr, nullptr};  // U_d^T, row-major len x len
  void* d_U[4] = {nullptr, nullptr, nullptr, nullptr};
  void* d_inv = nullptr;               // Lambda^{-1} over the free box
  void* d_base = nullptr;              // one block holding every U_d and U_d^T
  int64_t nsub = 0;
};

namespace iga {

// Y(o, i, t) = sum_k A[i][k] X(o, k, t) with A of shape n_out x n_in (row-major).
template <typename T>
static void mode_gemm_rect(const T* A, const T* X, T* Y, int64_t outer, int64_t n_out, int64_t n_in, int64_t inner,
                           const T* scale, hipStream_t s) {
  GemmDesc g{};
  dim3 grid;
  // 32 x 32 tiles for short factors (more work-groups, less padding), else 64 x 64
  const int tile = n_out <= 96 ? 32 : kTile;
  auto launch = [&](dim3 gr, const T* Lp, const T* Rp, T* Yp, const T* sc) {
    if (tile == 32)
      mfma_gemm_kernel<T, 32><<<gr, 256, 0, s>>>(g, Lp, Rp, Yp, sc);
    else
      mfma_gemm_kernel<T, kTile><<<gr, 256, 0, s>>>(g, Lp, Rp, Yp, sc);
    IGA_LAUNCH_CHECK();
  };
  if (inner > 1) {
    g.M = (int)n_out;
    g.N = (int)inner;
    g.K = (int)n_in;
    g.lm = n_in; g.lk = 1;         // L = A
    g.rk = inner; g.rj = 1;        // R = X(o)
    g.cm = inner; g.cj = 1;
    g.l_batch = 0; g.r_batch = n_in * inner; g.c_batch = n_out * inner;
    g.scale_m = inner; g.scale_j = 1; g.scale_batch = n_out * inner;
    IGA_CHECK_ARG(inner < (int64_t)1 << 30, "mode product too large for the grid");
    // batches beyond the grid's z limit go in further launches
    constexpr int64_t kZMax = 65535;
    for (int64_t o0 = 0; o0 < outer; o0 += kZMax) {
      const int64_t nb = std::min(kZMax, outer - o0);
      grid = dim3((unsigned)((n_out + tile - 1) / tile), (unsigned)((inner + tile - 1) / tile), (unsigned)nb);
      launch(grid, A, X + o0 * g.r_batch, Y + o0 * g.c_batch, scale ? scale + o0 * g.scale_batch : nullptr);
    }
    return;
  } else {
    g.M = (int)outer;
    g.N = (int)n_out;
    g.K = (int)n_in;
    g.lm = n_in; g.lk = 1;         // L = X rows
    g.rk = 1; g.rj = n_in;         // R(k, i) = A[i][k]
    g.cm = n_out; g.cj = 1;
    g.scale_m = n_out; g.scale_j = 1;
    IGA_CHECK_ARG(outer < (int64_t)1 << 30, "mode product too large for the grid");
    grid = dim3((unsigned)((outer + tile - 1) / tile), (unsigned)((n_out + tile - 1) / tile), 1);
  }
  launch(grid, X, A, Y, scale);
}

template <typename T>
static void mode_gemm(const T* A, const T* X, T* Y, int64_t outer, int64_t n, int64_t inner, const T* scale,
                      hipStream_t s) {
  mode_gemm_rect<T>(A, X, Y, outer, n, n, inner, scale, s);
}

// ---------------------------------------------------------------------------
// Three-launch FD apply of a three-direction free box (the spatial preconditioner):
//   plane kernel, forward:  for box plane i3 and a block S of 16 rows i1, gather the
//     plane of x, Y1(S, :) = U1^T X along direction 1, then W(S, :, i3) = Y1 U2 along
//     direction 2 (box-dense W, direction 1 fastest)
//   line kernel:            for 16 columns (i1, i2, rhs) of W: Z = U3^T W, Z *= Lambda^-1,
//     W = U3 Z, in place
//   plane kernel, backward: y(box) = U1 (W along 1), then U2 along 2, written into the
//     full vector; every entry outside the box is set to zero
// Each work-group stages its operands in LDS once (all loads issued together) and runs
// both of its products on the matrix cores: FP64 (v_mfma_f64_16x16x4f64), FP32
// (v_mfma_f32_16x16x4f32), or BF16 inputs with FP32 accumulation
// (v_mfma_f32_16x16x16bf16_1k; float storage, operands rounded to bf16 as they are
// read).  Three launches replace the gather, six GEMM launches and the scatter.
// ---------------------------------------------------------------------------

enum FdMode { FD_F64 = 0, FD_F32 = 1, FD_BF16 = 2 };

typedef short v4s __attribute__((ext_vector_type(4)));

__device__ inline short f32_to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (short)(u >> 16);
}

// Matrix-core operands.  For FP64 / FP32 a lane feeds one element of A (row lane & 15,
// k = k0 + lane / 16) and one of B (column lane & 15, same k) to the 16x16x4 MFMA; for
// BF16 it feeds four consecutive k (k0 + 4 (lane / 16) + 0..3) to the 16x16x16 MFMA,
// read as one 8-byte LDS word because BF16 operands are stored in LDS as bf16.  The
// same k set reaches A and B, so the sum over k is the full product.  KS is the K step
// whose operands are all read before its MFMAs.
template <int MODE>
struct FdMma;
template <>
struct FdMma<FD_F64> {
  using T = double;   // storage of vectors and accumulators
  using S = double;   // LDS operand storage
  using Op = double;  // one MFMA operand of a lane
  using V = v4d;
  static constexpr int KI = 4, KS = 8;
  __device__ static S to_s(T v) { return v; }
  // operand of a row starting at `rowp` (k contiguous)
  __device__ static Op op(const S* rowp, int k0, int lane) { return rowp[k0 + (lane >> 4)]; }
  // operand of column `c` of an array stored k-major with pitch p
  __device__ static Op op_t(const S* base, int p, int c, int k0, int lane) { return base[(k0 + (lane >> 4)) * p + c]; }
  template <int TENSOR>
  __device__ static V mma(Op a, Op b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct FdMma<FD_F32> {
  using T = float;
  using S = float;
  using Op = float;
  using V = v4f;
  static constexpr int KI = 4, KS = 8;
  __device__ static S to_s(T v) { return v; }
  __device__ static Op op(const S* rowp, int k0, int lane) { return rowp[k0 + (lane >> 4)]; }
  __device__ static Op op_t(const S* base, int p, int c, int k0, int lane) { return base[(k0 + (lane >> 4)) * p + c]; }
  template <int TENSOR>
  __device__ static V mma(Op a, Op b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
// BF16: LDS holds float; an eigenvector-matrix operand is rounded to bf16 as it is
// read, a vector (tensor) operand x is split into bf16 hi = bf16(x) and lo = bf16(x - hi)
// and enters two MFMAs.  The products then carry the bf16 rounding of the eigenvectors
// (a fixed linear map) and the tensor to about 16 significant bits, so the
// preconditioner stays linear to ~1e-5 (a single bf16 rounding of the tensor makes it
// nonlinear at the 1e-2 level, which stalls the Krylov solves it preconditions).
template <>
struct FdMma<FD_BF16> {
  using T = float;
  using S = float;
  using Op = v4f;
  using V = v4f;
  static constexpr int KI = 16, KS = 16;
  __device__ static S to_s(T v) { return v; }
  // rows start on 16-byte boundaries (pitch a multiple of 4)
  __device__ static Op op(const S* rowp, int k0, int lane) {
    return *reinterpret_cast<const v4f*>(rowp + k0 + 4 * (lane >> 4));
  }
  __device__ static Op op_t(const S* base, int p, int c, int k0, int lane) {
    const S* q = base + (k0 + 4 * (lane >> 4)) * p + c;
    return v4f{q[0], q[p], q[2 * p], q[3 * p]};
  }
  __device__ static v4s round(Op x) {
    return v4s{f32_to_bf16(x[0]), f32_to_bf16(x[1]), f32_to_bf16(x[2]), f32_to_bf16(x[3])};
  }
  __device__ static v4s low(Op x, v4s hi) {
    v4s lo;
#pragma unroll
    for (int e = 0; e < 4; ++e) lo[e] = f32_to_bf16(x[e] - __uint_as_float((uint32_t)(uint16_t)hi[e] << 16));
    return lo;
  }
  // TENSOR = 0: A is the vector operand; 1: B is
  template <int TENSOR>
  __device__ static V mma(Op a, Op b, V c) {
    if constexpr (TENSOR == 0) {
      const v4s ah = round(a), bb = round(b);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bb, c, 0, 0, 0);
      return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(low(a, ah), bb, c, 0, 0, 0);
    } else {
      const v4s aa = round(a), bh = round(b);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, bh, c, 0, 0, 0);
      return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, low(b, bh), c, 0, 0, 0);
    }
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kFdMaxTiles = 3;  // 16-wide tiles per wave: box lengths up to 4 * 3 * 16 = 192

// acc[u] (+)= A B_u over k < K (a multiple of KS) for the TPW tiles u of a wave; `fa(k0)`
// gives the A operand, `fb(u, k0)` the B operand of tile u; TENSOR names the operand
// holding vector data (0: A, 1: B) for the split BF16 products.  The LDS reads of the next K
// step are issued before the MFMAs of the current one (two operand buffers), and tiles
// alternate, so consecutive MFMAs are independent and LDS latency overlaps them.
template <int MODE, int TPW, int TENSOR, class FA, class FB>
__device__ inline void fd_mma_loop(typename FdMma<MODE>::V (&acc)[TPW], int K, const FA& fa, const FB& fb) {
  using M = FdMma<MODE>;
  constexpr int NS = M::KS / M::KI;
  typename M::Op av[2][NS], bv[2][NS][TPW];
  auto fetch = [&](int buf, int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      av[buf][s2] = fa(k0 + s2 * M::KI);
#pragma unroll
      for (int u = 0; u < TPW; ++u) bv[buf][s2][u] = fb(u, k0 + s2 * M::KI);
    }
  };
  auto run = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
      for (int u = 0; u < TPW; ++u) acc[u] = M::template mma<TENSOR>(av[buf][s2], bv[buf][s2][u], acc[u]);
  };
  fetch(0, 0);
  int k0 = 0;
  // two K steps per trip so that each buffer is addressed statically
  for (; k0 + 2 * M::KS <= K; k0 += 2 * M::KS) {
    fetch(1, k0 + M::KS);
    run(0);
    if (k0 + 2 * M::KS < K) fetch(0, k0 + 2 * M::KS);
    run(1);
  }
  if (k0 < K) run(0);
}

// Stage a rows x cols block into LDS, dst[r * p + c] = src(r)[c * cstride]: rows split
// over the four waves, columns over the 64 lanes (CC = ceil(cols / 64) <= 3 passes), RG
// rows per wave per round with every load of the round issued before the stores.
// Clamped indices keep every load in bounds; out-of-range elements are not stored.
// Rows of a staged operand per wave and round (CC lane-columns of 64 per row).
template <int CC>
struct FdStageShape {
  static constexpr int RG = CC == 1 ? 8 : CC == 2 ? 6 : 4;
};

// Loads of one round: rows r0 + wave + 4 i (clamped, so always in bounds), columns
// lane + 64 cc (clamped).
template <int MODE, int CC, typename RowPtr>
__device__ inline void fd_stage_load(typename FdMma<MODE>::T (&v)[FdStageShape<CC>::RG][CC], int r0, int rows,
                                     int cols, const RowPtr& rowptr, int64_t cstride, int wave, int lane) {
  constexpr int RG = FdStageShape<CC>::RG;
#pragma unroll
  for (int i = 0; i < RG; ++i) {
    const auto* src = rowptr(min(r0 + wave + 4 * i, rows - 1));
#pragma unroll
    for (int cc = 0; cc < CC; ++cc) v[i][cc] = src[(int64_t)min(lane + 64 * cc, cols - 1) * cstride];
  }
}

template <int MODE, int CC>
__device__ inline void fd_stage_store(typename FdMma<MODE>::S* dst, int p, const typename FdMma<MODE>::T (
                                          &v)[FdStageShape<CC>::RG][CC], int r0, int rows, int cols, int wave,
                                      int lane) {
  constexpr int RG = FdStageShape<CC>::RG;
#pragma unroll
  for (int i = 0; i < RG; ++i) {
    const int r = r0 + wave + 4 * i;
#pragma unroll
    for (int cc = 0; cc < CC; ++cc) {
      const int c = lane + 64 * cc;
      if (r < rows && c < cols) dst[r * p + c] = FdMma<MODE>::to_s(v[i][cc]);
    }
  }
}

// Stage rows [0, rows) x columns [0, cols) of a global operand into LDS (row pitch p,
// converted to the operand type): rows over the waves, columns over the lanes, two
// rounds in flight (the loads of the next round are issued before the current round
// is written), so a stage costs about one memory round trip instead of one per round.
// `pre`, when given, holds the loads of the first round, issued earlier by the caller.
template <int MODE, int CC, typename RowPtr>
__device__ inline void fd_stage(typename FdMma<MODE>::S* dst, int p, int rows, int cols, const RowPtr& rowptr,
                                int64_t cstride, int wave, int lane,
                                const typename FdMma<MODE>::T (*pre)[FdStageShape<CC>::RG][CC] = nullptr) {
  using T = typename FdMma<MODE>::T;
  constexpr int RG = FdStageShape<CC>::RG;
  constexpr int step = 4 * RG;
  T v0[RG][CC], v1[RG][CC];
  if (pre) {
#pragma unroll
    for (int i = 0; i < RG; ++i)
#pragma unroll
      for (int cc = 0; cc < CC; ++cc) v0[i][cc] = (*pre)[i][cc];
  } else {
    fd_stage_load<MODE, CC>(v0, 0, rows, cols, rowptr, cstride, wave, lane);
  }
  for (int r0 = 0; r0 < rows; r0 += 2 * step) {
    const bool next = r0 + step < rows;
    if (next) fd_stage_load<MODE, CC>(v1, r0 + step, rows, cols, rowptr, cstride, wave, lane);
    fd_stage_store<MODE, CC>(dst, p, v0, r0, rows, cols, wave, lane);
    if (next) {
      if (r0 + 2 * step < rows) fd_stage_load<MODE, CC>(v0, r0 + 2 * step, rows, cols, rowptr, cstride, wave, lane);
      fd_stage_store<MODE, CC>(dst, p, v1, r0 + step, rows, cols, wave, lane);
    }
  }
}

// Flat staging of a rows x cols operand through registers, for operands of at most
// 256 E elements: thread t holds elements t + 256 i (i < E), element e at row e / cols,
// column e % cols, so no lane idles on a short row.  load() issues every load (clamped
// past the end, so always in bounds) and store() writes them to LDS later: the caller
// places other work, or other loads, between the two, and a stage costs one memory
// round trip whatever its row count.
template <int MODE, int E>
struct FdFlat {
  using T = typename FdMma<MODE>::T;
  T v[E];
  template <typename RowPtr>
  __device__ void load(int rows, int cols, const RowPtr& rowptr, int64_t cstride, int tid) {
    const int q = cols <= 256 ? 256 / cols : 0, rr = cols <= 256 ? 256 % cols : 256;
    int r = tid / cols, c = tid % cols;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const bool ok = r < rows;
      v[i] = rowptr(ok ? r : rows - 1)[(int64_t)(ok ? c : cols - 1) * cstride];
      r += q;
      c += rr;
      if (c >= cols) {
        c -= cols;
        ++r;
      }
    }
  }
  __device__ void store(typename FdMma<MODE>::S* dst, int p, int rows, int cols, int tid) const {
    const int q = cols <= 256 ? 256 / cols : 0, rr = cols <= 256 ? 256 % cols : 256;
    int r = tid / cols, c = tid % cols;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if (r < rows) dst[r * p + c] = FdMma<MODE>::to_s(v[i]);
      r += q;
      c += rr;
      if (c >= cols) {
        c -= cols;
        ++r;
      }
    }
  }
};

// Zero columns [c0, c1) of rows [0, rows) (c1 - c0 < 64): the K padding of an operand.
template <int MODE>
__device__ inline void fd_zero_cols(typename FdMma<MODE>::S* dst, int p, int rows, int c0, int c1, int wave,
                                    int lane) {
  if (lane < c1 - c0)
    for (int r = wave; r < rows; r += 4) dst[r * p + c0 + lane] = FdMma<MODE>::to_s(0);
}

template <typename T>
struct FdPlaneArgs {
  int n1, n2, n3, nrhs;
  // element e of right-hand side k of the input / output at e * es + k * ks: interleaved
  // (N, nrhs) arrays have es = nrhs, ks = 1; right-hand-side-major ones es = 1, ks = N
  int64_t es_in, ks_in, es_out, ks_out;
  int64_t nf1, nf2, nf3;  // full lengths of the three directions
  int64_t lo1, lo2, lo3;  // box start
  const T* in;            // forward: full x; backward: box-dense W
  T* out;                 // forward: box-dense W; backward: full y
  const T* A1;            // n1 x n1 row-major: Y1(i1, .) = sum_k1 A1[i1][k1] X(k1, .)
  const T* A2;            // n2 x n2 row-major: Y(., i2) = sum_k2 A2[i2][k2] Y1(., k2)
  int p;                  // LDS row pitch (elements of the LDS operand type)
};

// LDS: P (n2 rows x p: the plane X[k2][k1], then A2[i2][k2]) and Q (16 rows x p: the A1
// rows of S, then Y1[i1][k2]); then the output tile staging [i2][17] in T over P.  Rows
// are zero-padded in k up to a multiple of KS.  Tiles past n2 read clamped (finite) rows:
// their garbage stays in output columns that are never stored, and Y1's columns past n2
// are stored as zeros.  TPW = ceil(ceil(n2 / 16) / 4) tiles per wave.  (Staging the
// operands flat, all loads at once, measured no faster for FP64 and slower for FP32 /
// BF16 than these row rounds.)
template <int MODE, bool BWD, int TPW>
__global__ void __launch_bounds__(256) fd_plane_kernel(FdPlaneArgs<typename FdMma<MODE>::T> a) {
  using M = FdMma<MODE>;
  using T = typename M::T;
  using V = typename M::V;
  using S = typename M::S;
  extern __shared__ __align__(16) unsigned char fd_smem[];
  S* P = reinterpret_cast<S*>(fd_smem);
  const int p = a.p, n1 = a.n1, n2 = a.n2;
  S* Q = P + (int64_t)n2 * p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = blockIdx.z;
  if (BWD && (int)blockIdx.y >= a.n3) {
    // a plane of y outside the box in direction 3: zeros, split over blockIdx.x
    const int t = blockIdx.y - a.n3;
    const int64_t i3f = t < a.lo3 ? t : a.lo3 + a.n3 + (t - a.lo3);
    const int64_t np = a.nf1 * a.nf2;
    for (int64_t q = (int64_t)blockIdx.x * 256 + tid; q < np; q += (int64_t)gridDim.x * 256)
      a.out[(q + np * i3f) * a.es_out + k * a.ks_out] = T(0);
    return;
  }
  const int i3 = blockIdx.y;
  const int i1s = blockIdx.x * 16;
  const int K1 = (n1 + M::KS - 1) / M::KS * M::KS;  // zero-padded inner lengths
  const int K2 = (n2 + M::KS - 1) / M::KS * M::KS;
  // the plane X[k2][k1] (rows k2, k1 contiguous in both layouts) and the A1 rows of S
  const int64_t row0 = BWD ? (int64_t)n1 * n2 * i3 * a.es_in + k * a.ks_in
                           : ((a.lo1) + a.nf1 * (a.lo2 + a.nf2 * (a.lo3 + i3))) * a.es_in + k * a.ks_in;
  const int64_t rstep = BWD ? (int64_t)n1 * a.es_in : a.nf1 * a.es_in;
  auto xrow = [&](int r) { return a.in + row0 + rstep * r; };
  auto a1row = [&](int r) { return a.A1 + (int64_t)(i1s + r) * n1; };
  auto a2row = [&](int r) { return a.A2 + (int64_t)r * n2; };
  const int qrows = min(16, n1 - i1s);  // rows of S inside the box; the rest are zero
  fd_stage<MODE, TPW>(P, p, n2, n1, xrow, a.es_in, wave, lane);
  fd_stage<MODE, TPW>(Q, p, qrows, n1, a1row, 1, wave, lane);
  for (int q = tid; q < (16 - qrows) * p; q += 256) Q[qrows * p + q] = M::to_s(T(0));
  fd_zero_cols<MODE>(P, p, n2, n1, K1, wave, lane);
  fd_zero_cols<MODE>(Q, p, qrows, n1, K1, wave, lane);
  // the first round of A2 is loaded now and stays in flight during the first product
  T a2pre[FdStageShape<TPW>::RG][TPW];
  fd_stage_load<MODE, TPW>(a2pre, 0, n2, n2, a2row, 1, wave, lane);
  __syncthreads();
  const int lr = lane & 15;
  const S* brow[TPW];  // this lane's B row (its output column) in each tile, clamped
#pragma unroll
  for (int u = 0; u < TPW; ++u) brow[u] = P + min((wave + 4 * u) * 16 + lr, n2 - 1) * p;
  const S* arow = Q + lr * p;
  V acc[TPW];
  // step 1: Y1(i1, k2) = sum_k1 A1[i1][k1] X(k1, k2); M = 16 rows of S, N = k2 tiles
#pragma unroll
  for (int u = 0; u < TPW; ++u) acc[u] = V{0, 0, 0, 0};
  fd_mma_loop<MODE, TPW, 1>(
      acc, K1, [&](int k0) { return M::op(arow, k0, lane); }, [&](int u, int k0) { return M::op(brow[u], k0, lane); });
  __syncthreads();
  // Y1 into Q (columns up to K2; past n2 zeros), A2 into P
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = (wave + 4 * u) * 16 + lr;
      if (col < K2) Q[M::row(lane, r) * p + col] = M::to_s(col < n2 ? acc[u][r] : T(0));
    }
  }
  fd_stage<MODE, TPW>(P, p, n2, n2, a2row, 1, wave, lane, &a2pre);
  fd_zero_cols<MODE>(P, p, n2, n2, K2, wave, lane);
  __syncthreads();
  // step 2: Y(i1, i2) = sum_k2 Y1(i1, k2) A2[i2][k2]; B(k2, i2) = A2[i2][k2]
#pragma unroll
  for (int u = 0; u < TPW; ++u) acc[u] = V{0, 0, 0, 0};
  fd_mma_loop<MODE, TPW, 0>(
      acc, K2, [&](int k0) { return M::op(arow, k0, lane); }, [&](int u, int k0) { return M::op(brow[u], k0, lane); });
  __syncthreads();
  // stage the 16 x n2 result as [i2][i1] (pitch 17, type T) for writes along direction 1
  constexpr int RP = 17;
  T* R = reinterpret_cast<T*>(fd_smem);
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = (wave + 4 * u) * 16 + lr;
      if (col < n2) R[col * RP + M::row(lane, r)] = acc[u][r];
    }
  }
  __syncthreads();
  for (int q = tid; q < 16 * n2; q += 256) {
    const int il = q & 15, i2 = q >> 4, i1 = i1s + il;
    if (i1 < n1) {
      const T v = R[i2 * RP + il];
      if constexpr (BWD)
        a.out[((a.lo1 + i1) + a.nf1 * ((a.lo2 + i2) + a.nf2 * (a.lo3 + i3))) * a.es_out + k * a.ks_out] = v;
      else
        a.out[((int64_t)i1 + (int64_t)n1 * ((int64_t)i2 + (int64_t)n2 * i3)) * a.es_out + k * a.ks_out] = v;
    }
  }
  if (BWD && blockIdx.x == 0) {
    // entries of this plane of y outside the box in directions 1 and 2: rows over the
    // waves, columns over the lanes
    const int nf1 = (int)a.nf1, nf2 = (int)a.nf2;
    T* pl = a.out + (a.nf1 * a.nf2 * (a.lo3 + i3)) * a.es_out + k * a.ks_out;
    for (int i2f = wave; i2f < nf2; i2f += 4) {
      const bool row_in = i2f >= a.lo2 && i2f < a.lo2 + n2;
      for (int i1f = lane; i1f < nf1; i1f += 64)
        if (!row_in || i1f < a.lo1 || i1f >= a.lo1 + n1) pl[((int64_t)i2f * nf1 + i1f) * a.es_out] = T(0);
    }
  }
}

template <typename T>
struct FdLineArgs {
  int n3, p, xp;
  int nrhs;        // right-hand sides interleaved in the columns of W (1: one per grid row)
  int64_t ncol;    // n1 n2 nrhs columns of W
  int64_t plane;   // n1 n2: stride of the eigen-index k3 in Lambda
  int64_t kstride; // right-hand-side-major W: elements per right-hand side (grid row y)
  T* w;            // box-dense (row i3 of ncol values), overwritten
  const T* Ut;     // U3^T row-major n3 x n3: Ut[k3][i3]
  const T* inv;    // Lambda^-1 over the box (direction 1 fastest)
};

// Both products are run transposed, with the 16 columns of the work-group as the shared
// A operand and the n3 rows spread over the tiles of the four waves:
//   Z^T(j, k3) = sum_i3 X^T(j, i3) U3^T[k3][i3],   Y^T(j, i3) = sum_k3 Z^T(j, k3) U3[i3][k3]
// LDS: U (K3 rows x p: U3^T[k3][i3], rows past n3 zero; the second product reads it
// transposed) and X (16 rows x xp: X^T[j][i3], then Z^T[j][k3]).  Tile rows past n3 are clamped: their garbage
// stays in output columns that are never stored, and Z^T's columns past n3 are zeros.
// E > 0 (n3 <= 128, U3^T of at most 256 E elements): every global load of the kernel
// (U3^T flat, the 16 columns of W, the Lambda^-1 entries of its outputs) is issued
// before the first LDS store, one round trip; E = 0: row rounds.
template <int MODE, int TPW, int E>
__global__ void __launch_bounds__(256) fd_line_kernel(FdLineArgs<typename FdMma<MODE>::T> a) {
  using M = FdMma<MODE>;
  using T = typename M::T;
  using V = typename M::V;
  using S = typename M::S;
  constexpr bool kTwoCopies = false;  // one copy of U3^T; the second product reads it transposed
  extern __shared__ __align__(16) unsigned char fd_smem[];
  const int n3 = a.n3, p = a.p, xp = a.xp;
  const int K3 = (n3 + M::KS - 1) / M::KS * M::KS;
  S* U = reinterpret_cast<S*>(fd_smem);
  S* Ub = U + (int64_t)K3 * p;
  S* X = Ub + (kTwoCopies ? (int64_t)n3 * p : 0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t j0 = (int64_t)blockIdx.x * 16;
  T* w = a.w + (int64_t)blockIdx.y * a.kstride;  // this grid row's right-hand side
  const int lr = lane & 15;
  int mrow[TPW];  // this lane's row of U3 / U3^T in each of its tiles, clamped
#pragma unroll
  for (int u = 0; u < TPW; ++u) mrow[u] = min((wave + 4 * u) * 16 + lr, n3 - 1);
  // Lambda^-1 of this lane's outputs of the first product (eigen-index k3 = tile row,
  // column j), loaded ahead
  T linv[TPW][4];
  auto load_inv = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < TPW; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t j = j0 + M::row(lane, r);
        const int64_t c = (j < a.ncol ? j : a.ncol - 1) / a.nrhs;
        linv[u][r] = a.inv[c + a.plane * mrow[u]];
      }
  };
  auto utrow = [&](int r) { return a.Ut + (int64_t)r * n3; };
  const int jc = tid & 15;
  const int64_t jg = min(j0 + jc, a.ncol - 1);
  if constexpr (E > 0) {
    // U3^T, the 16 columns of W (X[j][i3] = W(i3, j0 + j): 16 threads per W row, rows
    // i3 = tid / 16 + 16 e) and Lambda^-1: all loads first
    FdFlat<MODE, E> uflat;
    T xv[8];
    uflat.load(n3, n3, utrow, 1, tid);
#pragma unroll
    for (int e = 0; e < 8; ++e) xv[e] = w[(int64_t)min((tid >> 4) + 16 * e, n3 - 1) * a.ncol + jg];
    load_inv();
    uflat.store(U, p, n3, n3, tid);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int r = (tid >> 4) + 16 * e;
      if (r < n3) X[jc * xp + r] = M::to_s(j0 + jc < a.ncol ? xv[e] : T(0));
    }
  } else {
    fd_stage<MODE, TPW>(U, p, n3, n3, utrow, 1, wave, lane);
  }
  if constexpr (kTwoCopies) {
    // Ub[i3][k3] = U3^T[k3][i3]: row i3 of Ub is column i3 of U3^T
    fd_stage<MODE, TPW>(Ub, p, n3, n3, [&](int r) { return a.Ut + r; }, n3, wave, lane);
    fd_zero_cols<MODE>(Ub, p, n3, n3, K3, wave, lane);
  }
  for (int q = tid; q < (K3 - n3) * p; q += 256) U[n3 * p + q] = M::to_s(T(0));  // rows past n3
  fd_zero_cols<MODE>(U, p, n3, n3, K3, wave, lane);
  if constexpr (E == 0) {
    // X[j][i3] = W(i3, j0 + j): 16 threads per W row, rows i3 = tid / 16 + 16 e
    for (int r0 = 0; r0 < n3; r0 += 16 * 8) {
      T v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = w[(int64_t)min(r0 + (tid >> 4) + 16 * e, n3 - 1) * a.ncol + jg];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int r = r0 + (tid >> 4) + 16 * e;
        if (r < n3) X[jc * xp + r] = M::to_s(j0 + jc < a.ncol ? v[e] : T(0));
      }
    }
    load_inv();
  }
  for (int q = tid; q < (K3 - n3) * 16; q += 256) X[(q & 15) * xp + n3 + (q >> 4)] = M::to_s(T(0));
  __syncthreads();
  const S* arow = X + lr * xp;
  V acc[TPW];
  // step 1: Z^T(j, k3) = sum_i3 X^T(j, i3) U3^T[k3][i3]
#pragma unroll
  for (int u = 0; u < TPW; ++u) acc[u] = V{0, 0, 0, 0};
  fd_mma_loop<MODE, TPW, 0>(
      acc, K3, [&](int k0) { return M::op(arow, k0, lane); },
      [&](int u, int k0) { return M::op(U + mrow[u] * p, k0, lane); });
  __syncthreads();
  // scale by Lambda^-1 and store Z^T over X (columns past n3 zero)
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int m = (wave + 4 * u) * 16 + lr;  // k3
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int jr = M::row(lane, r);
      const int64_t j = j0 + jr;
      if (m < K3) X[jr * xp + m] = M::to_s((m < n3 && j < a.ncol) ? acc[u][r] * linv[u][r] : T(0));
    }
  }
  __syncthreads();
  // step 2: Y^T(j, i3) = sum_k3 Z^T(j, k3) U3[i3][k3]  (U3[i3][k3] = U3^T[k3][i3])
#pragma unroll
  for (int u = 0; u < TPW; ++u) acc[u] = V{0, 0, 0, 0};
  if constexpr (kTwoCopies) {
    fd_mma_loop<MODE, TPW, 0>(
        acc, K3, [&](int k0) { return M::op(arow, k0, lane); },
        [&](int u, int k0) { return M::op(Ub + mrow[u] * p, k0, lane); });
  } else {
    fd_mma_loop<MODE, TPW, 0>(
        acc, K3, [&](int k0) { return M::op(arow, k0, lane); },
        [&](int u, int k0) { return M::op_t(U, p, mrow[u], k0, lane); });
  }
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int m = (wave + 4 * u) * 16 + lr;  // i3
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t j = j0 + M::row(lane, r);
      if (m < n3 && j < a.ncol) w[(int64_t)m * a.ncol + j] = acc[u][r];
    }
  }
}

// LDS sizes and row pitches of the two kernels for box lengths len[0..2].  Pitches: k
// extent rounded to the K step, plus one element (FP64 / FP32, odd: conflict-free
// strided reads) or four (BF16: rows on 8-byte boundaries).
struct Fd3Shape {
  int p_plane, p_line, xp_line;
  size_t lds_plane, lds_line;
};

static Fd3Shape fd3_shape(const int64_t* len, int mode) {
  const int ks = mode == FD_BF16 ? 16 : 8;
  const size_t es = mode == FD_F64 ? 8 : 4;
  const size_t et = mode == FD_F64 ? 8 : 4;
  auto kpad = [&](int64_t n) { return (int)((n + ks - 1) / ks * ks); };
  auto pitch = [&](int64_t n) { return kpad(n) + (mode == FD_BF16 ? 4 : 1); };
  Fd3Shape s{};
  s.p_plane = pitch(std::max(len[0], len[1]));
  s.lds_plane = std::max(es * (size_t)(len[1] + 16) * s.p_plane, et * (size_t)len[1] * 17);
  s.p_line = pitch(len[2]);
  s.xp_line = pitch(len[2]);
  const size_t urows = (size_t)kpad(len[2]);
  s.lds_line = es * (urows * s.p_line + 16 * (size_t)s.xp_line);
  return s;
}

constexpr size_t kFdLdsMax = 160 * 1024;

static bool fd3_applicable(const iga_fd_s* f, int mode) {
  if (f->ndim != 3) return false;
  for (int d = 0; d < 3; ++d)
    if (f->len[d] < 1 || f->len[d] > 4 * kFdMaxTiles * 16) return false;
  const Fd3Shape s = fd3_shape(f->len, mode);
  return s.lds_plane <= kFdLdsMax && s.lds_line <= kFdLdsMax;
}

// Dynamic LDS above 64 KiB must be granted per kernel; done once per kernel and size.
template <typename Kern>
static void fd3_allow_lds(Kern kern, size_t bytes, std::atomic<size_t>& granted) {
  if (bytes <= 64 * 1024 || bytes <= granted.load(std::memory_order_relaxed)) return;
  IGA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes));
  granted.store(bytes, std::memory_order_relaxed);
}

template <int MODE, int TPW, int E>
static void fd3_launch(const iga_fd_s* f, FdPlaneArgs<typename FdMma<MODE>::T> pa,
                       FdLineArgs<typename FdMma<MODE>::T> la, const Fd3Shape& sh, int nrhs, int which,
                       hipStream_t s) {
  static std::atomic<size_t> granted[3];
  const unsigned nsb = (unsigned)((pa.n1 + 15) / 16);
  if (which == 0) {
    fd3_allow_lds(fd_plane_kernel<MODE, false, TPW>, sh.lds_plane, granted[0]);
    fd_plane_kernel<MODE, false, TPW><<<dim3(nsb, (unsigned)pa.n3, (unsigned)nrhs), 256, sh.lds_plane, s>>>(pa);
  } else if (which == 1) {
    fd3_allow_lds(fd_line_kernel<MODE, TPW, E>, sh.lds_line, granted[1]);
    const unsigned ky = la.kstride > 0 ? (unsigned)nrhs : 1u;
    fd_line_kernel<MODE, TPW, E><<<dim3((unsigned)((la.ncol + 15) / 16), ky), 256, sh.lds_line, s>>>(la);
  } else {
    const unsigned extra = (unsigned)(f->n[2] - pa.n3);
    fd3_allow_lds(fd_plane_kernel<MODE, true, TPW>, sh.lds_plane, granted[2]);
    fd_plane_kernel<MODE, true, TPW><<<dim3(nsb, (unsigned)pa.n3 + extra, (unsigned)nrhs), 256, sh.lds_plane, s>>>(
        pa);
  }
  IGA_LAUNCH_CHECK();
}

// Tiles per wave and, for the line kernel, the flat staging size from the longest
// operand side n of a launch: an n x n operand takes ceil(n^2 / 256) registers per
// thread when staged flat (up to n = 128); longer sides stage in row rounds.
template <int MODE>
static void fd3_launch_len(const iga_fd_s* f, FdPlaneArgs<typename FdMma<MODE>::T> pa,
                           FdLineArgs<typename FdMma<MODE>::T> la, const Fd3Shape& sh, int nrhs, int which,
                           int n, hipStream_t s) {
  if (n <= 64)
    fd3_launch<MODE, 1, 16>(f, pa, la, sh, nrhs, which, s);
  else if (n <= 78)
    fd3_launch<MODE, 2, 24>(f, pa, la, sh, nrhs, which, s);
  else if (n <= 101)
    fd3_launch<MODE, 2, 40>(f, pa, la, sh, nrhs, which, s);
  else if (n <= 128)
    fd3_launch<MODE, 2, 64>(f, pa, la, sh, nrhs, which, s);
  else
    fd3_launch<MODE, 3, 0>(f, pa, la, sh, nrhs, which, s);
}

// rhs_major: x, y and the box-dense intermediate hold right-hand side k at k * N
// (N, the length of one); otherwise (N, nrhs) interleaved.
template <int MODE>
static void fd3_apply(const iga_fd_s* f, const typename FdMma<MODE>::T* x, typename FdMma<MODE>::T* y, int nrhs,
                      typename FdMma<MODE>::T* w, hipStream_t s, bool rhs_major = false) {
  using T = typename FdMma<MODE>::T;
  IGA_CHECK_ARG(nrhs <= 65535, "too many right-hand sides for one FD launch");
  const Fd3Shape sh = fd3_shape(f->len, MODE);
  const int n1 = (int)f->len[0], n2 = (int)f->len[1], n3 = (int)f->len[2];
  FdPlaneArgs<T> pa{};
  pa.n1 = n1;
  pa.n2 = n2;
  pa.n3 = n3;
  pa.nrhs = nrhs;
  pa.nf1 = f->n[0];
  pa.nf2 = f->n[1];
  pa.nf3 = f->n[2];
  pa.lo1 = f->lo[0];
  pa.lo2 = f->lo[1];
  pa.lo3 = f->lo[2];
  pa.p = sh.p_plane;
  const int64_t nfull = f->n[0] * f->n[1] * f->n[2], nbox = (int64_t)n1 * n2 * n3;
  FdLineArgs<T> la{};
  la.n3 = n3;
  la.p = sh.p_line;
  la.xp = sh.xp_line;
  la.plane = (int64_t)n1 * n2;
  la.ncol = rhs_major ? la.plane : la.plane * nrhs;
  la.nrhs = rhs_major ? 1 : nrhs;
  la.kstride = rhs_major ? nbox : 0;
  la.w = w;
  la.Ut = (const T*)f->d_Ut[2];
  la.inv = (const T*)f->d_inv;
  // forward: x -> W (directions 1, 2)
  pa.in = x;
  pa.out = w;
  pa.es_in = rhs_major ? 1 : nrhs;
  pa.ks_in = rhs_major ? nfull : 1;
  pa.es_out = rhs_major ? 1 : nrhs;
  pa.ks_out = rhs_major ? nbox : 1;
  pa.A1 = (const T*)f->d_Ut[0];
  pa.A2 = (const T*)f->d_Ut[1];
  // the plane kernels' tiles per wave also set their column passes: cover both lengths
  const int plane_len = std::max(n1, n2);
  fd3_launch_len<MODE>(f, pa, la, sh, nrhs, 0, plane_len, s);
  // direction 3 with the eigenvalue scaling, in place on W
  fd3_launch_len<MODE>(f, pa, la, sh, nrhs, 1, n3, s);
  // backward: W -> y (directions 1, 2), zeros outside the box
  pa.in = w;
  pa.out = y;
  pa.ks_in = rhs_major ? nbox : 1;
  pa.ks_out = rhs_major ? nfull : 1;
  pa.A1 = (const T*)f->d_U[0];
  pa.A2 = (const T*)f->d_U[1];
  fd3_launch_len<MODE>(f, pa, la, sh, nrhs, 2, plane_len, s);
}

template <typename T>
static void fd_apply_t(const iga_fd_s* f, const T* x, T* y, int nrhs, T* w0, T* w1, hipStream_t s) {
  // gather the free box
  int64_t n[4], lo[4], len[4];
  for (int d = 0; d < f->ndim; ++d) {
    n[d] = f->n[d];
    lo[d] = f->lo[d];
    len[d] = f->len[d];
  }
  const int dt = std::is_same<T, double>::value ? IGA_F64 : IGA_F32;
  int rc = iga_box_gather(dt, f->ndim, n, lo, len, nrhs, x, w0, s);
  if (rc != IGA_OK) throw Error(rc, "box gather failed");
  T* cur = w0;
  T* nxt = w1;
  auto pass = [&](int d, bool transpose, const T* scale) {
    int64_t outer = 1, inner = nrhs;
    for (int e = d + 1; e < f->ndim; ++e) outer *= f->len[e];
    for (int e = 0; e < d; ++e) inner *= f->len[e];
    const T* A = (const T*)(transpose ? f->d_Ut[d] : f->d_U[d]);
    mode_gemm<T>(A, cur, nxt, outer, f->len[d], inner, scale, s);
    std::swap(cur, nxt);
  };
  // the epilogue scale indexes the free box without the right-hand-side index,
  // so it is fused only for one right-hand side
  for (int d = 0; d < f->ndim; ++d) pass(d, true, (d == f->ndim - 1 && nrhs == 1) ? (const T*)f->d_inv : nullptr);
  if (nrhs > 1) {
    rc = iga_qp_scale(dt, f->nsub, nrhs, f->d_inv, cur, cur, 0, 1.0, s);
    if (rc != IGA_OK) throw Error(rc, "eigenvalue scaling failed");
  }
  for (int d = f->ndim - 1; d >= 0; --d) pass(d, false, nullptr);
  rc = iga_box_scatter(dt, f->ndim, n, lo, len, nrhs, cur, y, 1, s);
  if (rc != IGA_OK) throw Error(rc, "box scatter failed");
}

}  // namespace iga

extern "C" {

int iga_fd_create(iga_fd_t* out, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                  const double* const* U, int dtype) {
  return guard([&] {
    IGA_CHECK_ARG(out && n && lo && len && U, "NULL argument");
    IGA_CHECK_ARG(ndim >= 1 && ndim <= 4, "ndim must be in 1..4");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32 || dtype == IGA_BF16, "unknown dtype");
    auto* f = new iga_fd_s();
    f->dtype = dtype;
    f->ndim = ndim;
    f->nsub = 1;
    try {
      const size_t es = dtype == IGA_F64 ? 8 : 4;
      size_t total = 0;
      for (int d = 0; d < ndim; ++d) {
        IGA_CHECK_ARG(len[d] >= 1 && lo[d] >= 0 && lo[d] + len[d] <= n[d], "free box does not fit");
        f->n[d] = n[d];
        f->lo[d] = lo[d];
        f->len[d] = len[d];
        f->nsub *= len[d];
        total += 2 * (((size_t)len[d] * len[d] * es + 255) & ~size_t(255));
      }
      // every U_d, U_d^T in one block and one transfer
      std::vector<char> host(total);
      size_t off = 0;
      size_t offs[4][2];
      for (int d = 0; d < ndim; ++d) {
        const int64_t m = len[d];
        for (int which = 0; which < 2; ++which) {
          offs[d][which] = off;
          for (int64_t i = 0; i < m; ++i)
            for (int64_t k = 0; k < m; ++k) {
              // which 0: U[i][k] (row-major, as numpy); 1: U^T[i][k]
              const double v = which == 0 ? U[d][i * m + k] : U[d][k * m + i];
              if (dtype == IGA_F64)
                reinterpret_cast<double*>(host.data() + off)[i * m + k] = v;
              else
                reinterpret_cast<float*>(host.data() + off)[i * m + k] = (float)v;
            }
          off += ((size_t)m * m * es + 255) & ~size_t(255);
        }
      }
      IGA_HIP(hipMalloc(&f->d_base, std::max<size_t>(total, 256)));
      IGA_HIP(hipMemcpy(f->d_base, host.data(), total, hipMemcpyHostToDevice));
      for (int d = 0; d < ndim; ++d) {
        f->d_U[d] = (char*)f->d_base + offs[d][0];
        f->d_Ut[d] = (char*)f->d_base + offs[d][1];
      }
    } catch (...) {
      (void)hipFree(f->d_base);
      delete f;
      throw;
    }
    *out = f;
  });
}

int iga_fd_set_eigenvalues(iga_fd_t f, const double* eigenvalues) {
  return guard([&] {
    IGA_CHECK_ARG(f && eigenvalues, "NULL argument");
    std::vector<double> inv((size_t)f->nsub);
    for (int64_t i = 0; i < f->nsub; ++i) {
      IGA_CHECK_ARG(eigenvalues[i] != 0.0, "zero eigenvalue");
      inv[i] = 1.0 / eigenvalues[i];
    }
    const size_t es = f->dtype == IGA_F64 ? 8 : 4;
    if (!f->d_inv) IGA_HIP(hipMalloc(&f->d_inv, es * f->nsub));
    if (f->dtype == IGA_F64) {
      IGA_HIP(hipMemcpy(f->d_inv, inv.data(), 8 * f->nsub, hipMemcpyHostToDevice));
    } else {
      std::vector<float> i32(inv.begin(), inv.end());
      IGA_HIP(hipMemcpy(f->d_inv, i32.data(), 4 * f->nsub, hipMemcpyHostToDevice));
    }
  });
}

int64_t iga_fd_workspace_bytes(iga_fd_t f, int nrhs) {
  if (!f || nrhs < 1) return 0;
  return 2 * f->nsub * (int64_t)nrhs * (f->dtype == IGA_F64 ? 8 : 4) + 512;
}

int iga_fd_apply(iga_fd_t f, const void* x, void* y, int nrhs, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && x && y && work && nrhs >= 1, "NULL argument");
    IGA_CHECK_ARG(f->d_inv, "eigenvalues are not set");
    const int64_t half = f->nsub * nrhs;
    const int mode = f->dtype == IGA_F64 ? FD_F64 : f->dtype == IGA_F32 ? FD_F32 : FD_BF16;
    if (fd3_applicable(f, mode)) {
      if (mode == FD_F64)
        fd3_apply<FD_F64>(f, (const double*)x, (double*)y, nrhs, (double*)work, as_stream(stream));
      else if (mode == FD_F32)
        fd3_apply<FD_F32>(f, (const float*)x, (float*)y, nrhs, (float*)work, as_stream(stream));
      else
        fd3_apply<FD_BF16>(f, (const float*)x, (float*)y, nrhs, (float*)work, as_stream(stream));
      return;
    }
    IGA_CHECK_ARG(f->dtype != IGA_BF16, "the BF16 FD apply takes three directions whose box planes fit the "
                                         "LDS (lengths up to about 190)");
    if (f->dtype == IGA_F64) {
      double* w0 = (double*)work;
      fd_apply_t<double>(f, (const double*)x, (double*)y, nrhs, w0, w0 + half, as_stream(stream));
    } else {
      float* w0 = (float*)work;
      fd_apply_t<float>(f, (const float*)x, (float*)y, nrhs, w0, w0 + half, as_stream(stream));
    }
  });
}

int iga_fd_apply_rhs_major(iga_fd_t f, const void* x, void* y, int nrhs, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && x && y && work && nrhs >= 1, "NULL argument");
    IGA_CHECK_ARG(f->d_inv, "eigenvalues are not set");
    const int mode = f->dtype == IGA_F64 ? FD_F64 : f->dtype == IGA_F32 ? FD_F32 : FD_BF16;
    if (fd3_applicable(f, mode)) {
      if (mode == FD_F64)
        fd3_apply<FD_F64>(f, (const double*)x, (double*)y, nrhs, (double*)work, as_stream(stream), true);
      else if (mode == FD_F32)
        fd3_apply<FD_F32>(f, (const float*)x, (float*)y, nrhs, (float*)work, as_stream(stream), true);
      else
        fd3_apply<FD_BF16>(f, (const float*)x, (float*)y, nrhs, (float*)work, as_stream(stream), true);
      return;
    }
    // other shapes: one apply per right-hand side
    int64_t nfull = 1;
    for (int d = 0; d < f->ndim; ++d) nfull *= f->n[d];
    const size_t es = f->dtype == IGA_F64 ? 8 : 4;
    for (int k = 0; k < nrhs; ++k)
      IGA_CHECK(iga_fd_apply(f, (const char*)x + (size_t)k * nfull * es, (char*)y + (size_t)k * nfull * es, 1, work,
                             stream));
  });
}

int iga_fd_destroy(iga_fd_t f) {
  return guard([&] {
    if (!f) return;
    IGA_HIP(hipFree(f->d_base));
    IGA_HIP(hipFree(f->d_inv));
    delete f;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Space-time arrowhead FD (time slowest):
//   y[free] = Re[(U_s (x) U_t) A^{-1} (U_s (x) U_t)^H x[free]],  y = 0 elsewhere,
// where, per spatial eigenvalue lambda_j, A_j = [[diag(D_i + lambda_j), b], [-b^H, D_L + lambda_j]]
// is the arrowhead of the time pencil (D complex, b complex).  The transposed apply
// swaps U_t and conj(U_t) and uses b -> -conj(b).  The space transforms are real;
// the time transform maps the real tensor to a complex one stored as a real block
// followed by an imaginary block, so every product is a real MFMA GEMM:
//   forward  [Y_re; Y_im] = F X,  F = [Re U_t^H; Im U_t^H]       (2L x L)
//   backward Re(U_t Z)    = G [Z_re; Z_im],  G = [Re U_t, -Im U_t]  (L x 2L)
// ---------------------------------------------------------------------------

namespace iga {

constexpr int kArrowMax = 512;  // time eigen-indices held in LDS

// One thread per column j (spatial eigen-index j / nrhs).  z holds the complex
// tensor as [2][nt][ncol] and is overwritten with the solution.
template <typename T>
__global__ void __launch_bounds__(256) arrowhead_kernel(int nt, int64_t ncol, int nrhs, const T* __restrict__ lam,
                                                         const T* __restrict__ dre, const T* __restrict__ dim,
                                                         const T* __restrict__ bre, const T* __restrict__ bim,
                                                         T* __restrict__ z) {
  __shared__ T sd[2][kArrowMax];
  __shared__ T sb[2][kArrowMax];
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    sd[0][i] = dre[i];
    sd[1][i] = dim[i];
    sb[0][i] = i < nt - 1 ? bre[i] : T(0);
    sb[1][i] = i < nt - 1 ? bim[i] : T(0);
  }
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  const T lj = lam[j / nrhs];
  T* zr = z + j;
  T* zi = z + (int64_t)nt * ncol + j;
  const int L = nt - 1;
  // Schur complement S = H_L + sum |b_i|^2 / H_i and  a = sum conj(b_i) r_i / H_i
  T s_re = sd[0][L] + lj, s_im = sd[1][L];
  T a_re = T(0), a_im = T(0);
  for (int i = 0; i < L; ++i) {
    const T hr = sd[0][i] + lj, hi = sd[1][i];
    const T den = T(1) / (hr * hr + hi * hi);
    const T ir = hr * den, ii = -hi * den;  // 1 / H_i
    const T br = sb[0][i], bi = sb[1][i];
    const T bb = br * br + bi * bi;
    s_re += bb * ir;
    s_im += bb * ii;
    const T rr = zr[(int64_t)i * ncol], ri = zi[(int64_t)i * ncol];
    const T yr = rr * ir - ri * ii, yi = rr * ii + ri * ir;
    a_re += br * yr + bi * yi;
    a_im += br * yi - bi * yr;
  }
  const T lr = zr[(int64_t)L * ncol] + a_re, li = zi[(int64_t)L * ncol] + a_im;
  const T sden = T(1) / (s_re * s_re + s_im * s_im);
  const T xr = (lr * s_re + li * s_im) * sden, xi = (li * s_re - lr * s_im) * sden;
  zr[(int64_t)L * ncol] = xr;
  zi[(int64_t)L * ncol] = xi;
  // x_i = (r_i - b_i x_L) / H_i
  for (int i = 0; i < L; ++i) {
    const T hr = sd[0][i] + lj, hi = sd[1][i];
    const T den = T(1) / (hr * hr + hi * hi);
    const T br = sb[0][i], bi = sb[1][i];
    const T tr = zr[(int64_t)i * ncol] - (br * xr - bi * xi);
    const T ti = zi[(int64_t)i * ncol] - (br * xi + bi * xr);
    zr[(int64_t)i * ncol] = (tr * hr + ti * hi) * den;
    zi[(int64_t)i * ncol] = (ti * hr - tr * hi) * den;
  }
}

}  // namespace iga

struct iga_stfd_s {
  int dtype = IGA_F64;
  int64_t n[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0}, len[4] = {0, 0, 0, 0};  // 3 space + time
  int64_t nspace = 0, nsub = 0;
  void* d_Ut[3] = {nullptr, nullptr, nullptr};
  void* d_U[3] = {nullptr, nullptr, nullptr};
  void* d_F[2] = {nullptr, nullptr};  // [apply, transpose]
  void* d_G[2] = {nullptr, nullptr};
  void* d_lam = nullptr;              // spatial eigenvalues over the space box
  void* d_D[2] = {nullptr, nullptr};  // adv * D  (re, im), length L
  void* d_b[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [apply/transpose][re/im], length L-1
  // legacy QZ form (time_legacy.py): adv S and T of the QZ pencil, row-major L x L (re, im)
  int qz = 0;
  void* d_S[2] = {nullptr, nullptr};
  void* d_T[2] = {nullptr, nullptr};
  bool blocks_set = false;
  std::vector<void*> owned;
};

namespace iga {

// Legacy QZ middle solve (time_legacy.py:88-101, 146-156): per column j (spatial
// eigen-index j / nrhs) the triangular system (adv S + lambda_j T) y = r, upper and by
// back substitution, or transposed (lower, forward substitution), complex, in place on
// z = [2][nt][ncol].  The matrix entries are the same for every thread (uniform loads).
template <typename T>
__global__ void __launch_bounds__(256) qz_solve_kernel(int nt, int64_t ncol, int nrhs, int transpose,
                                                        const T* __restrict__ lam, const T* __restrict__ sre,
                                                        const T* __restrict__ sim, const T* __restrict__ tre,
                                                        const T* __restrict__ tim, T* __restrict__ z) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  const T lj = lam[j / nrhs];
  T* zr = z + j;
  T* zi = z + (int64_t)nt * ncol + j;
  for (int s = 0; s < nt; ++s) {
    const int i = transpose ? s : nt - 1 - s;
    T ar = zr[(int64_t)i * ncol], ai = zi[(int64_t)i * ncol];
    // known unknowns: k > i (upper) or k < i (transposed, lower)
    const int k0 = transpose ? 0 : i + 1, k1 = transpose ? i : nt;
    for (int k = k0; k < k1; ++k) {
      const int64_t e = transpose ? (int64_t)k * nt + i : (int64_t)i * nt + k;  // m_ik or (m^T)_ik = m_ki
      const T mr = sre[e] + lj * tre[e], mi = sim[e] + lj * tim[e];
      const T yr = zr[(int64_t)k * ncol], yi = zi[(int64_t)k * ncol];
      ar -= mr * yr - mi * yi;
      ai -= mr * yi + mi * yr;
    }
    const int64_t d = (int64_t)i * nt + i;
    const T dr = sre[d] + lj * tre[d], di = sim[d] + lj * tim[d];
    const T den = T(1) / (dr * dr + di * di);
    zr[(int64_t)i * ncol] = (ar * dr + ai * di) * den;
    zi[(int64_t)i * ncol] = (ai * dr - ar * di) * den;
  }
}

static void* st_upload(iga_stfd_s* f, const std::vector<double>& h) {
  void* d = nullptr;
  const size_t es = f->dtype == IGA_F64 ? 8 : 4;
  IGA_HIP(hipMalloc(&d, es * std::max<size_t>(h.size(), 1)));
  f->owned.push_back(d);
  if (h.empty()) return d;
  if (f->dtype == IGA_F64) {
    IGA_HIP(hipMemcpy(d, h.data(), 8 * h.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h32(h.begin(), h.end());
    IGA_HIP(hipMemcpy(d, h32.data(), 4 * h.size(), hipMemcpyHostToDevice));
  }
  return d;
}

static void st_write(iga_stfd_s* f, void* d, const std::vector<double>& h) {
  if (f->dtype == IGA_F64) {
    IGA_HIP(hipMemcpy(d, h.data(), 8 * h.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h32(h.begin(), h.end());
    IGA_HIP(hipMemcpy(d, h32.data(), 4 * h.size(), hipMemcpyHostToDevice));
  }
}

template <typename T>
static void stfd_apply_t(const iga_stfd_s* f, const T* x, T* y, int nrhs, int tr, T* w0, T* w1, hipStream_t s) {
  const int dt = std::is_same<T, double>::value ? IGA_F64 : IGA_F32;
  int rc = iga_box_gather(dt, 4, f->n, f->lo, f->len, nrhs, x, w0, s);
  if (rc != IGA_OK) throw Error(rc, "box gather failed");
  const int64_t L = f->len[3];
  const int64_t cols = f->nspace * nrhs;
  T* cur = w0;
  T* nxt = w1;
  auto space_pass = [&](int d, bool transpose) {
    int64_t outer = L, inner = nrhs;
    for (int e = d + 1; e < 3; ++e) outer *= f->len[e];
    for (int e = 0; e < d; ++e) inner *= f->len[e];
    const T* A = (const T*)(transpose ? f->d_Ut[d] : f->d_U[d]);
    mode_gemm<T>(A, cur, nxt, outer, f->len[d], inner, nullptr, s);
    std::swap(cur, nxt);
  };
  for (int d = 0; d < 3; ++d) space_pass(d, true);
  mode_gemm_rect<T>((const T*)f->d_F[tr], cur, nxt, 1, 2 * L, L, cols, nullptr, s);
  std::swap(cur, nxt);
  if (f->qz)
    qz_solve_kernel<T><<<grid_for(cols, 256), 256, 0, s>>>((int)L, cols, nrhs, tr, (const T*)f->d_lam,
                                                           (const T*)f->d_S[0], (const T*)f->d_S[1],
                                                           (const T*)f->d_T[0], (const T*)f->d_T[1], cur);
  else
    arrowhead_kernel<T><<<grid_for(cols, 256), 256, 0, s>>>((int)L, cols, nrhs, (const T*)f->d_lam,
                                                            (const T*)f->d_D[0], (const T*)f->d_D[1],
                                                            (const T*)f->d_b[tr][0], (const T*)f->d_b[tr][1], cur);
  IGA_LAUNCH_CHECK();
  mode_gemm_rect<T>((const T*)f->d_G[tr], cur, nxt, 1, L, 2 * L, cols, nullptr, s);
  std::swap(cur, nxt);
  for (int d = 2; d >= 0; --d) space_pass(d, false);
  rc = iga_box_scatter(dt, 4, f->n, f->lo, f->len, nrhs, cur, y, 1, s);
  if (rc != IGA_OK) throw Error(rc, "box scatter failed");
}

static void free_stfd(iga_stfd_s* f) {
  for (int d = 0; d < 3; ++d) {
    (void)hipFree(f->d_U[d]);
    (void)hipFree(f->d_Ut[d]);
  }
  for (void* p : f->owned) (void)hipFree(p);
  f->owned.clear();
}

}  // namespace iga

extern "C" {

int iga_stfd_create(iga_stfd_t* out, const int64_t* n, const int64_t* lo, const int64_t* len, const double* const* U,
                    const double* ut_re, const double* ut_im, int dtype) {
  return guard([&] {
    IGA_CHECK_ARG(out && n && lo && len && U && ut_re && ut_im, "NULL argument");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    for (int d = 0; d < 4; ++d)
      IGA_CHECK_ARG(len[d] >= 1 && lo[d] >= 0 && lo[d] + len[d] <= n[d], "free box does not fit");
    IGA_CHECK_ARG(len[3] >= 2 && len[3] <= kArrowMax, "time block size must be in 2..512");
    auto* f = new iga_stfd_s();
    f->dtype = dtype;
    try {
      f->nspace = 1;
      for (int d = 0; d < 4; ++d) {
        f->n[d] = n[d];
        f->lo[d] = lo[d];
        f->len[d] = len[d];
        if (d < 3) f->nspace *= len[d];
      }
      f->nsub = f->nspace * len[3];
      const size_t es = dtype == IGA_F64 ? 8 : 4;
      for (int d = 0; d < 3; ++d) {
        const int64_t m = len[d];
        std::vector<double> u((size_t)m * m), ut((size_t)m * m);
        for (int64_t i = 0; i < m; ++i)
          for (int64_t k = 0; k < m; ++k) {
            u[i * m + k] = U[d][i * m + k];
            ut[i * m + k] = U[d][k * m + i];
          }
        IGA_HIP(hipMalloc(&f->d_U[d], es * m * m));
        IGA_HIP(hipMalloc(&f->d_Ut[d], es * m * m));
        st_write(f, f->d_U[d], u);
        st_write(f, f->d_Ut[d], ut);
      }
      const int64_t L = len[3];
      for (int tr = 0; tr < 2; ++tr) {
        // apply: U_t^H forward, U_t backward; transpose: U_t^T forward, conj(U_t) backward
        const double sg = tr == 0 ? -1.0 : 1.0;
        std::vector<double> F((size_t)2 * L * L), G((size_t)L * 2 * L);
        for (int64_t i = 0; i < L; ++i)
          for (int64_t k = 0; k < L; ++k) {
            F[i * L + k] = ut_re[k * L + i];
            F[(L + i) * L + k] = sg * ut_im[k * L + i];
            G[i * 2 * L + k] = ut_re[i * L + k];
            G[i * 2 * L + L + k] = sg * ut_im[i * L + k];
          }
        f->d_F[tr] = st_upload(f, F);
        f->d_G[tr] = st_upload(f, G);
      }
      f->d_lam = st_upload(f, std::vector<double>((size_t)f->nspace, 0.0));
      f->d_D[0] = st_upload(f, std::vector<double>((size_t)L, 0.0));
      f->d_D[1] = st_upload(f, std::vector<double>((size_t)L, 0.0));
      for (int tr = 0; tr < 2; ++tr)
        for (int c = 0; c < 2; ++c) f->d_b[tr][c] = st_upload(f, std::vector<double>((size_t)L, 0.0));
    } catch (...) {
      free_stfd(f);
      delete f;
      throw;
    }
    *out = f;
  });
}

int iga_stfd_create_qz(iga_stfd_t* out, const int64_t* n, const int64_t* lo, const int64_t* len, const double* const* U,
                       const double* q_re, const double* q_im, const double* z_re, const double* z_im, int dtype) {
  return guard([&] {
    IGA_CHECK_ARG(out && q_re && q_im && z_re && z_im, "NULL argument");
    // the space factors and the buffers of the arrowhead plan; Q stands in for U_t
    IGA_CHECK(iga_stfd_create(out, n, lo, len, U, q_re, q_im, dtype));
    iga_stfd_s* f = *out;
    try {
      const int64_t L = len[3];
      for (int tr = 0; tr < 2; ++tr) {
        // apply: Q^H forward, Z backward (time_legacy.py:81-105); transpose: Z^T
        // forward, conj(Q) backward (time_legacy.py:140-160).  F = [Re A; Im A],
        // G = [Re B, -Im B] for the forward A and backward B time matrices.
        std::vector<double> F((size_t)2 * L * L), G((size_t)L * 2 * L);
        for (int64_t i = 0; i < L; ++i)
          for (int64_t k = 0; k < L; ++k) {
            if (tr == 0) {
              F[i * L + k] = q_re[k * L + i];
              F[(L + i) * L + k] = -q_im[k * L + i];
              G[i * 2 * L + k] = z_re[i * L + k];
              G[i * 2 * L + L + k] = -z_im[i * L + k];
            } else {
              F[i * L + k] = z_re[k * L + i];
              F[(L + i) * L + k] = z_im[k * L + i];
              G[i * 2 * L + k] = q_re[i * L + k];
              G[i * 2 * L + L + k] = q_im[i * L + k];
            }
          }
        st_write(f, f->d_F[tr], F);
        st_write(f, f->d_G[tr], G);
      }
      for (int c = 0; c < 2; ++c) {
        f->d_S[c] = st_upload(f, std::vector<double>((size_t)L * L, 0.0));
        f->d_T[c] = st_upload(f, std::vector<double>((size_t)L * L, 0.0));
      }
      f->qz = 1;
    } catch (...) {
      free_stfd(f);
      delete f;
      *out = nullptr;
      throw;
    }
  });
}

int iga_stfd_set_qz(iga_stfd_t f, const double* lam_space, const double* s_re, const double* s_im,
                    const double* t_re, const double* t_im, double adv) {
  return guard([&] {
    IGA_CHECK_ARG(f && f->qz && lam_space && s_re && s_im && t_re && t_im, "NULL argument or not a QZ plan");
    const int64_t L = f->len[3];
    st_write(f, f->d_lam, std::vector<double>(lam_space, lam_space + f->nspace));
    std::vector<double> sr((size_t)L * L), si((size_t)L * L);
    for (int64_t e = 0; e < L * L; ++e) {
      sr[e] = adv * s_re[e];
      si[e] = adv * s_im[e];
    }
    st_write(f, f->d_S[0], sr);
    st_write(f, f->d_S[1], si);
    st_write(f, f->d_T[0], std::vector<double>(t_re, t_re + L * L));
    st_write(f, f->d_T[1], std::vector<double>(t_im, t_im + L * L));
    f->blocks_set = true;
  });
}

int iga_stfd_set_blocks(iga_stfd_t f, const double* lam_space, const double* d_re, const double* d_im,
                        const double* g_re, const double* g_im, double adv) {
  return guard([&] {
    IGA_CHECK_ARG(f && lam_space && d_re && d_im && g_re && g_im, "NULL argument");
    IGA_CHECK_ARG(!f->qz, "a QZ plan takes iga_stfd_set_qz");
    const int64_t L = f->len[3];
    st_write(f, f->d_lam, std::vector<double>(lam_space, lam_space + f->nspace));
    std::vector<double> dr(L), di(L), b[2][2];
    for (int64_t i = 0; i < L; ++i) {
      dr[i] = adv * d_re[i];
      di[i] = adv * d_im[i];
    }
    for (int tr = 0; tr < 2; ++tr)
      for (int c = 0; c < 2; ++c) b[tr][c].assign((size_t)L, 0.0);
    for (int64_t i = 0; i + 1 < L; ++i) {
      b[0][0][i] = adv * g_re[i];   // apply: b = adv g
      b[0][1][i] = adv * g_im[i];
      b[1][0][i] = -adv * g_re[i];  // transpose: b = -adv conj(g)
      b[1][1][i] = adv * g_im[i];
    }
    st_write(f, f->d_D[0], dr);
    st_write(f, f->d_D[1], di);
    for (int tr = 0; tr < 2; ++tr)
      for (int c = 0; c < 2; ++c) st_write(f, f->d_b[tr][c], b[tr][c]);
    f->blocks_set = true;
  });
}

int64_t iga_stfd_workspace_bytes(iga_stfd_t f, int nrhs) {
  if (!f || nrhs < 1) return 0;
  return 2 * (2 * f->nsub * (int64_t)nrhs) * (f->dtype == IGA_F64 ? 8 : 4) + 512;
}

int iga_stfd_apply(iga_stfd_t f, const void* x, void* y, int nrhs, int transpose, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(f && x && y && work && nrhs >= 1, "NULL argument");
    IGA_CHECK_ARG(f->blocks_set, "time blocks and eigenvalues are not set");
    const int64_t half = 2 * f->nsub * nrhs;
    const int tr = transpose ? 1 : 0;
    if (f->dtype == IGA_F64) {
      double* w0 = (double*)work;
      stfd_apply_t<double>(f, (const double*)x, (double*)y, nrhs, tr, w0, w0 + half, as_stream(stream));
    } else {
      float* w0 = (float*)work;
      stfd_apply_t<float>(f, (const float*)x, (float*)y, nrhs, tr, w0, w0 + half, as_stream(stream));
    }
  });
}

int iga_stfd_destroy(iga_stfd_t f) {
  return guard([&] {
    if (!f) return;
    free_stfd(f);
    delete f;
  });
}

}  // extern "C"
