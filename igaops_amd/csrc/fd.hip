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

// C = L R (+ epilogue scale).  Lane layout of the 16x16x4 MFMA (measured on gfx950):
// A operand: lane l -> A[l % 16][l / 16]; B operand: lane l -> B[l / 16][l % 16];
// D register r of lane l -> D[Acc<T>::row(l, r)][l % 16].
template <typename T>
__global__ void __launch_bounds__(256) mfma_gemm_kernel(GemmDesc g, const T* __restrict__ L, const T* __restrict__ R,
                                                        T* __restrict__ Cout, const T* __restrict__ scale) {
  using V = typename Acc<T>::V;
  __shared__ T ls[kTile][kKStep + 1];
  __shared__ T rs[kKStep][kTile + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kTile, j0 = blockIdx.y * kTile;
  const int64_t bz = blockIdx.z;
  const T* Lb = L + bz * g.l_batch;
  const T* Rb = R + bz * g.r_batch;
  T* Cb = Cout + bz * g.c_batch;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{0, 0, 0, 0};

  for (int k0 = 0; k0 < g.K; k0 += kKStep) {
    __syncthreads();
    // L tile: 64 rows x 16 k  (k fastest among the threads when lk == 1)
    for (int e = tid; e < kTile * kKStep; e += 256) {
      const int mm = e / kKStep, kk = e % kKStep;
      const int m = m0 + mm, k = k0 + kk;
      ls[mm][kk] = (m < g.M && k < g.K) ? Lb[m * g.lm + k * g.lk] : T(0);
    }
    // R tile: 16 k x 64 columns (columns fastest among the threads when rj == 1)
    for (int e = tid; e < kTile * kKStep; e += 256) {
      int kk, jj;
      if (g.rj == 1) {
        kk = e / kTile;
        jj = e % kTile;
      } else {
        jj = e / kKStep;
        kk = e % kKStep;
      }
      const int k = k0 + kk, j = j0 + jj;
      rs[kk][jj] = (k < g.K && j < g.N) ? Rb[k * g.rk + j * g.rj] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < kKStep; k4 += 4) {
      T av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = ls[wm + a * 16 + (lane & 15)][k4 + (lane >> 4)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = rs[k4 + (lane >> 4)][wn + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc<T>::mfma(av[a], bv[b], acc[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
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
  int64_t nsub = 0;
};

namespace iga {

template <typename T>
static void mode_gemm(const T* A, const T* X, T* Y, int64_t outer, int64_t n, int64_t inner, const T* scale,
                      hipStream_t s) {
  GemmDesc g{};
  dim3 grid;
  if (inner > 1) {
    g.M = (int)n;
    g.N = (int)inner;
    g.K = (int)n;
    g.lm = n; g.lk = 1;            // L = A
    g.rk = inner; g.rj = 1;        // R = X(o)
    g.cm = inner; g.cj = 1;
    g.l_batch = 0; g.r_batch = n * inner; g.c_batch = n * inner;
    g.scale_m = inner; g.scale_j = 1; g.scale_batch = n * inner;
    IGA_CHECK_ARG(outer < 65536 && inner < (int64_t)1 << 30, "mode product too large for the grid");
    grid = dim3((unsigned)((n + kTile - 1) / kTile), (unsigned)((inner + kTile - 1) / kTile), (unsigned)outer);
  } else {
    g.M = (int)outer;
    g.N = (int)n;
    g.K = (int)n;
    g.lm = n; g.lk = 1;            // L = X rows
    g.rk = 1; g.rj = n;            // R(k, i) = A[i][k]
    g.cm = n; g.cj = 1;
    g.scale_m = n; g.scale_j = 1;
    IGA_CHECK_ARG(outer < (int64_t)1 << 30, "mode product too large for the grid");
    grid = dim3((unsigned)((outer + kTile - 1) / kTile), (unsigned)((n + kTile - 1) / kTile), 1);
  }
  const T* Lp = inner > 1 ? A : X;
  const T* Rp = inner > 1 ? X : A;
  mfma_gemm_kernel<T><<<grid, 256, 0, s>>>(g, Lp, Rp, Y, scale);
  IGA_LAUNCH_CHECK();
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
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    auto* f = new iga_fd_s();
    f->dtype = dtype;
    f->ndim = ndim;
    f->nsub = 1;
    try {
      for (int d = 0; d < ndim; ++d) {
        IGA_CHECK_ARG(len[d] >= 1 && lo[d] >= 0 && lo[d] + len[d] <= n[d], "free box does not fit");
        f->n[d] = n[d];
        f->lo[d] = lo[d];
        f->len[d] = len[d];
        f->nsub *= len[d];
        const int64_t m = len[d];
        std::vector<double> ut((size_t)m * m), u((size_t)m * m);
        for (int64_t i = 0; i < m; ++i)
          for (int64_t k = 0; k < m; ++k) {
            u[i * m + k] = U[d][i * m + k];   // U[i][k]  (row-major, as numpy)
            ut[i * m + k] = U[d][k * m + i];  // U^T[i][k]
          }
        const size_t es = dtype == IGA_F64 ? 8 : 4;
        IGA_HIP(hipMalloc(&f->d_U[d], es * m * m));
        IGA_HIP(hipMalloc(&f->d_Ut[d], es * m * m));
        if (dtype == IGA_F64) {
          IGA_HIP(hipMemcpy(f->d_U[d], u.data(), 8 * m * m, hipMemcpyHostToDevice));
          IGA_HIP(hipMemcpy(f->d_Ut[d], ut.data(), 8 * m * m, hipMemcpyHostToDevice));
        } else {
          std::vector<float> u32(u.begin(), u.end()), ut32(ut.begin(), ut.end());
          IGA_HIP(hipMemcpy(f->d_U[d], u32.data(), 4 * m * m, hipMemcpyHostToDevice));
          IGA_HIP(hipMemcpy(f->d_Ut[d], ut32.data(), 4 * m * m, hipMemcpyHostToDevice));
        }
      }
    } catch (...) {
      for (int d = 0; d < 4; ++d) {
        (void)hipFree(f->d_U[d]);
        (void)hipFree(f->d_Ut[d]);
      }
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
    if (f->dtype == IGA_F64) {
      double* w0 = (double*)work;
      fd_apply_t<double>(f, (const double*)x, (double*)y, nrhs, w0, w0 + half, as_stream(stream));
    } else {
      float* w0 = (float*)work;
      fd_apply_t<float>(f, (const float*)x, (float*)y, nrhs, w0, w0 + half, as_stream(stream));
    }
  });
}

int iga_fd_destroy(iga_fd_t f) {
  return guard([&] {
    if (!f) return;
    for (int d = 0; d < 4; ++d) {
      IGA_HIP(hipFree(f->d_U[d]));
      IGA_HIP(hipFree(f->d_Ut[d]));
    }
    IGA_HIP(hipFree(f->d_inv));
    delete f;
  });
}

}  // extern "C"
