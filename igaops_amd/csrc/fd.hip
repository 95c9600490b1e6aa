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
  bool blocks_set = false;
  std::vector<void*> owned;
};

namespace iga {

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

int iga_stfd_set_blocks(iga_stfd_t f, const double* lam_space, const double* d_re, const double* d_im,
                        const double* g_re, const double* g_im, double adv) {
  return guard([&] {
    IGA_CHECK_ARG(f && lam_space && d_re && d_im && g_re && g_im, "NULL argument");
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
