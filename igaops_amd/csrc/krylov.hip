// Krylov vector kernels: many-dot products (Gram-Schmidt projections, norms),
// many-axpy updates, axpby and precision casts.  Reductions are deterministic:
// fixed per-block chunks, a 64-lane wavefront shuffle tree, then a fixed-order
// second pass over the per-block partials.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace iga {

constexpr int kRedThreads = 256;
constexpr int kRedPerThread = 8;
constexpr int kRedChunk = kRedThreads * kRedPerThread;  // elements per block
constexpr int kRedMaxBlocks = 2048;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum; the result is valid in thread 0.
__device__ inline double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
  return s;
}

static int reduce_blocks(int64_t n) {
  int64_t g = (n + kRedChunk - 1) / kRedChunk;
  if (g < 1) g = 1;
  if (g > kRedMaxBlocks) g = kRedMaxBlocks;
  return (int)g;
}

// Many-dot products V w: blocks tile (element range) x (group of kDotVec vectors);
// every thread keeps one partial sum per vector of its group over its elements, the
// 64 lanes of a wave are combined by shuffles, and each wave writes its partials
// (deterministic: fixed ranges, fixed combination order).
constexpr int kDotVec = 16;
constexpr int kDotChunk = 1024;  // elements per block when n is small enough
constexpr int kDotMaxChunks = 4096;

static int dot_chunks(int64_t n) {
  int64_t g = (n + kDotChunk - 1) / kDotChunk;
  if (g < 1) g = 1;
  if (g > kDotMaxChunks) g = kDotMaxChunks;
  return (int)g;
}

// partial[(b * 4 + wave) * k + j] for chunk b
template <typename T>
__global__ void __launch_bounds__(kRedThreads) dot_many_stage1(int64_t n, int k, const T* __restrict__ V,
                                                               int64_t ldv, const T* __restrict__ w,
                                                               double* __restrict__ partial) {
  const int nb = gridDim.x;
  const int64_t per = ((n + nb - 1) / nb + kRedThreads - 1) / kRedThreads * kRedThreads;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  const int j0 = blockIdx.y * kDotVec;
  const int kb = k - j0 < kDotVec ? k - j0 : kDotVec;
  double acc[kDotVec];
#pragma unroll
  for (int jj = 0; jj < kDotVec; ++jj) acc[jj] = 0.0;
  const T* Vb = V + (int64_t)j0 * ldv;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kRedThreads) {
    const double wv = (double)w[i];
#pragma unroll
    for (int jj = 0; jj < kDotVec; ++jj)
      if (jj < kb) acc[jj] += (double)Vb[(int64_t)jj * ldv + i] * wv;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* out = partial + ((int64_t)blockIdx.x * (kRedThreads / 64) + wave) * k + j0;
#pragma unroll
  for (int jj = 0; jj < kDotVec; ++jj) {
    if (jj < kb) {
      const double v = wave_sum(acc[jj]);
      if (lane == 0) out[jj] = v;
    }
  }
}

__global__ void __launch_bounds__(kRedThreads) dot_many_stage2(int nb, int k, const double* __restrict__ partial,
                                                               double* __restrict__ out) {
  __shared__ double lds[kRedThreads / 64];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[(int64_t)b * k + j];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) out[j] = s;
}

template <typename T>
__global__ void axpy_many_kernel(int64_t n, int k, const T* __restrict__ V, int64_t ldv,
                                 const double* __restrict__ h, double alpha, T* __restrict__ w) {
  extern __shared__ double hs[];
  for (int j = threadIdx.x; j < k; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += hs[j] * (double)V[(int64_t)j * ldv + i];
    w[i] = (T)((double)w[i] + alpha * acc);
  }
}

template <typename T>
__global__ void axpby_kernel(int64_t n, double a, const double* a_dev, int invert_a, const T* __restrict__ x,
                             double b, T* __restrict__ y) {
  double av = a_dev ? a_dev[0] : a;
  if (invert_a) av = 1.0 / av;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double yv = (b == 0.0) ? 0.0 : b * (double)y[i];
    y[i] = (T)(av * (double)x[i] + yv);
  }
}

// Complex (interleaved re, im) variants: partial[b][2j..2j+1] = sum conj(V_j[i]) w[i]
__global__ void __launch_bounds__(kRedThreads) zdot_many_stage1(int64_t n, int k, const double* __restrict__ V,
                                                                int64_t ldv, const double* __restrict__ w,
                                                                double* __restrict__ partial) {
  __shared__ double lds[kRedThreads / 64];
  const int nb = gridDim.x;
  const int64_t per = (n + nb - 1) / nb;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  for (int j = 0; j < k; ++j) {
    const double* v = V + 2 * (int64_t)j * ldv;
    double re = 0.0, im = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const double vr = v[2 * i], vi = v[2 * i + 1], wr = w[2 * i], wi = w[2 * i + 1];
      re += vr * wr + vi * wi;
      im += vr * wi - vi * wr;
    }
    const double sr = block_sum(re, lds);
    const double si = block_sum(im, lds);
    if (threadIdx.x == 0) {
      partial[((int64_t)blockIdx.x * k + j) * 2] = sr;
      partial[((int64_t)blockIdx.x * k + j) * 2 + 1] = si;
    }
  }
}

__global__ void __launch_bounds__(kRedThreads) zdot_many_stage2(int nb, int k, const double* __restrict__ partial,
                                                                double* __restrict__ out) {
  __shared__ double lds[kRedThreads / 64];
  const int j = blockIdx.x >> 1, part = blockIdx.x & 1;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[((int64_t)b * k + j) * 2 + part];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) out[2 * j + part] = s;
}

// w[i] += alpha * sum_j h[j] V_j[i]   (complex h, complex alpha)
__global__ void zaxpy_many_kernel(int64_t n, int k, const double* __restrict__ V, int64_t ldv,
                                  const double* __restrict__ h, double ar, double ai, double* __restrict__ w) {
  extern __shared__ double hs[];
  for (int j = threadIdx.x; j < 2 * k; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < k; ++j) {
      const double vr = V[2 * ((int64_t)j * ldv + i)], vi = V[2 * ((int64_t)j * ldv + i) + 1];
      const double hr = hs[2 * j], hi = hs[2 * j + 1];
      re += hr * vr - hi * vi;
      im += hr * vi + hi * vr;
    }
    w[2 * i] += ar * re - ai * im;
    w[2 * i + 1] += ar * im + ai * re;
  }
}

template <typename S, typename D>
__global__ void cast_kernel(int64_t n, const S* __restrict__ x, D* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (D)x[i];
}

// ---- device-side GMRES step -------------------------------------------------
// Per-solve state (float64) for Krylov dimension m, at fixed offsets:
//   [0, m)       first Gram-Schmidt pass, [m, 2m) second pass,
//   2m           ||w||^2, 2m+1 the scale of the new basis vector (1/||w||, or 1/beta),
//   R            m x m, column-major, the rotated Hessenberg columns,
//   cs, sn       the Givens rotations, g (m+1) the rotated right-hand side beta e_1,
//   hist         2 (m+1): (g[k], ||w_k||) per step, the values a host polls.
constexpr int kGmMaxDim = 2048;  // the rotation kernel stages 3 k doubles in LDS

struct GmState {
  int m;
  __host__ __device__ int64_t R() const { return 2 * (int64_t)m + 2; }
  __host__ __device__ int64_t cs() const { return R() + (int64_t)m * m; }
  __host__ __device__ int64_t sn() const { return cs() + m; }
  __host__ __device__ int64_t g() const { return sn() + m; }
  __host__ __device__ int64_t hist() const { return g() + m + 1; }
  __host__ __device__ int64_t size() const { return hist() + 2 * ((int64_t)m + 1); }
};

__global__ void gm_zero_kernel(int64_t n, double* __restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// beta = ||r|| from ls[2m]; g[0] = beta, scale = 1/beta (0 for r = 0)
__global__ void gm_start_kernel(GmState st, double* ls) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double beta = sqrt(fmax(ls[2 * st.m], 0.0));
  ls[st.g()] = beta;
  ls[2 * st.m + 1] = beta > 0.0 ? 1.0 / beta : 0.0;
  ls[st.hist()] = beta;
  ls[st.hist() + 1] = beta;
}

// Step k (1-based): Hessenberg column h = pass1 + pass2, h[k] = ||w||, rotated by the
// k-1 previous rotations; a new rotation annihilates h[k] and updates g.  The column
// and the rotations are staged through LDS by one wave, the sequential rotation
// sweep runs on one lane.
__global__ void __launch_bounds__(64) gm_rotate_kernel(GmState st, int k, double* ls) {
  extern __shared__ double sh[];  // h[k], c[k], s[k]
  double* h = sh;
  double* c = sh + k;
  double* sv = sh + 2 * k;
  const int m = st.m;
  double* Rk = ls + st.R() + (int64_t)(k - 1) * m;
  double* cs = ls + st.cs();
  double* sn = ls + st.sn();
  double* g = ls + st.g();
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    h[i] = ls[i] + ls[m + i];
    c[i] = cs[i];
    sv[i] = sn[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nw = sqrt(fmax(ls[2 * m], 0.0));
    for (int i = 0; i + 1 < k; ++i) {
      const double a = h[i], b = h[i + 1];
      h[i] = c[i] * a + sv[i] * b;
      h[i + 1] = -sv[i] * a + c[i] * b;
    }
    const double a = h[k - 1];
    double cc = 1.0, ss = 0.0;
    if (nw != 0.0) {
      const double r = hypot(a, nw);
      cc = a / r;
      ss = nw / r;
    }
    cs[k - 1] = cc;
    sn[k - 1] = ss;
    h[k - 1] = cc * a + ss * nw;
    const double gk1 = g[k - 1];
    g[k] = -ss * gk1;
    g[k - 1] = cc * gk1;
    ls[2 * m + 1] = nw > 0.0 ? 1.0 / nw : 0.0;
    ls[st.hist() + 2 * k] = -ss * gk1;
    ls[st.hist() + 2 * k + 1] = nw;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) Rk[i] = h[i];
}

}  // namespace iga

using namespace iga;

extern "C" {

int64_t iga_reduce_workspace_bytes(int64_t n, int k) {
  if (n < 0 || k < 0) return 0;
  const int64_t kk = k > 0 ? k : 1;
  const int64_t dots = (int64_t)dot_chunks(n) * (kRedThreads / 64) * kk;  // iga_dot_many
  const int64_t zdots = (int64_t)reduce_blocks(n) * 2 * kk;              // iga_zdot_many
  return (dots > zdots ? dots : zdots) * (int64_t)sizeof(double);
}

int iga_dot_many(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* w, double* out,
                 void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && out && work && n >= 0 && k >= 0 && ldv >= n, "bad argument");
    if (k == 0) return;
    hipStream_t s = as_stream(stream);
    const int nb = dot_chunks(n);
    double* partial = (double*)work;
    const dim3 grid(nb, (k + kDotVec - 1) / kDotVec);
    if (dtype == IGA_F64)
      dot_many_stage1<double><<<grid, kRedThreads, 0, s>>>(n, k, (const double*)V, ldv, (const double*)w, partial);
    else if (dtype == IGA_F32)
      dot_many_stage1<float><<<grid, kRedThreads, 0, s>>>(n, k, (const float*)V, ldv, (const float*)w, partial);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
    dot_many_stage2<<<k, kRedThreads, 0, s>>>(nb * (kRedThreads / 64), k, partial, out);
    IGA_LAUNCH_CHECK();
  });
}

int iga_axpy_many(int dtype, int64_t n, int k, const void* V, int64_t ldv, const double* h, double alpha, void* w,
                  void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && h && n >= 0 && k >= 0 && ldv >= n, "bad argument");
    IGA_CHECK_ARG(k <= 8192, "too many vectors");
    if (k == 0 || n == 0) return;
    hipStream_t s = as_stream(stream);
    dim3 grid(grid_for(n, 256, 8192)), block(256);
    const size_t shm = sizeof(double) * (size_t)k;
    if (dtype == IGA_F64)
      axpy_many_kernel<double><<<grid, block, shm, s>>>(n, k, (const double*)V, ldv, h, alpha, (double*)w);
    else if (dtype == IGA_F32)
      axpy_many_kernel<float><<<grid, block, shm, s>>>(n, k, (const float*)V, ldv, h, alpha, (float*)w);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
  });
}

int iga_zdot_many(int64_t n, int k, const double* V, int64_t ldv, const double* w, double* out, void* work,
                  void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && out && work && n >= 0 && k >= 0 && ldv >= n, "bad argument");
    if (k == 0) return;
    hipStream_t s = as_stream(stream);
    const int nb = reduce_blocks(n);
    zdot_many_stage1<<<nb, kRedThreads, 0, s>>>(n, k, V, ldv, w, (double*)work);
    IGA_LAUNCH_CHECK();
    zdot_many_stage2<<<2 * k, kRedThreads, 0, s>>>(nb, k, (const double*)work, out);
    IGA_LAUNCH_CHECK();
  });
}

int iga_zaxpy_many(int64_t n, int k, const double* V, int64_t ldv, const double* h, double alpha_re,
                   double alpha_im, double* w, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && h && n >= 0 && k >= 0 && ldv >= n, "bad argument");
    IGA_CHECK_ARG(k <= 4096, "too many vectors");
    if (k == 0 || n == 0) return;
    dim3 grid(grid_for(n, 256, 8192)), block(256);
    zaxpy_many_kernel<<<grid, block, sizeof(double) * 2 * (size_t)k, as_stream(stream)>>>(n, k, V, ldv, h, alpha_re,
                                                                                       alpha_im, w);
    IGA_LAUNCH_CHECK();
  });
}

int iga_axpby(int dtype, int64_t n, double a, const double* a_dev, int invert_a, const void* x, double b, void* y,
              void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(x && y && n >= 0, "bad argument");
    if (n == 0) return;
    dim3 grid(grid_for(n, 256, 8192)), block(256);
    hipStream_t s = as_stream(stream);
    if (dtype == IGA_F64)
      axpby_kernel<double><<<grid, block, 0, s>>>(n, a, a_dev, invert_a, (const double*)x, b, (double*)y);
    else if (dtype == IGA_F32)
      axpby_kernel<float><<<grid, block, 0, s>>>(n, a, a_dev, invert_a, (const float*)x, b, (float*)y);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
  });
}

int iga_cast(int dtype_x, int dtype_y, int64_t n, const void* x, void* y, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(x && y && n >= 0, "bad argument");
    if (n == 0) return;
    dim3 grid(grid_for(n, 256, 8192)), block(256);
    hipStream_t s = as_stream(stream);
    if (dtype_x == IGA_F64 && dtype_y == IGA_F32)
      cast_kernel<double, float><<<grid, block, 0, s>>>(n, (const double*)x, (float*)y);
    else if (dtype_x == IGA_F32 && dtype_y == IGA_F64)
      cast_kernel<float, double><<<grid, block, 0, s>>>(n, (const float*)x, (double*)y);
    else if (dtype_x == IGA_F64 && dtype_y == IGA_F64)
      cast_kernel<double, double><<<grid, block, 0, s>>>(n, (const double*)x, (double*)y);
    else if (dtype_x == IGA_F32 && dtype_y == IGA_F32)
      cast_kernel<float, float><<<grid, block, 0, s>>>(n, (const float*)x, (float*)y);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
  });
}

int64_t iga_gmres_state_doubles(int m) {
  if (m < 1) return 0;
  return GmState{m}.size();
}

int iga_gmres_start(int dtype, int64_t n, int m, const void* r, void* V0, double* ls, void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(r && V0 && ls && work && n >= 1 && m >= 1, "bad argument");
    IGA_CHECK_ARG(m <= kGmMaxDim, "Krylov dimension above 2048");
    hipStream_t s = as_stream(stream);
    const GmState st{m};
    gm_zero_kernel<<<grid_for(st.size(), 256, 1024), 256, 0, s>>>(st.size(), ls);
    IGA_LAUNCH_CHECK();
    IGA_CHECK(iga_dot_many(dtype, n, 1, r, n, r, ls + 2 * m, work, stream));
    gm_start_kernel<<<1, 64, 0, s>>>(st, ls);
    IGA_LAUNCH_CHECK();
    IGA_CHECK(iga_axpby(dtype, n, 0.0, ls + 2 * m + 1, 0, r, 0.0, V0, stream));
  });
}

int iga_gmres_step(int dtype, int64_t n, int k, int m, void* V, int64_t ldv, void* w, double* ls, void* work,
                   void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && ls && work && n >= 1 && ldv >= n, "bad argument");
    IGA_CHECK_ARG(k >= 1 && k <= m, "step outside the Krylov dimension");
    IGA_CHECK_ARG(m <= kGmMaxDim, "Krylov dimension above 2048");
    hipStream_t s = as_stream(stream);
    const GmState st{m};
    // classical Gram-Schmidt, two passes
    IGA_CHECK(iga_dot_many(dtype, n, k, V, ldv, w, ls, work, stream));
    IGA_CHECK(iga_axpy_many(dtype, n, k, V, ldv, ls, -1.0, w, stream));
    IGA_CHECK(iga_dot_many(dtype, n, k, V, ldv, w, ls + m, work, stream));
    IGA_CHECK(iga_axpy_many(dtype, n, k, V, ldv, ls + m, -1.0, w, stream));
    IGA_CHECK(iga_dot_many(dtype, n, 1, w, n, w, ls + 2 * m, work, stream));
    gm_rotate_kernel<<<1, 64, 3 * sizeof(double) * (size_t)k, s>>>(st, k, ls);
    IGA_LAUNCH_CHECK();
    const size_t es = dtype == IGA_F64 ? 8 : 4;
    void* vk = (char*)V + (size_t)k * ldv * es;
    IGA_CHECK(iga_axpby(dtype, n, 0.0, ls + 2 * m + 1, 0, w, 0.0, vk, stream));
  });
}

}  // extern "C"
