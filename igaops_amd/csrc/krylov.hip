// Krylov vector kernels: many-dot products (Gram-Schmidt projections, norms),
// many-axpy updates, axpby and precision casts.  Reductions are deterministic:
// fixed per-block chunks, a 64-lane wavefront shuffle tree, then a fixed-order
// second pass over the per-block partials.
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.hpp"
#include "internal.hpp"

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

// Many-dot products of the rows of V with one or two vectors x0, x1.  Stage 1:
// blocks tile (element range) x (group of rows); each thread keeps kDotAcc
// partial sums (one per row of its group and vector: 16 rows for one vector, 8
// for two; more accumulators cost occupancy and measured slower), reads 16-byte vectors where
// the layout allows it, the 64 lanes of a wave are combined by a transposed
// butterfly, the waves of a block through LDS, and each block writes one partial
// per (vector, row).  Stage 2 sums the partials of every block in a fixed order
// (one wave per output), so the result does not depend on the scheduling of the
// blocks.
#ifndef IGA_DOT_BLOCKS
#define IGA_DOT_BLOCKS 4096
#endif
#ifndef IGA_DOT_EPT
#define IGA_DOT_EPT 8
#endif
constexpr int kDotAcc = 16;
constexpr int kDotBlocksTarget = IGA_DOT_BLOCKS;
constexpr int kDotMaxVectors = 16384;
// 16-byte vectors per thread and chunk.  GMRES step at n = 314 432: 4 took FP64 at basis
// size 64 from 77 to 68 us against 2 (1 was slower); 8 against 4, at basis sizes 16 / 32 /
// 64: FP64 34.1 -> 31.8, 48.0 -> 44.1, 67.9 -> 67.9 us, FP32 24.9 -> 26.0, 32.2 -> 30.0,
// 46.1 -> 43.8 us (the 2000-request serving stream: FP64 49.1 -> 50.0 solves/s, mixed equal)
constexpr int kDotElemsPerThread = IGA_DOT_EPT;

template <typename T> struct Vec16;
template <> struct Vec16<double> {
  using type = double2;
  static constexpr int width = 2;
  __device__ static double dot(const type& a, const type& b) { return a.x * b.x + a.y * b.y; }
};
template <> struct Vec16<float> {
  using type = float4;
  static constexpr int width = 4;
  __device__ static double dot(const type& a, const type& b) {
    return (double)a.x * (double)b.x + (double)a.y * (double)b.y + (double)a.z * (double)b.z +
           (double)a.w * (double)b.w;
  }
};

// element chunks (blockIdx.x) for `groups` row groups (blockIdx.y)
static int dot_chunks(int64_t n, int groups, int esize) {
  int64_t cap = kDotBlocksTarget / (groups > 0 ? groups : 1);
  if (cap < 64) cap = 64;
  const int64_t chunk = (int64_t)kRedThreads * (16 / esize) * kDotElemsPerThread;
  int64_t g = (n + chunk - 1) / chunk;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Transposed butterfly over the 64 lanes of a wave for N = 16 or 32 values per
// lane: each exchange halves the values a lane holds, so lane q * (64 / N) ends
// with the wave's sum of acc[q] (N + log2(64 / N) - 1 exchanges instead of N
// separate 6-step trees).  Returns the lane's value.
template <int N>
__device__ inline double wave_sum_t(double (&acc)[N], int lane) {
  static_assert(N == 16 || N == 32, "16 or 32 values per lane");
  constexpr int kLevels = N == 16 ? 4 : 5;
#pragma unroll
  for (int st = 0; st < kLevels; ++st) {
    const int off = 32 >> st, c = N >> (st + 1);
    const bool upper = (lane & off) != 0;
#pragma unroll
    for (int q = 0; q < c; ++q) {
      const double give = upper ? acc[q] : acc[q + c];
      const double keep = upper ? acc[q + c] : acc[q];
      acc[q] = keep + __shfl_xor(give, off, 64);
    }
  }
  double v = acc[0];
#pragma unroll
  for (int off = 32 >> kLevels; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// partial[(x * k + j) * nb + b] = sum over chunk b of V_j x_x.
// One-dimensional grid: block L runs chunk b = (L / 8 / groups) * 8 + L % 8 and
// row group (L / 8) % groups, so that the row groups of one chunk run on one XCD
// (work-groups go round-robin to the eight XCDs) one after the other and read
// the chunk of x0, x1 from that XCD's L2.
template <typename T, bool VECTOR, int NX>
__device__ __forceinline__ void dot_many_stage1_body(int L, int64_t n, int k, int nb, int groups,
                                                     const T* __restrict__ V, int64_t ldv, const T* __restrict__ x0,
                                                     const T* __restrict__ x1, double* __restrict__ partial) {
  constexpr int NA = kDotAcc;  // accumulators per thread
  constexpr int G = NA / NX;    // rows per group
  __shared__ double red[kRedThreads / 64][NA];
  const int bx = (L / 8 / groups) * 8 + L % 8;
  if (bx >= nb) return;
  const int by = (L / 8) % groups;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j0 = by * G;
  const int kb = k - j0 < G ? k - j0 : G;
  double acc[NA];  // acc[x * G + jj]
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = 0.0;
  const T* Vb = V + (int64_t)j0 * ldv;
  int64_t body = 0;
  if (VECTOR) {
    using VT = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::width;
    const int64_t nv = n / W;
    body = nv * W;
    const int64_t per = (nv + nb - 1) / nb;
    const int64_t lo = (int64_t)bx * per;
    const int64_t hi = lo + per < nv ? lo + per : nv;
    for (int64_t i = lo + tid; i < hi; i += kRedThreads) {
      const VT a = reinterpret_cast<const VT*>(x0)[i];
      VT b{};
      if (NX == 2) b = reinterpret_cast<const VT*>(x1)[i];
      // every load is issued unconditionally (rows past the group's end re-read its
      // last row, their sums are never written): a guarded load would be waited for
      // before the next one is issued
      VT v[G];
#pragma unroll
      for (int jj = 0; jj < G; ++jj) v[jj] = reinterpret_cast<const VT*>(Vb + (int64_t)(jj < kb ? jj : kb - 1) * ldv)[i];
#pragma unroll
      for (int jj = 0; jj < G; ++jj) {
        acc[jj] += Vec16<T>::dot(v[jj], a);
        if (NX == 2) acc[G + jj] += Vec16<T>::dot(v[jj], b);
      }
    }
  }
  {
    // scalar elements: all of them (no vector layout) or the tail beyond the last whole vector
    const int64_t m = n - body;
    const int64_t per = VECTOR ? m : (m + nb - 1) / nb;
    const int64_t lo = VECTOR ? (bx == nb - 1 ? 0 : m) : (int64_t)bx * per;
    const int64_t hi = lo + per < m ? lo + per : m;
    for (int64_t i = body + lo + tid; i < body + hi; i += kRedThreads) {
      const double a = (double)x0[i];
      const double b = NX == 2 ? (double)x1[i] : 0.0;
#pragma unroll
      for (int jj = 0; jj < G; ++jj) {
        const int r = jj < kb ? jj : kb - 1;
        const double v = (double)Vb[(int64_t)r * ldv + i];
        acc[jj] += v * a;
        if (NX == 2) acc[G + jj] += v * b;
      }
    }
  }
  const double wsum = wave_sum_t<NA>(acc, lane);
  constexpr int kSpread = 64 / NA;  // lanes per accumulator after the butterfly
  if (lane % kSpread == 0) red[wave][lane / kSpread] = wsum;
  __syncthreads();
  if (tid < NA && tid % G < kb) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < kRedThreads / 64; ++q) sum += red[q][tid];
    const int x = tid / G, jj = tid % G;
    partial[((int64_t)x * k + j0 + jj) * nb + bx] = sum;
  }
}

template <typename T, bool VECTOR, int NX>
__global__ void __launch_bounds__(kRedThreads) __attribute__((amdgpu_waves_per_eu(1, 4)))
dot_many_stage1(int64_t n, int k, int nb, int groups, const T* __restrict__ V, int64_t ldv, const T* __restrict__ x0,
                const T* __restrict__ x1, double* __restrict__ partial) {
  dot_many_stage1_body<T, VECTOR, NX>(blockIdx.x, n, k, nb, groups, V, ldv, x0, x1, partial);
}

// out[x * ostride + j] = sum over b of partial[(x * k + j) * nb + b]: one wave per output
__device__ __forceinline__ void dot_many_stage2_body(int blk, int nb, int k, int nx, const double* __restrict__ partial,
                                                     double* __restrict__ out, int64_t ostride) {
  const int o = blk * (kRedThreads / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= nx * k) return;
  const double* po = partial + (int64_t)o * nb;
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += po[b];
  s = wave_sum(s);
  if (lane == 0) out[(int64_t)(o / k) * ostride + o % k] = s;
}

__global__ void __launch_bounds__(kRedThreads) dot_many_stage2(int nb, int k, int nx,
                                                               const double* __restrict__ partial,
                                                               double* __restrict__ out, int64_t ostride) {
  dot_many_stage2_body(blockIdx.x, nb, k, nx, partial, out, ostride);
}

// Both stages; nx = 1 (x1 unused) or 2.
static void dot_many_launch(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* x0,
                            const void* x1, int nx, double* out, int64_t ostride, void* work, hipStream_t s) {
  const int G = kDotAcc / nx;
  const int groups = (k + G - 1) / G;
  const int nb = dot_chunks(n, groups, dtype == IGA_F32 ? 4 : 8);
  double* partial = (double*)work;
  const dim3 grid((unsigned)(((nb + 7) / 8) * 8 * groups));
  auto aligned = [](const void* p) { return p == nullptr || ((uintptr_t)p & 15) == 0; };
  const int width = dtype == IGA_F32 ? 4 : 2;
  // with one row the leading dimension is never used
  const bool vec = (k == 1 || ldv % width == 0) && aligned(V) && aligned(x0) && aligned(x1);
#define IGA_DOT1(T, VEC, NX) \
  dot_many_stage1<T, VEC, NX><<<grid, kRedThreads, 0, s>>>(n, k, nb, groups, (const T*)V, ldv, (const T*)x0, \
                                                           (const T*)x1, partial)
#define IGA_DOT2(T, VEC) \
  do {                    \
    if (nx == 2)          \
      IGA_DOT1(T, VEC, 2); \
    else                  \
      IGA_DOT1(T, VEC, 1); \
  } while (0)
  if (dtype == IGA_F64) {
    if (vec) IGA_DOT2(double, true); else IGA_DOT2(double, false);
  } else if (dtype == IGA_F32) {
    if (vec) IGA_DOT2(float, true); else IGA_DOT2(float, false);
  } else {
    throw Error(IGA_ERR_INVALID, "unknown dtype");
  }
#undef IGA_DOT2
#undef IGA_DOT1
  IGA_LAUNCH_CHECK();
  const int outs = nx * k;
  dot_many_stage2<<<(outs + kRedThreads / 64 - 1) / (kRedThreads / 64), kRedThreads, 0, s>>>(nb, k, nx, partial,
                                                                                            out, ostride);
  IGA_LAUNCH_CHECK();
}

template <typename T>
__global__ void axpy_many_kernel(int64_t n, int k, const T* __restrict__ V, int64_t ldv,
                                 const double* __restrict__ h, double alpha, T* __restrict__ w) {
  extern __shared__ double hs[];
  for (int j = threadIdx.x; j < k; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
#pragma unroll 8
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
// Arnoldi with delayed classical Gram-Schmidt reorthogonalisation (DCGS2:
// Bielich, Langou, Thomas, Swirydowicz, Yamazaki, Boman, "Low-synch Gram-Schmidt
// with delayed reorthogonalization for Krylov solvers", Parallel Computing 2022).
// Step k (1-based) finds the basis q_0..q_{j-1} (j = k - 1) final in V[0..j) and
// in V[j] the tentative vector u_j = A' q_{j-1} - Q hcol_{j-1}, projected once and
// not normalised; w = A' u_j.  One pass over V[0..j] and w gives
//   s = Q^T u_j, a = u_j.u_j, t = Q^T w, c = u_j.w,
// from which, on the device: rho = sqrt(a - s.s) and q_j = (u_j - Q s) / rho; the
// Hessenberg column j-1 becomes final (hcol_{j-1} + s, rho) and its Givens
// rotation yields the residual estimate of Krylov dimension j; with
// y = H[0:j+1, 0:j] s, A' q_j = (w - Q_{j+1} y) / rho is projected once:
//   hcol_j = (Q_{j+1}^T w - y) / rho,   u_{j+1} = w / rho - Q_{j+1} (y / rho + hcol_j).
// A second pass over V[0..j] and w writes q_j into V[j] and u_{j+1} into V[j+1].
// Two passes over the basis per step instead of the four of two-pass classical
// Gram-Schmidt; the residual estimate of dimension j arrives with step j + 1.
//
// Per-solve state (float64) for Krylov dimension m, at fixed offsets:
//   R (m x m, column-major, the rotated Hessenberg columns), cs, sn (m): the
//   Givens rotations, g (m+1): the rotated beta e_1, hist 2 (m+1): (g[d], rho_d)
//   for dimension d, the values a host polls; U, W (m+1 each): the dots of the
//   first pass with u and w; cq, cw (m+1 each) and scal: the second pass's
//   coefficients; H ((m+1) x (m+1), row-major): the Hessenberg matrix, its
//   last column pending.
constexpr int kGmMaxDim = 2000;  // the step kernel stages 4 m + 1 doubles in LDS (64 KiB)
constexpr int kGmThreads = 256;

struct GmState {
  int m;
  __host__ __device__ int64_t R() const { return 0; }
  __host__ __device__ int64_t cs() const { return (int64_t)m * m; }
  __host__ __device__ int64_t sn() const { return cs() + m; }
  __host__ __device__ int64_t g() const { return sn() + m; }
  __host__ __device__ int64_t hist() const { return g() + m + 1; }
  __host__ __device__ int64_t U() const { return hist() + 2 * ((int64_t)m + 1); }
  __host__ __device__ int64_t W() const { return U() + m + 1; }
  __host__ __device__ int64_t cq() const { return W() + m + 1; }
  __host__ __device__ int64_t cw() const { return cq() + m + 1; }
  __host__ __device__ int64_t scal() const { return cw() + m + 1; }  // 1/rho, z_j/rho, beta^2, 1/beta
  __host__ __device__ int64_t H() const { return scal() + 8; }
  __host__ __device__ int64_t ldh() const { return (int64_t)m + 1; }
  __host__ __device__ int64_t Y() const { return H() + ldh() * ldh(); }  // y = H s (m + 1)
  __host__ __device__ int64_t size() const { return Y() + m + 1; }
};

// From this basis size on, y = H[0:j, 0:j] s is formed by gm_hs_kernel (one wave per
// row, many work-groups) before the small dense kernel, whose single work-group
// otherwise spends most of its time in the O(j^2) product.
constexpr int kGmHsFrom = 24;

// y_i = sum_l H[i][l] s_l for i < j with column j - 1 taken final (H[i][j-1] + s_i),
// s = U[0..j); the Hessenberg rows start at column i - 1.  Row j (rho s_{j-1}) is
// left to the small dense kernel.  One wave per row.
__device__ __forceinline__ void gm_hs_body(int blk, GmState st, int j, double* ls) {
  const int lane = threadIdx.x & 63;
  const int i = blk * 4 + (threadIdx.x >> 6);
  if (i >= j) return;
  const double* H = ls + st.H();
  const double* s = ls + st.U();
  const int64_t ldh = st.ldh();
  double acc = 0.0;
  for (int l = (i > 0 ? i - 1 : 0) + lane; l < j; l += 64) {
    double h = H[(int64_t)i * ldh + l];
    if (l == j - 1) h += s[i];
    acc += h * s[l];
  }
  acc = wave_sum(acc);
  if (lane == 0) ls[st.Y() + i] = acc;
}

__global__ void __launch_bounds__(256) gm_hs_kernel(GmState st, int j, double* ls) { gm_hs_body(blockIdx.x, st, j, ls); }

__global__ void gm_zero_kernel(int64_t n, double* __restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Hist entry d (g[d], rho_d) also goes to host-visible memory when `rec` is set:
// rec[2 + 2d], rec[3 + 2d], then the release store of tag + d into rec[0] (as int64),
// so that a host that reads rec[0] and finds tag + d sees the values of every entry up
// to d.  The tag tells one solve's entries from another's in a reused buffer.
// Increment-criterion record of a solve (host-visible, written by the small dense
// kernel): entry j, at 1 + j * gm_inc_stride(m), holds R[0..j-1][j-1] (m slots), g[j - 1]
// (slot m) and the Gram row z_j . z_0..z_j of the preconditioned vectors (slots m + 1
// on); the int64 at 0 carries tag + j once entry j is complete.
__host__ __device__ constexpr int64_t gm_inc_stride(int m) { return 2 * (int64_t)m + 3; }

__device__ inline void gm_publish(double* rec, int64_t tag, int d, double g, double rho) {
  if (!rec) return;
  rec[2 + 2 * d] = g;
  rec[3 + 2 * d] = rho;
  __hip_atomic_store(reinterpret_cast<int64_t*>(rec), tag + d, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// beta = ||r|| from scal[2]; g[0] = beta, scal[3] = 1/beta (0 for r = 0), hist[0] = (beta, beta)
__global__ void gm_start_kernel(GmState st, double* ls, double* rec, int64_t tag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* sc = ls + st.scal();
  const double beta = sqrt(fmax(sc[2], 0.0));
  ls[st.g()] = beta;
  sc[3] = beta > 0.0 ? 1.0 / beta : 0.0;
  ls[st.hist()] = beta;
  ls[st.hist() + 1] = beta;
  gm_publish(rec, tag, 0, beta, beta);
}

// The small dense part of step j + 1 (one block): finalise column j - 1, rotate
// it, form hcol_j and the second pass's coefficients.  Every global value the
// kernel needs before the product y = H s is loaded up front, unconditionally
// (clamped indices), so that they arrive in one memory round trip.
constexpr int kGmPer = (kGmMaxDim + kGmThreads - 1) / kGmThreads;  // indices l per thread

// sh: s[j] | col[j+1] | cr[j] | sr[j]   (y[j+1] reuses col), dynamic LDS of the kernel
__device__ __forceinline__ void gm_dcgs2_body(GmState st, int j, double* ls, double* rec, int64_t tag,
                                              const GmIncrement& inc, double* sh) {
  __shared__ double red[2][kGmThreads / 64];
  __shared__ double pub[3];  // residual estimate, rho, final g[j - 1]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ldh = st.ldh();
  double* H = ls + st.H();
  double* s = sh;
  double* col = sh + j;
  double* cr = col + j + 1;
  double* sr = cr + j;
  const double* U = ls + st.U();
  const double* Wd = ls + st.W();
  double* sc = ls + st.scal();
  const int jm = j > 0 ? j - 1 : 0;
  // l = tid + c * kGmThreads: s_l, t_l, H[l, j-1] (hcol_{j-1}), cs_l, sn_l
  double sl[kGmPer], tl[kGmPer], hl[kGmPer], cl[kGmPer], snl[kGmPer];
#pragma unroll
  for (int c = 0; c < kGmPer; ++c) {
    const int l = tid + c * kGmThreads;
    const int lc = l < j ? l : 0;
    const int lr = l < jm ? l : 0;
    sl[c] = U[lc];
    tl[c] = Wd[lc];
    hl[c] = H[(int64_t)lc * ldh + jm];
    cl[c] = ls[st.cs() + lr];
    snl[c] = ls[st.sn() + lr];
  }
  const double a = U[j], cw_j = Wd[j];
  const double gj1 = ls[st.g() + jm];
  double ss = 0.0, stw = 0.0;
#pragma unroll
  for (int c = 0; c < kGmPer; ++c) {
    const int l = tid + c * kGmThreads;
    if (l < j) {
      s[l] = sl[c];
      ss += sl[c] * sl[c];
      stw += sl[c] * tl[c];
    }
  }
  ss = wave_sum(ss);
  stw = wave_sum(stw);
  if (lane == 0) {
    red[0][wave] = ss;
    red[1][wave] = stw;
  }
  __syncthreads();
  ss = 0.0;
  stw = 0.0;
#pragma unroll
  for (int q = 0; q < kGmThreads / 64; ++q) {
    ss += red[0][q];
    stw += red[1][q];
  }
  double rho = 1.0;  // step 1: V[0] = r / beta is final already
  if (j > 0) {
    const double r2 = a - ss;
    rho = r2 > 1e-14 * a ? sqrt(r2) : 0.0;  // breakdown: u_j lies in the span of Q
  }
  const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
  if (j > 0) {
    // final column j - 1: hcol_{j-1} + s, rho; staged with the rotations
#pragma unroll
    for (int c = 0; c < kGmPer; ++c) {
      const int l = tid + c * kGmThreads;
      if (l < j) {
        const double v = hl[c] + sl[c];
        H[(int64_t)l * ldh + jm] = v;
        col[l] = v;
        if (l < jm) {
          cr[l] = cl[c];
          sr[l] = snl[c];
        }
      }
    }
    if (tid == 0) {
      H[(int64_t)j * ldh + jm] = rho;
      col[j] = rho;
    }
    __syncthreads();
    if (tid == 0) {
      // previous rotations, the carried entry in a register and the LDS reads of
      // four rotations issued ahead of their use
      double x = col[0];
      int i = 0;
      for (; i + 4 < j; i += 4) {
        const double b0 = col[i + 1], b1 = col[i + 2], b2 = col[i + 3], b3 = col[i + 4];
        const double c0 = cr[i], c1 = cr[i + 1], c2 = cr[i + 2], c3 = cr[i + 3];
        const double s0 = sr[i], s1 = sr[i + 1], s2 = sr[i + 2], s3 = sr[i + 3];
        col[i] = c0 * x + s0 * b0;
        x = -s0 * x + c0 * b0;
        col[i + 1] = c1 * x + s1 * b1;
        x = -s1 * x + c1 * b1;
        col[i + 2] = c2 * x + s2 * b2;
        x = -s2 * x + c2 * b2;
        col[i + 3] = c3 * x + s3 * b3;
        x = -s3 * x + c3 * b3;
      }
      for (; i + 1 < j; ++i) {
        const double b = col[i + 1];
        col[i] = cr[i] * x + sr[i] * b;
        x = -sr[i] * x + cr[i] * b;
      }
      double cc = 1.0, sv = 0.0;
      if (rho != 0.0) {
        const double r = hypot(x, rho);
        cc = x / r;
        sv = rho / r;
      }
      col[j - 1] = cc * x + sv * rho;
      ls[st.cs() + j - 1] = cc;
      ls[st.sn() + j - 1] = sv;
      ls[st.g() + j] = -sv * gj1;
      ls[st.g() + j - 1] = cc * gj1;
      ls[st.hist() + 2 * j] = -sv * gj1;
      ls[st.hist() + 2 * j + 1] = rho;
      pub[0] = -sv * gj1;
      pub[1] = rho;
      pub[2] = cc * gj1;
    }
    __syncthreads();
    double* Rc = ls + st.R() + (int64_t)(j - 1) * st.m;
    for (int l = tid; l < j; l += kGmThreads) Rc[l] = col[l];
    if (inc.rec) {
      double* e = inc.rec + 1 + (int64_t)j * gm_inc_stride(st.m);
      for (int l = tid; l < j; l += kGmThreads) e[l] = col[l];
      if (tid == 0) e[st.m] = pub[2];
    }
    __syncthreads();
  }
  if (inc.rec) {
    // increment record entry j: R column j - 1 and g[j - 1] (above), the Gram row of the
    // preconditioned vector z_j against z_0..z_j; every thread's writes reach the
    // system before thread 0 releases the entry's tag
    double* e = inc.rec + 1 + (int64_t)j * gm_inc_stride(st.m);
    for (int l = tid; l <= j; l += kGmThreads) e[st.m + 1 + l] = inc.gram[l];
    __threadfence_system();
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(reinterpret_cast<int64_t*>(inc.rec), inc.tag + j, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (j > 0 && tid == 0) gm_publish(rec, tag, j, pub[0], pub[1]);
  // y = H[0:j+1, 0:j] s (Hessenberg: row i starts at column i - 1), into col: from
  // gm_hs_kernel for rows below j at larger j (row j is rho s_{j-1}); otherwise the
  // rows of the row-major H dotted with s, 16 rows per wave at a time, lanes along
  // the row (the entries below the band are zero)
  double* y = col;
  if (j >= kGmHsFrom) {
    for (int l = tid; l <= j; l += kGmThreads) y[l] = l < j ? ls[st.Y() + l] : rho * s[j - 1];
  } else
  for (int r0 = wave * 16; r0 <= j; r0 += 16 * (kGmThreads / 64)) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int l = (r0 > 0 ? r0 - 1 : 0) + lane; l < j; l += 64) {
      const double slv = s[l];
      double h[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)  // unconditional loads (rows past j re-read row j)
        h[q] = H[(int64_t)(r0 + q <= j ? r0 + q : j) * ldh + l];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] += h[q] * slv;
    }
    const double v = wave_sum_t<16>(acc, lane);
    if ((lane & 3) == 0 && r0 + (lane >> 2) <= j) y[r0 + (lane >> 2)] = v;
  }
  __syncthreads();
  // hcol_j = (Q_{j+1}^T w - y) / rho with q_j.w = (c - s.t) / rho; z = y / rho + hcol_j
  const double qw = (cw_j - stw) * inv;
  const double hj = (qw - y[j]) * inv;
  const double zj = y[j] * inv + hj;
  double* Hj = H + j;  // column j
  double* cq = ls + st.cq();
  double* cw = ls + st.cw();
#pragma unroll
  for (int c = 0; c < kGmPer; ++c) {
    const int l = tid + c * kGmThreads;
    if (l < j) {
      const double h = (tl[c] - y[l]) * inv;
      const double z = y[l] * inv + h;
      Hj[(int64_t)l * ldh] = h;
      cq[l] = sl[c] * inv;           // q_j = u_j / rho - Q (s / rho)
      cw[l] = z - zj * sl[c] * inv;  // u_{j+1} = w / rho - (z_j / rho) u_j - Q (z - z_j s / rho)
    }
  }
  if (tid == 0) {
    Hj[(int64_t)j * ldh] = hj;
    sc[0] = inv;
    sc[1] = zj * inv;
  }
}

__global__ void __launch_bounds__(kGmThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
gm_dcgs2_kernel(GmState st, int j, double* ls, double* rec, int64_t tag, GmIncrement inc) {
  extern __shared__ double gm_sh[];
  gm_dcgs2_body(st, j, ls, rec, tag, inc, gm_sh);
}

// q_j = inv u_j - Q cq -> V[j];  u_{j+1} = inv w - zinv u_j - Q cw -> V[j+1]
template <typename T>
__device__ __forceinline__ void gm_update_body(int64_t n, int j, T* __restrict__ V, int64_t ldv,
                                               const T* __restrict__ w, const double* __restrict__ cq,
                                               const double* __restrict__ cw, const double* __restrict__ scal,
                                               double* hs) {
  // hs: cq[j], cw[j] (dynamic LDS of the kernel)
  for (int l = threadIdx.x; l < j; l += blockDim.x) {
    hs[l] = cq[l];
    hs[j + l] = cw[l];
  }
  __syncthreads();
  const double inv = scal[0], zinv = scal[1];
  T* Vj = V + (int64_t)j * ldv;
  T* Vn = Vj + ldv;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double aq = 0.0, aw = 0.0;
#pragma unroll 8
    for (int l = 0; l < j; ++l) {
      const double v = (double)V[(int64_t)l * ldv + i];
      aq += hs[l] * v;
      aw += hs[j + l] * v;
    }
    const double u = (double)Vj[i];
    Vj[i] = (T)(inv * u - aq);
    Vn[i] = (T)(inv * (double)w[i] - zinv * u - aw);
  }
}

template <typename T>
__global__ void gm_update_kernel(int64_t n, int j, T* __restrict__ V, int64_t ldv, const T* __restrict__ w,
                                 const double* __restrict__ cq, const double* __restrict__ cw,
                                 const double* __restrict__ scal) {
  extern __shared__ double gm_hs[];
  gm_update_body<T>(n, j, V, ldv, w, cq, cw, scal, gm_hs);
}

}  // namespace iga

using namespace iga;

extern "C" {

int64_t iga_reduce_workspace_bytes(int64_t n, int k) {
  if (n < 0 || k < 0) return 0;
  const int64_t kk = k > 0 ? k : 1;
  // many-dots with one or two vectors: one partial per (vector, row, chunk), for every k' <= k
  int64_t dots = 0;
  for (int nx = 1; nx <= 2; ++nx) {
    const int G = kDotAcc / nx;
    for (int64_t g = 1; g <= (kk + G - 1) / G; ++g) {
      const int64_t kg = g * G < kk ? g * G : kk;
      const int64_t b = (int64_t)dot_chunks(n, (int)g, 8) * kg * nx;
      if (b > dots) dots = b;
    }
  }
  const int64_t zdots = (int64_t)reduce_blocks(n) * 2 * kk;  // iga_zdot_many
  return (dots > zdots ? dots : zdots) * (int64_t)sizeof(double);
}

int iga_dot_many(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* w, double* out,
                 void* work, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && w && out && work && n >= 0 && k >= 0 && ldv >= n, "bad argument");
    IGA_CHECK_ARG(k <= kDotMaxVectors, "too many vectors");
    if (k == 0) return;
    dot_many_launch(dtype, n, k, V, ldv, w, nullptr, 1, out, k, work, as_stream(stream));
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

int iga_gmres_layout(int m, int64_t* offsets) {
  return guard([&] {
    IGA_CHECK_ARG(offsets && m >= 1, "bad argument");
    const GmState st{m};
    offsets[0] = st.R();
    offsets[1] = st.g();
    offsets[2] = st.hist();
  });
}

int iga_gmres_start(int dtype, int64_t n, int m, const void* r, void* V0, double* ls, void* work, void* stream) {
  return guard([&] { gmres_start_impl(dtype, n, m, r, V0, ls, work, nullptr, 0, stream); });
}

int iga_gmres_step(int dtype, int64_t n, int k, int m, void* V, int64_t ldv, void* w, double* ls, void* work,
                   void* stream) {
  return guard([&] { gmres_step_impl(dtype, n, k, m, V, ldv, w, ls, work, nullptr, 0, stream); });
}

int iga_host_alloc(int64_t bytes, void** out) {
  return guard([&] {
    IGA_CHECK_ARG(out && bytes >= 1, "bad argument");
    void* p = nullptr;
    IGA_HIP(hipHostMalloc(&p, (size_t)bytes, hipHostMallocCoherent));
    std::memset(p, 0, (size_t)bytes);
    *out = p;
  });
}

int iga_host_free(void* p) {
  return guard([&] {
    if (p) IGA_HIP(hipHostFree(p));
  });
}

}  // extern "C"

namespace iga {

void gmres_start_impl(int dtype, int64_t n, int m, const void* r, void* V0, double* ls, void* work, double* rec,
                      int64_t tag, void* stream) {
  IGA_CHECK_ARG(r && V0 && ls && work && n >= 1 && m >= 1, "bad argument");
  IGA_CHECK_ARG(m <= kGmMaxDim, "Krylov dimension above 2000");
  hipStream_t s = as_stream(stream);
  const GmState st{m};
  gm_zero_kernel<<<grid_for(st.size(), 256, 1024), 256, 0, s>>>(st.size(), ls);
  IGA_LAUNCH_CHECK();
  IGA_CHECK(iga_dot_many(dtype, n, 1, r, n, r, ls + st.scal() + 2, work, stream));
  gm_start_kernel<<<1, 64, 0, s>>>(st, ls, rec, tag);
  IGA_LAUNCH_CHECK();
  IGA_CHECK(iga_axpby(dtype, n, 0.0, ls + st.scal() + 3, 0, r, 0.0, V0, stream));
}

void gmres_step_impl(int dtype, int64_t n, int k, int m, void* V, int64_t ldv, void* w, double* ls, void* work,
                     double* rec, int64_t tag, void* stream, const GmIncrement* inc) {
  {
    IGA_CHECK_ARG(V && w && ls && work && n >= 1 && ldv >= n, "bad argument");
    IGA_CHECK_ARG(k >= 1 && k <= m + 1, "step outside the Krylov dimension");
    IGA_CHECK_ARG(m <= kGmMaxDim, "Krylov dimension above 2000");
    IGA_CHECK_ARG(dtype == IGA_F64 || dtype == IGA_F32, "unknown dtype");
    hipStream_t s = as_stream(stream);
    const GmState st{m};
    const int j = k - 1;
    const size_t es = dtype == IGA_F64 ? 8 : 4;
    const void* u = (const char*)V + (size_t)j * ldv * es;
    // first pass: V[0..j] against u = V[j] and w
    dot_many_launch(dtype, n, j + 1, V, ldv, u, w, 2, ls + st.U(), st.W() - st.U(), work, s);
    if (j >= kGmHsFrom) {
      gm_hs_kernel<<<(unsigned)((j + 3) / 4), 256, 0, s>>>(st, j, ls);
      IGA_LAUNCH_CHECK();
    }
    gm_dcgs2_kernel<<<1, kGmThreads, sizeof(double) * (4 * (size_t)j + 1), s>>>(st, j, ls, rec, tag,
                                                                                 inc ? *inc : GmIncrement{});
    IGA_LAUNCH_CHECK();
    // second pass
    const dim3 grid(grid_for(n, 256, 8192)), block(256);
    const size_t shm = sizeof(double) * 2 * (size_t)(j > 0 ? j : 1);
    if (dtype == IGA_F64)
      gm_update_kernel<double><<<grid, block, shm, s>>>(n, j, (double*)V, ldv, (const double*)w, ls + st.cq(),
                                                        ls + st.cw(), ls + st.scal());
    else
      gm_update_kernel<float><<<grid, block, shm, s>>>(n, j, (float*)V, ldv, (const float*)w, ls + st.cq(),
                                                       ls + st.cw(), ls + st.scal());
    IGA_LAUNCH_CHECK();
  }
}

// Grouped launches of the step's kernels: member blockIdx.y (blockIdx.x for the small
// dense kernel), each with its own arguments; blocks past a member's own grid exit.
template <typename T>
struct Dot1Member {
  int k, nb, groups;
  const T* V;
  int64_t ldv;
  const T* x0;
  const T* x1;
  double* partial;
};
template <typename T>
struct Dot1Group {
  Dot1Member<T> m[kGmGroupMax];
};
template <typename T, bool VECTOR>
__global__ void __launch_bounds__(kRedThreads) __attribute__((amdgpu_waves_per_eu(1, 4)))
dot_many_stage1_group(int64_t n, Dot1Group<T> a) {
  const Dot1Member<T>& d = a.m[blockIdx.y];
  dot_many_stage1_body<T, VECTOR, 2>(blockIdx.x, n, d.k, d.nb, d.groups, d.V, d.ldv, d.x0, d.x1, d.partial);
}

struct Dot2Member {
  int nb, k;
  const double* partial;
  double* out;
  int64_t ostride;
};
struct Dot2Group {
  Dot2Member m[kGmGroupMax];
};
__global__ void __launch_bounds__(kRedThreads) dot_many_stage2_group(Dot2Group a) {
  const Dot2Member& d = a.m[blockIdx.y];
  dot_many_stage2_body(blockIdx.x, d.nb, d.k, 2, d.partial, d.out, d.ostride);
}

struct GmMember {
  GmState st;
  int j;
  double* ls;
  double* rec;
  int64_t tag;
};
struct GmGroup {
  GmMember m[kGmGroupMax];
};
// members below kGmHsFrom carry j = 0 here: their blocks exit at once
__global__ void __launch_bounds__(256) gm_hs_group(GmGroup a) {
  const GmMember& d = a.m[blockIdx.y];
  gm_hs_body(blockIdx.x, d.st, d.j, d.ls);
}

__global__ void __launch_bounds__(kGmThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) gm_dcgs2_group(GmGroup a) {
  extern __shared__ double gm_sh[];
  const GmMember& d = a.m[blockIdx.x];
  gm_dcgs2_body(d.st, d.j, d.ls, d.rec, d.tag, GmIncrement{}, gm_sh);
}

template <typename T>
struct UpdMember {
  int j;
  T* V;
  int64_t ldv;
  const T* w;
  const double* ls;
  GmState st;
};
template <typename T>
struct UpdGroup {
  UpdMember<T> m[kGmGroupMax];
};
template <typename T>
__global__ void gm_update_group(int64_t n, UpdGroup<T> a) {
  extern __shared__ double gm_hs[];
  const UpdMember<T>& d = a.m[blockIdx.y];
  gm_update_body<T>(n, d.j, d.V, d.ldv, d.w, d.ls + d.st.cq(), d.ls + d.st.cw(), d.ls + d.st.scal(), gm_hs);
}

template <typename T>
static bool gmres_step_group_t(int64_t n, int g, const GmGroupMember* mem, hipStream_t s) {
  constexpr int width = sizeof(T) == 4 ? 4 : 2;
  constexpr int G = kDotAcc / 2;  // rows per group of the two-vector pass
  auto aligned = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  // the vector path of each member's own step, which the group must share
  bool vec0 = false;
  for (int i = 0; i < g; ++i) {
    const GmGroupMember& d = mem[i];
    const int j = d.k - 1;
    const size_t es = sizeof(T);
    const void* u = (const char*)d.V + (size_t)j * d.ldv * es;
    const bool v = (j + 1 == 1 || d.ldv % width == 0) && aligned(d.V) && aligned(u) && aligned(d.w);
    if (i == 0) vec0 = v;
    if (v != vec0) return false;
  }
  Dot1Group<T> a1{};
  Dot2Group a2{};
  GmGroup gs{};
  UpdGroup<T> au{};
  unsigned grid1 = 1, grid2 = 1, gridh = 0;
  int jmax = 0;
  for (int i = 0; i < g; ++i) {
    const GmGroupMember& d = mem[i];
    const GmState st{d.m};
    const int j = d.k - 1, k = j + 1;
    const int groups = (k + G - 1) / G;
    const int nb = dot_chunks(n, groups, sizeof(T));
    a1.m[i] = Dot1Member<T>{k, nb, groups, (const T*)d.V, d.ldv, (const T*)d.V + (int64_t)j * d.ldv,
                            (const T*)d.w, (double*)d.work};
    grid1 = std::max(grid1, (unsigned)(((nb + 7) / 8) * 8 * groups));
    a2.m[i] = Dot2Member{nb, k, (const double*)d.work, d.ls + st.U(), st.W() - st.U()};
    grid2 = std::max(grid2, (unsigned)((2 * k + kRedThreads / 64 - 1) / (kRedThreads / 64)));
    gs.m[i] = GmMember{st, j >= kGmHsFrom ? j : 0, d.ls, d.rec, d.tag};
    if (j >= kGmHsFrom) gridh = std::max(gridh, (unsigned)((j + 3) / 4));
    au.m[i] = UpdMember<T>{j, (T*)d.V, d.ldv, (const T*)d.w, d.ls, st};
    jmax = std::max(jmax, j);
  }
  if (vec0)
    dot_many_stage1_group<T, true><<<dim3(grid1, (unsigned)g), kRedThreads, 0, s>>>(n, a1);
  else
    dot_many_stage1_group<T, false><<<dim3(grid1, (unsigned)g), kRedThreads, 0, s>>>(n, a1);
  IGA_LAUNCH_CHECK();
  dot_many_stage2_group<<<dim3(grid2, (unsigned)g), kRedThreads, 0, s>>>(a2);
  IGA_LAUNCH_CHECK();
  if (gridh > 0) {
    gm_hs_group<<<dim3(gridh, (unsigned)g), 256, 0, s>>>(gs);
    IGA_LAUNCH_CHECK();
  }
  // the small dense kernel takes each member's own j (above, members below kGmHsFrom
  // carried 0 for the product kernel)
  for (int i = 0; i < g; ++i) gs.m[i].j = mem[i].k - 1;
  gm_dcgs2_group<<<(unsigned)g, kGmThreads, sizeof(double) * (4 * (size_t)jmax + 1), s>>>(gs);
  IGA_LAUNCH_CHECK();
  const dim3 gridu(grid_for(n, 256, 8192), (unsigned)g);
  gm_update_group<T><<<gridu, 256, sizeof(double) * 2 * (size_t)(jmax > 0 ? jmax : 1), s>>>(n, au);
  IGA_LAUNCH_CHECK();
  return true;
}

bool gmres_step_group_impl(int dtype, int64_t n, int g, const GmGroupMember* mem, void* stream) {
  IGA_CHECK_ARG(g >= 1 && g <= kGmGroupMax && mem && n >= 1, "bad group");
  for (int i = 0; i < g; ++i) {
    const GmGroupMember& d = mem[i];
    IGA_CHECK_ARG(d.V && d.w && d.ls && d.work && d.ldv >= n, "bad member");
    IGA_CHECK_ARG(d.k >= 1 && d.k <= d.m + 1, "step outside the Krylov dimension");
    IGA_CHECK_ARG(d.m <= kGmMaxDim, "Krylov dimension above 2000");
  }
  hipStream_t s = as_stream(stream);
  if (dtype == IGA_F64) return gmres_step_group_t<double>(n, g, mem, s);
  if (dtype == IGA_F32) return gmres_step_group_t<float>(n, g, mem, s);
  throw Error(IGA_ERR_INVALID, "unknown dtype");
}

int64_t gmres_increment_doubles(int m) { return 1 + ((int64_t)m + 1) * gm_inc_stride(m); }

void dot_many_impl(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* x, double* out, void* work,
                   void* stream) {
  IGA_CHECK_ARG(V && x && out && work && n >= 0 && k >= 0 && ldv >= n, "bad argument");
  IGA_CHECK_ARG(k <= kDotMaxVectors, "too many vectors");
  if (k == 0) return;
  dot_many_launch(dtype, n, k, V, ldv, x, nullptr, 1, out, k, work, as_stream(stream));
}

}  // namespace iga
