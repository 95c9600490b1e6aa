// Block operations on sets of long complex vectors (interleaved re, im float64),
// for the recycled Krylov solver: one pass over the data per call.
//   gram:    G[a][b] = sum_i conj(X_a[i]) Y_b[i]          (X, Y: lists of row pointers)
//   combine: Out_c[i] = beta Out_c[i] + sum_j coef[j][c] V_j[i]
// Rows are passed as device arrays of device addresses, so a block may mix rows of
// different allocations (the recycled basis C and the Arnoldi basis V, say).
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace iga {

constexpr int kGramThreads = 256;
constexpr int kGramChunk = 32;  // complex elements of every row staged per step
constexpr int kGramMaxRows = 64;

// partial[blk][a][b] (complex) over the block's chunks; then a fixed-order sum.
__global__ void __launch_bounds__(kGramThreads) zgram_stage1(int64_t n, int na, int nb,
                                                             const double* const* __restrict__ X,
                                                             const double* const* __restrict__ Y,
                                                             double* __restrict__ partial) {
  extern __shared__ double sm[];
  double* xs = sm;                             // [na][kGramChunk][2]
  double* ys = sm + (size_t)na * kGramChunk * 2;  // [nb][kGramChunk][2]
  const int npair = na * nb;
  const int nblk = gridDim.x;
  const int64_t per = (n + nblk - 1) / nblk;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  // each thread accumulates pairs p = tid, tid + 256, ...
  constexpr int kMaxPairsPerThread = (kGramMaxRows * kGramMaxRows + kGramThreads - 1) / kGramThreads;
  double accr[kMaxPairsPerThread], acci[kMaxPairsPerThread];
  const int mine = (npair - (int)threadIdx.x + kGramThreads - 1) / kGramThreads;
#pragma unroll
  for (int q = 0; q < kMaxPairsPerThread; ++q) accr[q] = acci[q] = 0.0;
  for (int64_t c0 = lo; c0 < hi; c0 += kGramChunk) {
    const int len = (int)((hi - c0) < kGramChunk ? (hi - c0) : kGramChunk);
    __syncthreads();
    for (int e = threadIdx.x; e < na * kGramChunk; e += kGramThreads) {
      const int a = e / kGramChunk, c = e % kGramChunk;
      const bool ok = c < len;
      xs[(a * kGramChunk + c) * 2] = ok ? X[a][2 * (c0 + c)] : 0.0;
      xs[(a * kGramChunk + c) * 2 + 1] = ok ? X[a][2 * (c0 + c) + 1] : 0.0;
    }
    for (int e = threadIdx.x; e < nb * kGramChunk; e += kGramThreads) {
      const int b = e / kGramChunk, c = e % kGramChunk;
      const bool ok = c < len;
      ys[(b * kGramChunk + c) * 2] = ok ? Y[b][2 * (c0 + c)] : 0.0;
      ys[(b * kGramChunk + c) * 2 + 1] = ok ? Y[b][2 * (c0 + c) + 1] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxPairsPerThread; ++q) {
      if (q < mine) {
        const int p = threadIdx.x + q * kGramThreads;
        const int a = p / nb, b = p % nb;
        const double* xr = xs + a * kGramChunk * 2;
        const double* yr = ys + b * kGramChunk * 2;
        double re = 0.0, im = 0.0;
        for (int c = 0; c < kGramChunk; ++c) {
          const double x0 = xr[2 * c], x1 = xr[2 * c + 1], y0 = yr[2 * c], y1 = yr[2 * c + 1];
          re += x0 * y0 + x1 * y1;
          im += x0 * y1 - x1 * y0;
        }
        accr[q] += re;
        acci[q] += im;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kMaxPairsPerThread; ++q) {
    if (q < mine) {
      const int p = threadIdx.x + q * kGramThreads;
      partial[((int64_t)blockIdx.x * npair + p) * 2] = accr[q];
      partial[((int64_t)blockIdx.x * npair + p) * 2 + 1] = acci[q];
    }
  }
}

__global__ void zgram_stage2(int nblk, int npair, const double* __restrict__ partial, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= 2 * npair) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * npair * 2 + p];
  out[p] = s;
}

constexpr int kCombMaxOut = 24;

__global__ void __launch_bounds__(256) zcombine_kernel(int64_t n, int nv, int nout, const double* const* __restrict__ V,
                                                       const double* __restrict__ coef, double beta,
                                                       double* const* __restrict__ out) {
  extern __shared__ double cs[];  // [nv][nout][2]
  for (int e = threadIdx.x; e < nv * nout * 2; e += blockDim.x) cs[e] = coef[e];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double re[kCombMaxOut], im[kCombMaxOut];
#pragma unroll
    for (int c = 0; c < kCombMaxOut; ++c) re[c] = im[c] = 0.0;
    for (int j = 0; j < nv; ++j) {
      const double vr = V[j][2 * i], vi = V[j][2 * i + 1];
#pragma unroll
      for (int c = 0; c < kCombMaxOut; ++c) {
        if (c < nout) {
          const double hr = cs[(j * nout + c) * 2], hi = cs[(j * nout + c) * 2 + 1];
          re[c] += hr * vr - hi * vi;
          im[c] += hr * vi + hi * vr;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kCombMaxOut; ++c) {
      if (c < nout) {
        double* o = out[c];
        const double br = beta == 0.0 ? 0.0 : beta * o[2 * i];
        const double bi = beta == 0.0 ? 0.0 : beta * o[2 * i + 1];
        o[2 * i] = br + re[c];
        o[2 * i + 1] = bi + im[c];
      }
    }
  }
}

static int gram_blocks(int64_t n) {
  int64_t g = (n + 8192 - 1) / 8192;
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return (int)g;
}

}  // namespace iga

using namespace iga;

extern "C" {

int64_t iga_zgram_workspace_bytes(int64_t n, int na, int nb) {
  if (n < 0 || na < 0 || nb < 0) return 0;
  return (int64_t)gram_blocks(n) * na * nb * 2 * (int64_t)sizeof(double) + 256;
}

int iga_zgram(int64_t n, int na, const double* const* X, int nb, const double* const* Y, double* out, void* work,
              void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(X && Y && out && work && n >= 0, "bad argument");
    IGA_CHECK_ARG(na >= 1 && nb >= 1 && na <= kGramMaxRows && nb <= kGramMaxRows, "row counts must be in 1..64");
    const int nblk = gram_blocks(n);
    const size_t shm = sizeof(double) * 2 * kGramChunk * (size_t)(na + nb);
    IGA_CHECK_ARG(shm <= 64 * 1024, "too many rows for one gram call");
    hipStream_t s = as_stream(stream);
    zgram_stage1<<<nblk, kGramThreads, shm, s>>>(n, na, nb, X, Y, (double*)work);
    IGA_LAUNCH_CHECK();
    const int np2 = 2 * na * nb;
    zgram_stage2<<<(np2 + 255) / 256, 256, 0, s>>>(nblk, na * nb, (const double*)work, out);
    IGA_LAUNCH_CHECK();
  });
}

int iga_zcombine(int64_t n, int nv, const double* const* V, int nout, const double* coef, double beta,
                 double* const* out, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && coef && out && n >= 0, "bad argument");
    IGA_CHECK_ARG(nv >= 1 && nout >= 1 && nout <= kCombMaxOut, "nout must be in 1..24");
    const size_t shm = sizeof(double) * 2 * (size_t)nv * nout;
    IGA_CHECK_ARG(shm <= 64 * 1024, "too many coefficients");
    if (n == 0) return;
    zcombine_kernel<<<grid_for(n, 256, 8192), 256, shm, as_stream(stream)>>>(n, nv, nout, V, coef, beta, out);
    IGA_LAUNCH_CHECK();
  });
}

}  // extern "C"
