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
constexpr int kGramMaxRows = 64;
constexpr int kGramMaxChunks = 1024;

typedef double v4d __attribute__((ext_vector_type(4)));

// X^H Y as a real product with inner dimension 2n over the interleaved doubles:
// output row a is X_a; output column 2b is Y_b itself (sum xr yr + xi yi = Re G_ab) and
// column 2b + 1 is Y_b with its parts swapped and the new imaginary part negated,
// (yi, -yr) (sum xr yi - xi yr = Im G_ab).  One 16-double step of the inner dimension
// is 8 complex elements: lane l reads the 2 complex elements 8 s + 2 (l / 16) + {0, 1}
// (32 contiguous bytes) of the row it feeds, and the four v_mfma_f64_16x16x4f64 of the
// step take element j of those four doubles (A[m][g] = X_m[16 s + 4 g + j],
// B[g][c] = Y~_c[16 s + 4 g + j]: the same inner index on both sides).  A block of
// 4 waves owns MT x NT fragments of 16 x 16 outputs and one chunk of steps; its waves
// take the chunk's steps round-robin, then add their fragments in LDS and write one
// partial per output: partial[o * nchunks + chunk], o = a * 2 nb + column.
template <int MT, int NT>
__global__ void __launch_bounds__(kGramThreads) zgram_mfma_kernel(int64_t n, int na, int nb, int tiles_n, int ntiles,
                                                                  int nchunks, const double* const* __restrict__ X,
                                                                  const double* const* __restrict__ Y,
                                                                  double* __restrict__ partial) {
  extern __shared__ double red[];  // [4 waves][MT * NT fragments][4 registers][64 lanes]
  // blocks go round-robin to the 8 XCDs: the tiles of one chunk share an XCD (and its L2)
  const int L = blockIdx.x;
  const int chunk = (L / 8 / ntiles) * 8 + L % 8;
  const int tile = (L / 8) % ntiles;
  if (chunk >= nchunks) return;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  const bool swap = (lane & 1) != 0;  // odd output column: imaginary part

  const double* xp[MT];
  const double* yp[NT];
#pragma unroll
  for (int mi = 0; mi < MT; ++mi) {
    const int a = tm * MT * 16 + mi * 16 + c16;
    xp[mi] = X[a < na ? a : na - 1];  // rows past na re-read the last row; their sums are dropped
  }
#pragma unroll
  for (int ni = 0; ni < NT; ++ni) {
    const int b = (tn * NT * 16 + ni * 16 + c16) >> 1;
    yp[ni] = Y[b < nb ? b : nb - 1];
  }

  const int64_t nsteps = (n + 7) / 8;
  const int64_t per = (nsteps + nchunks - 1) / nchunks;
  const int64_t lo = (int64_t)chunk * per;
  const int64_t hi = lo + per < nsteps ? lo + per : nsteps;
  const int64_t mine = hi > lo + wave ? (hi - lo - wave + 3) / 4 : 0;

  v4d acc[MT][NT];
#pragma unroll
  for (int mi = 0; mi < MT; ++mi)
#pragma unroll
    for (int ni = 0; ni < NT; ++ni) acc[mi][ni] = v4d{0.0, 0.0, 0.0, 0.0};

  // every load is issued unconditionally at a clamped address; elements past n or
  // steps past the chunk are zeroed by a select after the load
  auto load = [&](int64_t it, double2 (&xa)[MT][2], double2 (&yb)[NT][2]) __attribute__((always_inline)) {
    const int64_t s = lo + wave + 4 * it;
    const int64_t i0 = 8 * s + 2 * g;
    const bool ok0 = s < hi && i0 < n, ok1 = s < hi && i0 + 1 < n;
    const int64_t e0 = i0 < n ? i0 : n - 1, e1 = i0 + 1 < n ? i0 + 1 : n - 1;
#pragma unroll
    for (int mi = 0; mi < MT; ++mi) {
      xa[mi][0] = reinterpret_cast<const double2*>(xp[mi])[e0];
      xa[mi][1] = reinterpret_cast<const double2*>(xp[mi])[e1];
    }
#pragma unroll
    for (int ni = 0; ni < NT; ++ni) {
      yb[ni][0] = reinterpret_cast<const double2*>(yp[ni])[e0];
      yb[ni][1] = reinterpret_cast<const double2*>(yp[ni])[e1];
    }
#pragma unroll
    for (int mi = 0; mi < MT; ++mi) {
      if (!ok0) xa[mi][0] = double2{0.0, 0.0};
      if (!ok1) xa[mi][1] = double2{0.0, 0.0};
    }
#pragma unroll
    for (int ni = 0; ni < NT; ++ni) {
      if (!ok0) yb[ni][0] = double2{0.0, 0.0};
      if (!ok1) yb[ni][1] = double2{0.0, 0.0};
    }
  };
  auto step = [&](const double2 (&xa)[MT][2], const double2 (&yb)[NT][2]) __attribute__((always_inline)) {
    double bv[NT][4];
#pragma unroll
    for (int ni = 0; ni < NT; ++ni)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        bv[ni][2 * h] = swap ? yb[ni][h].y : yb[ni][h].x;
        bv[ni][2 * h + 1] = swap ? -yb[ni][h].x : yb[ni][h].y;
      }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int mi = 0; mi < MT; ++mi) {
        const double av = (j & 1) ? xa[mi][j >> 1].y : xa[mi][j >> 1].x;
#pragma unroll
        for (int ni = 0; ni < NT; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[ni][j], acc[mi][ni], 0, 0, 0);
      }
    }
  };

  double2 xa0[MT][2], yb0[NT][2], xa1[MT][2], yb1[NT][2];
  load(0, xa0, yb0);
  load(1, xa1, yb1);
  for (int64_t it = 0; it < mine; it += 2) {
    step(xa0, yb0);
    load(it + 2, xa0, yb0);
    step(xa1, yb1);  // past the wave's last step this adds zeros
    load(it + 3, xa1, yb1);
  }

  constexpr int F = MT * NT;
#pragma unroll
  for (int mi = 0; mi < MT; ++mi)
#pragma unroll
    for (int ni = 0; ni < NT; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[((wave * F + mi * NT + ni) * 4 + r) * 64 + lane] = acc[mi][ni][r];
  __syncthreads();
  const int ncol = 2 * nb;
  for (int e = threadIdx.x; e < F * 256; e += kGramThreads) {
    const double v = red[e] + red[F * 256 + e] + red[2 * F * 256 + e] + red[3 * F * 256 + e];
    const int f = e >> 8, r = (e >> 6) & 3, l = e & 63;
    const int mi = f / NT, ni = f % NT;
    // register r of lane l holds row l / 16 + 4 r, column l % 16 of the fragment
    const int a = tm * MT * 16 + mi * 16 + (l >> 4) + 4 * r;
    const int col = tn * NT * 16 + ni * 16 + (l & 15);
    if (a < na && col < ncol) partial[((int64_t)a * ncol + col) * nchunks + chunk] = v;
  }
}

// The same Gram for at most 2 x 2 rows (norms, a row against one or two others):
// the matrix-core tile would have its 16 lanes per operand row read the same
// addresses, so here each thread takes one complex element per row and step
// (16-byte loads) and keeps NA x NB complex sums; waves reduce by shuffles, the
// block through LDS, and the partial layout is the one of zgram_mfma_kernel.
template <int NA, int NB>
__global__ void __launch_bounds__(kGramThreads) zgram_small_kernel(int64_t n, int na, int nb, int nchunks,
                                                                   const double* const* __restrict__ X,
                                                                   const double* const* __restrict__ Y,
                                                                   double* __restrict__ partial) {
  constexpr int NV = 2 * NA * NB;
  __shared__ double red[kGramThreads / 64][NV];
  const int chunk = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double2* xp[NA];
  const double2* yp[NB];
#pragma unroll
  for (int a = 0; a < NA; ++a) xp[a] = reinterpret_cast<const double2*>(X[a < na ? a : na - 1]);
#pragma unroll
  for (int b = 0; b < NB; ++b) yp[b] = reinterpret_cast<const double2*>(Y[b < nb ? b : nb - 1]);
  const int64_t per = (n + nchunks - 1) / nchunks;
  const int64_t lo = (int64_t)chunk * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kGramThreads) {
    double2 xv[NA], yv[NB];
#pragma unroll
    for (int a = 0; a < NA; ++a) xv[a] = xp[a][i];
#pragma unroll
    for (int b = 0; b < NB; ++b) yv[b] = yp[b][i];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        acc[(a * NB + b) * 2] += xv[a].x * yv[b].x + xv[a].y * yv[b].y;
        acc[(a * NB + b) * 2 + 1] += xv[a].x * yv[b].y - xv[a].y * yv[b].x;
      }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double v = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[q] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[wave][q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kGramThreads / 64; ++w) v += red[w][q];
    const int a = q / (2 * NB), b = (q / 2) % NB, part = q & 1;
    if (a < na && b < nb) partial[((int64_t)a * 2 * nb + 2 * b + part) * nchunks + chunk] = v;
  }
}

// out[o] = sum over chunks of partial[o * nchunks + chunk], in chunk order: one wave per output
__global__ void __launch_bounds__(256) zgram_finish_kernel(int nchunks, int nout, const double* __restrict__ partial,
                                                           double* __restrict__ out) {
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= nout) return;
  const double* po = partial + (int64_t)o * nchunks;
  double s = 0.0;
  for (int b = lane; b < nchunks; b += 64) s += po[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[o] = s;
}

__global__ void zero_kernel(int64_t n, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
}

constexpr int kCombMaxOut = 24;

__global__ void __launch_bounds__(256) zcombine_kernel(int64_t n, int nv, int nout, const double* const* __restrict__ V,
                                                       const double* __restrict__ coef, double alpha, double beta,
                                                       double* const* __restrict__ out) {
  extern __shared__ double cs[];  // [nv][nout][2], scaled by alpha
  for (int e = threadIdx.x; e < nv * nout * 2; e += blockDim.x) cs[e] = alpha * coef[e];
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

// chunks of 16-double steps: at least 64 steps per chunk, at most kGramMaxChunks chunks
static int gram_chunks(int64_t n) {
  const int64_t nsteps = (n + 7) / 8;
  int64_t c = (nsteps + 63) / 64;
  if (c < 1) c = 1;
  if (c > kGramMaxChunks) c = kGramMaxChunks;
  return (int)c;
}

}  // namespace iga

using namespace iga;

extern "C" {

int64_t iga_zgram_workspace_bytes(int64_t n, int na, int nb) {
  if (n < 0 || na < 0 || nb < 0) return 0;
  return (int64_t)gram_chunks(n) * na * nb * 2 * (int64_t)sizeof(double) + 256;
}

int iga_zgram(int64_t n, int na, const double* const* X, int nb, const double* const* Y, double* out, void* work,
              void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(X && Y && out && work && n >= 0, "bad argument");
    IGA_CHECK_ARG(na >= 1 && nb >= 1 && na <= kGramMaxRows && nb <= kGramMaxRows, "row counts must be in 1..64");
    hipStream_t s = as_stream(stream);
    const int nout = 2 * na * nb;
    if (n == 0) {
      zero_kernel<<<(nout + 255) / 256, 256, 0, s>>>(nout, out);
      IGA_LAUNCH_CHECK();
      return;
    }
    const int nchunks_small = gram_chunks(n);
    if (na <= 2 && nb <= 2) {
      double* partial = (double*)work;
      if (na == 1 && nb == 1)
        zgram_small_kernel<1, 1><<<nchunks_small, kGramThreads, 0, s>>>(n, na, nb, nchunks_small, X, Y, partial);
      else
        zgram_small_kernel<2, 2><<<nchunks_small, kGramThreads, 0, s>>>(n, na, nb, nchunks_small, X, Y, partial);
      IGA_LAUNCH_CHECK();
      zgram_finish_kernel<<<(nout + 3) / 4, 256, 0, s>>>(nchunks_small, nout, partial, out);
      IGA_LAUNCH_CHECK();
      return;
    }
    const int fm = (na + 15) / 16, fn = (2 * nb + 15) / 16;  // 16 x 16 output fragments
    const int MT = fm < 2 ? fm : 2, NT = fn < 4 ? fn : 4;
    const int tiles_m = (fm + MT - 1) / MT, tiles_n = (fn + NT - 1) / NT, ntiles = tiles_m * tiles_n;
    const int nchunks = gram_chunks(n);
    const dim3 grid((unsigned)(((nchunks + 7) / 8) * 8 * ntiles));
    const size_t shm = sizeof(double) * 4 * 4 * 64 * (size_t)(MT * NT);
    double* partial = (double*)work;
#define IGA_ZGRAM(A, B)                                                                                            \
  if (MT == A && NT == B) {                                                                                        \
    zgram_mfma_kernel<A, B><<<grid, kGramThreads, shm, s>>>(n, na, nb, tiles_n, ntiles, nchunks, X, Y, partial);  \
  }
    IGA_ZGRAM(1, 1) IGA_ZGRAM(1, 2) IGA_ZGRAM(1, 3) IGA_ZGRAM(1, 4)
    IGA_ZGRAM(2, 1) IGA_ZGRAM(2, 2) IGA_ZGRAM(2, 3) IGA_ZGRAM(2, 4)
#undef IGA_ZGRAM
    IGA_LAUNCH_CHECK();
    zgram_finish_kernel<<<(nout + 3) / 4, 256, 0, s>>>(nchunks, nout, partial, out);
    IGA_LAUNCH_CHECK();
  });
}

int iga_zcombine(int64_t n, int nv, const double* const* V, int nout, const double* coef, double alpha, double beta,
                 double* const* out, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(V && coef && out && n >= 0, "bad argument");
    IGA_CHECK_ARG(nv >= 1 && nout >= 1 && nout <= kCombMaxOut, "nout must be in 1..24");
    const size_t shm = sizeof(double) * 2 * (size_t)nv * nout;
    IGA_CHECK_ARG(shm <= 64 * 1024, "too many coefficients");
    if (n == 0) return;
    zcombine_kernel<<<grid_for(n, 256, 8192), 256, shm, as_stream(stream)>>>(n, nv, nout, V, coef, alpha, beta,
                                                                              out);
    IGA_LAUNCH_CHECK();
  });
}

}  // extern "C"
