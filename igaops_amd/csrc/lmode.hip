// Generic banded L-mode products, point-wise coefficient products and box
// gather/scatter.  These are the building blocks of the composed operator path
// (any dimension, any quadrature type, space-time) and of MatrixFree.apply.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace iga {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace iga

struct iga_band_s {
  int nrows = 0, ncols = 0, width = 0, nmat = 0;
  int32_t* d_first = nullptr;
  double* d_vals64 = nullptr;
  float* d_vals32 = nullptr;
};

namespace iga {

// y_m(o, i, t) = alpha * sum_j V_m[i, j] x(o, first[i] + j, t) (+ y_m)
// One thread per output element; t (the inner, fastest index) maps to
// consecutive lanes so that x and y accesses coalesce whenever inner > 1.
template <typename T, int NMAT>
__global__ void __launch_bounds__(256) band_mode_kernel(
    int64_t total, int64_t inner, int nout, int nin, int width, const int32_t* __restrict__ first,
    const T* __restrict__ vals, const T* __restrict__ x, T* y0, T* y1, T* y2, T* y3, T alpha,
    int accumulate) {
  T* ys[4] = {y0, y1, y2, y3};
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = idx % inner;
    const int64_t oi = idx / inner;
    const int i = (int)(oi % nout);
    const int64_t o = oi / nout;
    const T* xp = x + ((o * nin + first[i]) * inner + t);
    T acc[NMAT];
#pragma unroll
    for (int m = 0; m < NMAT; ++m) acc[m] = T(0);
    for (int j = 0; j < width; ++j) {
      const T xv = xp[(int64_t)j * inner];
#pragma unroll
      for (int m = 0; m < NMAT; ++m) acc[m] += vals[((int64_t)m * nout + i) * width + j] * xv;
    }
#pragma unroll
    for (int m = 0; m < NMAT; ++m) {
      T v = alpha * acc[m];
      if (accumulate) v += ys[m][idx];
      ys[m][idx] = v;
    }
  }
}

template <typename T>
void launch_band(const iga_band_s* b, int64_t outer, int64_t inner, const void* x, void* const* y,
                 int accumulate, double alpha, hipStream_t s) {
  const int64_t total = outer * (int64_t)b->nrows * inner;
  if (total == 0) return;
  const T* vals = std::is_same<T, double>::value ? (const T*)b->d_vals64 : (const T*)b->d_vals32;
  T* ys[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int m = 0; m < b->nmat; ++m) ys[m] = (T*)y[m];
  dim3 grid(grid_for(total, 256)), block(256);
  switch (b->nmat) {
#define IGA_CASE(N)                                                                          \
  case N:                                                                                    \
    band_mode_kernel<T, N><<<grid, block, 0, s>>>(total, inner, b->nrows, b->ncols, b->width, \
                                                  b->d_first, vals, (const T*)x, ys[0], ys[1], \
                                                  ys[2], ys[3], (T)alpha, accumulate);        \
    break;
    IGA_CASE(1)
    IGA_CASE(2)
    IGA_CASE(3)
    IGA_CASE(4)
#undef IGA_CASE
    default:
      throw Error(IGA_ERR_INVALID, "nmat must be in 1..4");
  }
  IGA_LAUNCH_CHECK();
}

template <typename T>
__global__ void qp_scale_kernel(int64_t nq, int64_t r, const T* __restrict__ coef,
                                const T* __restrict__ x, T* y, T alpha, int accumulate) {
  const int64_t total = nq * r;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    T v = alpha * coef[idx / r] * x[idx];
    if (accumulate) v += y[idx];
    y[idx] = v;
  }
}

struct Box {
  int ndim;
  int64_t n[4], lo[4], len[4];
};

// Index arithmetic in I = int32_t when every index of the call, plus one grid
// stride (at most 2^28 threads), stays below 2^31 -- the hosts use it below 2^30
// elements -- else int64_t: a 32-bit division is a short sequence, a 64-bit one a
// long one.

// sub index s -> full index; the box is stored with direction 1 fastest.
template <typename I>
__device__ inline I box_offset(const Box& b, I s) {
  I off = 0, stride = 1;
  for (int d = 0; d < b.ndim; ++d) {
    const I len = (I)b.len[d];
    const I c = s % len;
    s /= len;
    off += ((I)b.lo[d] + c) * stride;
    stride *= (I)b.n[d];
  }
  return off;
}

// full index f -> whether it lies in the box, and its sub index when it does
template <typename I>
__device__ inline bool box_sub_index(const Box& b, I f, I& sub) {
  bool in = true;
  I s = 0, stride = 1;
  for (int d = 0; d < b.ndim; ++d) {
    const I n = (I)b.n[d];
    const I c = f % n - (I)b.lo[d];
    f /= n;
    in = in && c >= 0 && c < (I)b.len[d];
    s += c * stride;
    stride *= (I)b.len[d];
  }
  sub = s;
  return in;
}

template <typename T, typename I>
__global__ void box_gather_kernel(Box b, I nsub, I r, const T* __restrict__ full, T* __restrict__ sub) {
  const I total = nsub * r;
  for (I idx = blockIdx.x * (I)blockDim.x + threadIdx.x; idx < total; idx += (I)gridDim.x * blockDim.x) {
    const I s = idx / r, k = idx % r;
    sub[idx] = full[box_offset<I>(b, s) * r + k];
  }
}

template <typename T, typename I>
__global__ void box_scatter_kernel(Box b, I nsub, I r, const T* __restrict__ sub, T* __restrict__ full) {
  const I total = nsub * r;
  for (I idx = blockIdx.x * (I)blockDim.x + threadIdx.x; idx < total; idx += (I)gridDim.x * blockDim.x) {
    const I s = idx / r, k = idx % r;
    full[box_offset<I>(b, s) * r + k] = sub[idx];
  }
}

// every element of the full array: the box value inside, zero outside (one pass)
template <typename T, typename I>
__global__ void box_scatter_full_kernel(Box b, I nfull, I r, const T* __restrict__ sub, T* __restrict__ full) {
  const I total = nfull * r;
  for (I idx = blockIdx.x * (I)blockDim.x + threadIdx.x; idx < total; idx += (I)gridDim.x * blockDim.x) {
    const I f = idx / r, k = idx % r;
    I si;
    const bool in = box_sub_index<I>(b, f, si);
    full[idx] = in ? sub[si * r + k] : T(0);
  }
}

template <typename T, typename I>
__global__ void box_zero_outside_kernel(Box b, I nfull, I r, T* __restrict__ full) {
  const I total = nfull * r;
  for (I idx = blockIdx.x * (I)blockDim.x + threadIdx.x; idx < total; idx += (I)gridDim.x * blockDim.x) {
    I si;
    if (!box_sub_index<I>(b, idx / r, si)) full[idx] = T(0);
  }
}

template <typename T>
__global__ void index_zero_kernel(int64_t nidx, const int64_t* __restrict__ idx, int64_t r, T* __restrict__ x) {
  const int64_t total = nidx * r;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    x[idx[i / r] * r + i % r] = T(0);
}

static Box make_box(int ndim, const int64_t* n, const int64_t* lo, const int64_t* len, int64_t* nsub,
                    int64_t* nfull) {
  IGA_CHECK_ARG(ndim >= 1 && ndim <= 4, "box ndim must be in 1..4");
  Box b{};
  b.ndim = ndim;
  *nsub = 1;
  *nfull = 1;
  for (int d = 0; d < ndim; ++d) {
    IGA_CHECK_ARG(n[d] > 0 && lo[d] >= 0 && len[d] >= 0 && lo[d] + len[d] <= n[d],
                  "box does not fit in the array");
    b.n[d] = n[d];
    b.lo[d] = lo[d];
    b.len[d] = len[d];
    *nsub *= len[d];
    *nfull *= n[d];
  }
  return b;
}

}  // namespace iga

using namespace iga;

extern "C" {

int iga_version(void) { return 100; }

int iga_last_error(char* buf, int buflen) {
  if (!buf || buflen <= 0) return IGA_ERR_INVALID;
  std::strncpy(buf, g_last_error.c_str(), (size_t)buflen - 1);
  buf[buflen - 1] = '\0';
  return IGA_OK;
}

int iga_device_count(int* count) {
  return guard([&] {
    IGA_CHECK_ARG(count, "count is NULL");
    IGA_HIP(hipGetDeviceCount(count));
  });
}

int iga_band_create(iga_band_t* out, int nrows, int ncols, int width, int nmat, const int32_t* first,
                    const double* vals) {
  return guard([&] {
    IGA_CHECK_ARG(out && first && vals, "NULL argument");
    IGA_CHECK_ARG(nrows > 0 && ncols > 0 && width > 0 && width <= ncols, "bad band shape");
    IGA_CHECK_ARG(nmat >= 1 && nmat <= 4, "nmat must be in 1..4");
    for (int i = 0; i < nrows; ++i)
      IGA_CHECK_ARG(first[i] >= 0 && first[i] + width <= ncols, "band row leaves the matrix");
    auto* b = new iga_band_s();
    b->nrows = nrows;
    b->ncols = ncols;
    b->width = width;
    b->nmat = nmat;
    const size_t nv = (size_t)nmat * nrows * width;
    std::vector<float> v32(nv);
    for (size_t k = 0; k < nv; ++k) v32[k] = (float)vals[k];
    try {
      IGA_HIP(hipMalloc(&b->d_first, sizeof(int32_t) * nrows));
      IGA_HIP(hipMalloc(&b->d_vals64, sizeof(double) * nv));
      IGA_HIP(hipMalloc(&b->d_vals32, sizeof(float) * nv));
      IGA_HIP(hipMemcpy(b->d_first, first, sizeof(int32_t) * nrows, hipMemcpyHostToDevice));
      IGA_HIP(hipMemcpy(b->d_vals64, vals, sizeof(double) * nv, hipMemcpyHostToDevice));
      IGA_HIP(hipMemcpy(b->d_vals32, v32.data(), sizeof(float) * nv, hipMemcpyHostToDevice));
    } catch (...) {
      (void)hipFree(b->d_first);
      (void)hipFree(b->d_vals64);
      (void)hipFree(b->d_vals32);
      delete b;
      throw;
    }
    *out = b;
  });
}

int iga_band_destroy(iga_band_t band) {
  return guard([&] {
    if (!band) return;
    IGA_HIP(hipFree(band->d_first));
    IGA_HIP(hipFree(band->d_vals64));
    IGA_HIP(hipFree(band->d_vals32));
    delete band;
  });
}

int iga_band_mode_product(iga_band_t band, int dtype, int64_t outer, int64_t inner, const void* x,
                          void* const* y, int accumulate, double alpha, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(band && x && y, "NULL argument");
    IGA_CHECK_ARG(outer >= 0 && inner >= 1, "bad tensor view");
    for (int m = 0; m < band->nmat; ++m) IGA_CHECK_ARG(y[m], "NULL output");
    if (dtype == IGA_F64)
      launch_band<double>(band, outer, inner, x, y, accumulate, alpha, as_stream(stream));
    else if (dtype == IGA_F32)
      launch_band<float>(band, outer, inner, x, y, accumulate, alpha, as_stream(stream));
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
  });
}

int iga_qp_scale(int dtype, int64_t nq, int64_t nrhs, const void* coef, const void* x, void* y,
                 int accumulate, double alpha, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(coef && x && y && nq >= 0 && nrhs >= 1, "bad argument");
    const int64_t total = nq * nrhs;
    if (total == 0) return;
    dim3 grid(grid_for(total, 256)), block(256);
    if (dtype == IGA_F64)
      qp_scale_kernel<double><<<grid, block, 0, as_stream(stream)>>>(
          nq, nrhs, (const double*)coef, (const double*)x, (double*)y, alpha, accumulate);
    else if (dtype == IGA_F32)
      qp_scale_kernel<float><<<grid, block, 0, as_stream(stream)>>>(
          nq, nrhs, (const float*)coef, (const float*)x, (float*)y, (float)alpha, accumulate);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
  });
}

int iga_box_gather(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                   int64_t nrhs, const void* full, void* sub, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(n && lo && len && full && sub && nrhs >= 1, "bad argument");
    int64_t nsub, nfull;
    Box b = make_box(ndim, n, lo, len, &nsub, &nfull);
    if (nsub == 0) return;
    hipStream_t s = as_stream(stream);
    // the index type is chosen by the largest index the kernel forms (full offsets)
    const int64_t count = nfull * nrhs;
    if (count < ((int64_t)1 << 30)) {
      if (dtype == IGA_F64)
        box_gather_kernel<double, int32_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, (int32_t)nsub, (int32_t)nrhs, (const double*)full, (double*)sub);
      else if (dtype == IGA_F32)
        box_gather_kernel<float, int32_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, (int32_t)nsub, (int32_t)nrhs, (const float*)full, (float*)sub);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    } else {
      if (dtype == IGA_F64)
        box_gather_kernel<double, int64_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(b, nsub, nrhs,
                                                                                       (const double*)full,
                                                                                       (double*)sub);
      else if (dtype == IGA_F32)
        box_gather_kernel<float, int64_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(b, nsub, nrhs,
                                                                                      (const float*)full,
                                                                                      (float*)sub);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    }
    IGA_LAUNCH_CHECK();
  });
}

int iga_box_scatter(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                    int64_t nrhs, const void* sub, void* full, int zero_outside, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(n && lo && len && full && sub && nrhs >= 1, "bad argument");
    int64_t nsub, nfull;
    Box b = make_box(ndim, n, lo, len, &nsub, &nfull);
    hipStream_t s = as_stream(stream);
    const int64_t count = nfull * nrhs;
    const bool small = count < ((int64_t)1 << 30);
    if (zero_outside) {
      if (nfull == 0) return;
      if (small) {
        if (dtype == IGA_F64)
          box_scatter_full_kernel<double, int32_t><<<grid_for(count, 256), 256, 0, s>>>(
              b, (int32_t)nfull, (int32_t)nrhs, (const double*)sub, (double*)full);
        else if (dtype == IGA_F32)
          box_scatter_full_kernel<float, int32_t><<<grid_for(count, 256), 256, 0, s>>>(
              b, (int32_t)nfull, (int32_t)nrhs, (const float*)sub, (float*)full);
        else
          throw Error(IGA_ERR_INVALID, "unknown dtype");
      } else {
        if (dtype == IGA_F64)
          box_scatter_full_kernel<double, int64_t><<<grid_for(count, 256), 256, 0, s>>>(
              b, nfull, nrhs, (const double*)sub, (double*)full);
        else if (dtype == IGA_F32)
          box_scatter_full_kernel<float, int64_t><<<grid_for(count, 256), 256, 0, s>>>(
              b, nfull, nrhs, (const float*)sub, (float*)full);
        else
          throw Error(IGA_ERR_INVALID, "unknown dtype");
      }
      IGA_LAUNCH_CHECK();
      return;
    }
    if (nsub == 0) return;
    if (small) {
      if (dtype == IGA_F64)
        box_scatter_kernel<double, int32_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, (int32_t)nsub, (int32_t)nrhs, (const double*)sub, (double*)full);
      else if (dtype == IGA_F32)
        box_scatter_kernel<float, int32_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, (int32_t)nsub, (int32_t)nrhs, (const float*)sub, (float*)full);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    } else {
      if (dtype == IGA_F64)
        box_scatter_kernel<double, int64_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, nsub, nrhs, (const double*)sub, (double*)full);
      else if (dtype == IGA_F32)
        box_scatter_kernel<float, int64_t><<<grid_for(nsub * nrhs, 256), 256, 0, s>>>(
            b, nsub, nrhs, (const float*)sub, (float*)full);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    }
    IGA_LAUNCH_CHECK();
  });
}

int iga_box_mask(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len, int64_t nrhs,
                 void* full, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(n && lo && len && full && nrhs >= 1, "bad argument");
    int64_t nsub, nfull;
    Box b = make_box(ndim, n, lo, len, &nsub, &nfull);
    if (nfull == 0) return;
    hipStream_t s = as_stream(stream);
    const int64_t count = nfull * nrhs;
    if (count < ((int64_t)1 << 30)) {
      if (dtype == IGA_F64)
        box_zero_outside_kernel<double, int32_t><<<grid_for(count, 256), 256, 0, s>>>(b, (int32_t)nfull,
                                                                                       (int32_t)nrhs, (double*)full);
      else if (dtype == IGA_F32)
        box_zero_outside_kernel<float, int32_t><<<grid_for(count, 256), 256, 0, s>>>(b, (int32_t)nfull,
                                                                                      (int32_t)nrhs, (float*)full);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    } else {
      if (dtype == IGA_F64)
        box_zero_outside_kernel<double, int64_t><<<grid_for(count, 256), 256, 0, s>>>(b, nfull, nrhs, (double*)full);
      else if (dtype == IGA_F32)
        box_zero_outside_kernel<float, int64_t><<<grid_for(count, 256), 256, 0, s>>>(b, nfull, nrhs, (float*)full);
      else
        throw Error(IGA_ERR_INVALID, "unknown dtype");
    }
    IGA_LAUNCH_CHECK();
  });
}

int iga_index_zero(int dtype, int64_t nidx, const int64_t* idx, int64_t nrhs, void* x, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(x && nrhs >= 1 && nidx >= 0, "bad argument");
    if (nidx == 0) return;
    IGA_CHECK_ARG(idx, "NULL index array");
    dim3 g(grid_for(nidx * nrhs, 256)), blk(256);
    if (dtype == IGA_F64)
      index_zero_kernel<double><<<g, blk, 0, as_stream(stream)>>>(nidx, idx, nrhs, (double*)x);
    else if (dtype == IGA_F32)
      index_zero_kernel<float><<<g, blk, 0, as_stream(stream)>>>(nidx, idx, nrhs, (float*)x);
    else
      throw Error(IGA_ERR_INVALID, "unknown dtype");
    IGA_LAUNCH_CHECK();
  });
}

}  // extern "C"
