// Sum-factorised sparse assembly (operators/utils/operations.py:59-115):
//   A[i, j] = sum_t s_t sum_q prod_d V_{t,d}(i_d, j_d, q_d) c_t[q]
// with V_{t,d}(i, j, q) = W_{t,d}[i, q] B_{t,d}[q, j], directions d = 1 .. ndim
// (direction 1 fastest in i, j and q), on the Kronecker product of the 1-D
// patterns {(i_d, j_d): some V_{t,d}(i_d, j_d, .) != 0}.  Rows are in C order,
// the columns of a row ascending (direction ndim slowest), so the values and
// column indices fill a CSR matrix directly.  The optional diagonal scaling
// A[i, j] *= w[i] w[j] is the NURBS scaling diag(w) A diag(w)
// (operators/nurbs_ops.py:194-196).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "common.hpp"

namespace iga {

constexpr int kAsmMaxDim = 4;
constexpr int kAsmMaxTerms = 32;

struct AsmArgs {
  int ndim, nterms;
  int64_t nr[kAsmMaxDim], nq[kAsmMaxDim], nc[kAsmMaxDim];
  const int32_t* pat_ptr[kAsmMaxDim];  // [nr_d + 1]: pairs of row i_d
  const int32_t* pat_col[kAsmMaxDim];  // [pairs_d]: j_d of each pair
  // per (term, direction): [pairs_d + 1] offsets into the (q, value) lists
  const int32_t* lst_ptr[kAsmMaxTerms][kAsmMaxDim];
  const int32_t* lst_q[kAsmMaxDim];
  const double* lst_v[kAsmMaxDim];
  const double* coef[kAsmMaxTerms];
  double scale[kAsmMaxTerms];
  const int64_t* row_ptr;  // [N_r + 1]
  const double* w;         // optional diagonal scaling (N_r == N_c)
  int64_t* cols;
  double* vals;
};

// One wave per row; lanes over the row's entries.
template <int NDIM>
__global__ void __launch_bounds__(64) assemble_kernel(AsmArgs a) {
  int64_t row = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
  int64_t nrows = 1;
  for (int d = 0; d < NDIM; ++d) nrows *= a.nr[d];
  if (row >= nrows) return;
  int id[NDIM], p0[NDIM], cnt[NDIM];
  int64_t r = row;
  for (int d = 0; d < NDIM; ++d) {
    id[d] = (int)(r % a.nr[d]);
    r /= a.nr[d];
    p0[d] = a.pat_ptr[d][id[d]];
    cnt[d] = a.pat_ptr[d][id[d] + 1] - p0[d];
  }
  const int64_t base = a.row_ptr[row];
  const int64_t nnz = a.row_ptr[row + 1] - base;
  for (int64_t e = threadIdx.x; e < nnz; e += blockDim.x) {
    int p[NDIM];
    int64_t rem = e, col = 0, cstride = 1;
    for (int d = 0; d < NDIM; ++d) {
      p[d] = p0[d] + (int)(rem % cnt[d]);
      rem /= cnt[d];
      col += (int64_t)a.pat_col[d][p[d]] * cstride;
      cstride *= a.nc[d];
    }
    double total = 0.0;
    for (int t = 0; t < a.nterms; ++t) {
      int lo[NDIM], hi[NDIM];
      for (int d = 0; d < NDIM; ++d) {
        lo[d] = a.lst_ptr[t][d][p[d]];
        hi[d] = a.lst_ptr[t][d][p[d] + 1];
      }
      const double* c = a.coef[t];
      // innermost sum over direction 1, outermost over direction NDIM
      double s_out = 0.0;
      if constexpr (NDIM == 3) {
        for (int k3 = lo[2]; k3 < hi[2]; ++k3) {
          const int64_t o3 = (int64_t)a.lst_q[2][k3] * a.nq[1];
          double s2 = 0.0;
          for (int k2 = lo[1]; k2 < hi[1]; ++k2) {
            const int64_t o2 = (o3 + a.lst_q[1][k2]) * a.nq[0];
            double s1 = 0.0;
            for (int k1 = lo[0]; k1 < hi[0]; ++k1) s1 += a.lst_v[0][k1] * c[o2 + a.lst_q[0][k1]];
            s2 += a.lst_v[1][k2] * s1;
          }
          s_out += a.lst_v[2][k3] * s2;
        }
      } else if constexpr (NDIM == 4) {
        for (int k4 = lo[3]; k4 < hi[3]; ++k4) {
          const int64_t o4 = (int64_t)a.lst_q[3][k4] * a.nq[2];
          double s3 = 0.0;
          for (int k3 = lo[2]; k3 < hi[2]; ++k3) {
            const int64_t o3 = (o4 + a.lst_q[2][k3]) * a.nq[1];
            double s2 = 0.0;
            for (int k2 = lo[1]; k2 < hi[1]; ++k2) {
              const int64_t o2 = (o3 + a.lst_q[1][k2]) * a.nq[0];
              double s1 = 0.0;
              for (int k1 = lo[0]; k1 < hi[0]; ++k1) s1 += a.lst_v[0][k1] * c[o2 + a.lst_q[0][k1]];
              s2 += a.lst_v[1][k2] * s1;
            }
            s3 += a.lst_v[2][k3] * s2;
          }
          s_out += a.lst_v[3][k4] * s3;
        }
      } else if constexpr (NDIM == 2) {
        for (int k2 = lo[1]; k2 < hi[1]; ++k2) {
          const int64_t o2 = (int64_t)a.lst_q[1][k2] * a.nq[0];
          double s1 = 0.0;
          for (int k1 = lo[0]; k1 < hi[0]; ++k1) s1 += a.lst_v[0][k1] * c[o2 + a.lst_q[0][k1]];
          s_out += a.lst_v[1][k2] * s1;
        }
      } else {
        for (int k1 = lo[0]; k1 < hi[0]; ++k1) s_out += a.lst_v[0][k1] * c[a.lst_q[0][k1]];
      }
      total += a.scale[t] * s_out;
    }
    if (a.w) total *= a.w[row] * a.w[col];
    a.cols[base + e] = col;
    a.vals[base + e] = total;
  }
}

}  // namespace iga

using namespace iga;

namespace {

// Row offsets of the Kronecker pattern: nnz(row) = prod_d (pairs of row i_d).
std::vector<int64_t> kron_row_ptr(int ndim, const int64_t* nr, const int32_t* const* pat_ptr) {
  int64_t nrows = 1;
  for (int d = 0; d < ndim; ++d) nrows *= nr[d];
  std::vector<int64_t> rp((size_t)nrows + 1, 0);
  int64_t id[kAsmMaxDim] = {0, 0, 0, 0};
  for (int64_t row = 0; row < nrows; ++row) {
    int64_t cnt = 1;
    for (int d = 0; d < ndim; ++d) cnt *= pat_ptr[d][id[d] + 1] - pat_ptr[d][id[d]];
    rp[(size_t)row + 1] = rp[(size_t)row] + cnt;
    for (int d = 0; d < ndim; ++d) {  // next row, direction 1 fastest
      if (++id[d] < nr[d]) break;
      id[d] = 0;
    }
  }
  return rp;
}

void check_monotone(const int32_t* ptr, int64_t n, int64_t last, const char* what) {
  IGA_CHECK_ARG(ptr[0] == 0 && ptr[n] == last, what);
  for (int64_t i = 0; i < n; ++i) IGA_CHECK_ARG(ptr[i] <= ptr[i + 1], what);
}

void check_range(const int32_t* v, int64_t n, int64_t hi, const char* what) {
  for (int64_t i = 0; i < n; ++i) IGA_CHECK_ARG(v[i] >= 0 && v[i] < hi, what);
}

}  // namespace

extern "C" {

int64_t iga_sumfact_nnz(int ndim, const int64_t* nr, const int32_t* const* pat_ptr) {
  if (ndim < 1 || ndim > kAsmMaxDim || !nr || !pat_ptr) return -1;
  for (int d = 0; d < ndim; ++d)
    if (nr[d] < 1 || !pat_ptr[d]) return -1;
  return kron_row_ptr(ndim, nr, pat_ptr).back();
}

int iga_sumfact_assemble(int ndim, int nterms, const int64_t* nr, const int64_t* nq, const int64_t* nc,
                         const int32_t* const* pat_ptr, const int32_t* const* pat_col, const int64_t* npairs,
                         const int32_t* const* lst_ptr, const int32_t* const* lst_q, const double* const* lst_v,
                         const int64_t* nlist, const double* const* coef, const double* scale,
                         int64_t* row_ptr, const double* w, int64_t* cols, double* vals, void* stream) {
  return guard([&] {
    IGA_CHECK_ARG(ndim >= 1 && ndim <= kAsmMaxDim, "ndim must be in 1..4");
    IGA_CHECK_ARG(nterms >= 1 && nterms <= kAsmMaxTerms, "between 1 and 32 terms");
    IGA_CHECK_ARG(nr && nq && nc && pat_ptr && pat_col && npairs && lst_ptr && lst_q && lst_v && nlist && coef &&
                      scale && row_ptr && cols && vals,
                  "NULL argument");
    // every table index the kernel follows is checked here
    for (int d = 0; d < ndim; ++d) {
      IGA_CHECK_ARG(nr[d] >= 1 && nq[d] >= 1 && nc[d] >= 1 && npairs[d] >= 0 && nlist[d] >= 0, "bad sizes");
      IGA_CHECK_ARG(pat_ptr[d] && (pat_col[d] || npairs[d] == 0), "NULL pattern");
      check_monotone(pat_ptr[d], nr[d], npairs[d], "pattern offsets do not match the pair count");
      check_range(pat_col[d], npairs[d], nc[d], "pattern column outside the trial space");
      IGA_CHECK_ARG((lst_q[d] && lst_v[d]) || nlist[d] == 0, "NULL lists");
      check_range(lst_q[d], nlist[d], nq[d], "list point outside the quadrature rule");
      for (int t = 0; t < nterms; ++t) {
        const int32_t* lp = lst_ptr[t * ndim + d];
        IGA_CHECK_ARG(lp, "NULL list offsets");
        IGA_CHECK_ARG(lp[0] >= 0 && lp[npairs[d]] <= nlist[d], "list offsets outside the lists");
        for (int64_t i = 0; i < npairs[d]; ++i) IGA_CHECK_ARG(lp[i] <= lp[i + 1], "list offsets not monotone");
      }
    }
    if (w)
      for (int d = 0; d < ndim; ++d) IGA_CHECK_ARG(nr[d] == nc[d], "the diagonal scaling needs a square matrix");
    const std::vector<int64_t> rp = kron_row_ptr(ndim, nr, pat_ptr);
    hipStream_t s = as_stream(stream);
    AsmArgs a{};
    a.ndim = ndim;
    a.nterms = nterms;
    int64_t nrows = 1;
    for (int d = 0; d < ndim; ++d) {
      a.nr[d] = nr[d];
      a.nq[d] = nq[d];
      a.nc[d] = nc[d];
      nrows *= nr[d];
    }
    // stage the 1-D tables in one device block: per direction the pattern, then
    // per (term, direction) the list offsets, then per direction the lists
    std::vector<size_t> off;
    size_t total = 0;
    auto reserve = [&](size_t bytes) {
      off.push_back(total);
      total += (bytes + 255) / 256 * 256;
    };
    for (int d = 0; d < ndim; ++d) {
      reserve(sizeof(int32_t) * (size_t)(nr[d] + 1));
      reserve(sizeof(int32_t) * (size_t)(npairs[d] > 0 ? npairs[d] : 1));
    }
    for (int t = 0; t < nterms; ++t)
      for (int d = 0; d < ndim; ++d) reserve(sizeof(int32_t) * (size_t)(npairs[d] + 1));
    for (int d = 0; d < ndim; ++d) {
      reserve(sizeof(int32_t) * (size_t)(nlist[d] > 0 ? nlist[d] : 1));
      reserve(sizeof(double) * (size_t)(nlist[d] > 0 ? nlist[d] : 1));
    }
    std::vector<unsigned char> host(total, 0);
    size_t k = 0;
    for (int d = 0; d < ndim; ++d) {
      std::memcpy(&host[off[k++]], pat_ptr[d], sizeof(int32_t) * (size_t)(nr[d] + 1));
      if (npairs[d] > 0) std::memcpy(&host[off[k]], pat_col[d], sizeof(int32_t) * (size_t)npairs[d]);
      ++k;
    }
    for (int t = 0; t < nterms; ++t)
      for (int d = 0; d < ndim; ++d) {
        std::memcpy(&host[off[k++]], lst_ptr[t * ndim + d], sizeof(int32_t) * (size_t)(npairs[d] + 1));
      }
    for (int d = 0; d < ndim; ++d) {
      if (nlist[d] > 0) std::memcpy(&host[off[k]], lst_q[d], sizeof(int32_t) * (size_t)nlist[d]);
      ++k;
      if (nlist[d] > 0) std::memcpy(&host[off[k]], lst_v[d], sizeof(double) * (size_t)nlist[d]);
      ++k;
    }
    unsigned char* dev = nullptr;
    IGA_HIP(hipMalloc(&dev, total));
    try {
      IGA_HIP(hipMemcpyAsync(dev, host.data(), total, hipMemcpyHostToDevice, s));
      IGA_HIP(hipMemcpyAsync(row_ptr, rp.data(), sizeof(int64_t) * rp.size(), hipMemcpyHostToDevice, s));
      k = 0;
      for (int d = 0; d < ndim; ++d) {
        a.pat_ptr[d] = (const int32_t*)(dev + off[k++]);
        a.pat_col[d] = (const int32_t*)(dev + off[k++]);
      }
      for (int t = 0; t < nterms; ++t)
        for (int d = 0; d < ndim; ++d) a.lst_ptr[t][d] = (const int32_t*)(dev + off[k++]);
      for (int d = 0; d < ndim; ++d) {
        a.lst_q[d] = (const int32_t*)(dev + off[k++]);
        a.lst_v[d] = (const double*)(dev + off[k++]);
      }
      for (int t = 0; t < nterms; ++t) {
        IGA_CHECK_ARG(coef[t], "NULL coefficients");
        a.coef[t] = coef[t];
        a.scale[t] = scale[t];
      }
      a.row_ptr = row_ptr;
      a.w = w;
      a.cols = cols;
      a.vals = vals;
      const int64_t gx = nrows < 65535 ? nrows : 65535;
      const dim3 grid((unsigned)gx, (unsigned)((nrows + gx - 1) / gx));
      switch (ndim) {
        case 1: assemble_kernel<1><<<grid, 64, 0, s>>>(a); break;
        case 2: assemble_kernel<2><<<grid, 64, 0, s>>>(a); break;
        case 3: assemble_kernel<3><<<grid, 64, 0, s>>>(a); break;
        default: assemble_kernel<4><<<grid, 64, 0, s>>>(a); break;
      }
      IGA_LAUNCH_CHECK();
      // the staging block is released once the kernel has read it
      IGA_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamSynchronize(s);  // the host tables stay alive until queued copies are done
      (void)hipFree(dev);
      throw;
    }
    IGA_HIP(hipFree(dev));
  });
}

}  // extern "C"
