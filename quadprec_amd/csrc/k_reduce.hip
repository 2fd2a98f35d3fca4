// k_reduce.hip -- dd dot / gemv (crates/quad/src/linalg.rs:7-40) and the Neumaier
// compensated sums (crates/compensated-sum/src/lib.rs:6-66) on the GPU, in the reference's
// summation orders:
//   * dot: below 1024 elements the serial dd sum; else one thread per 1024-element chunk
//     (serial inside the chunk, as dot_serial) and a left fold of the chunk sums from ZERO,
//     one of the association orders rayon's reduce may produce;
//   * gemv: y_i = dot_serial(row_i, x), one thread per row, the rows staged through LDS in
//     64 x 64 tiles so the loads are coalesced and each thread reads its row across banks;
//   * par_neumaier_sum(_f64): one thread per 4096-element chunk, then the Neumaier sum of the
//     partials in order; neumaier_sum(_f64): the whole sum on one thread (the reference's
//     serial order; use the par form for long inputs).
#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr long DOT_CHUNK = 1024, NEU_CHUNK = 4096;

__global__ void k_dot_chunks(const double* __restrict__ ah, const double* __restrict__ al,
                             const double* __restrict__ bh, const double* __restrict__ bl, long n,
                             double* __restrict__ ph, double* __restrict__ pl) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * DOT_CHUNK;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + DOT_CHUNK);
  dd s = mkdd(0.0);
  for (long i = c0; i < c1; ++i) s = dd_add(s, dd_mul(mkdd(ah[i], al[i]), mkdd(bh[i], bl[i])));
  ph[c] = s.hi;
  pl[c] = s.lo;
}
// n >= 1024: out = ((ZERO + p0) + p1) + ...; else out = p0, the serial dot (ZERO for n = 0)
__global__ void k_fold(const double* __restrict__ ph, const double* __restrict__ pl, long np, long n,
                       double* out_h, double* out_l) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool chunked = n >= DOT_CHUNK;
  dd acc = n ? mkdd(ph[0], pl[0]) : mkdd(0.0);
  if (chunked) {
    acc = mkdd(0.0);
    for (long c = 0; c < np; ++c) acc = dd_add(acc, mkdd(ph[c], pl[c]));
  }
  out_h[0] = acc.hi;
  out_l[0] = acc.lo;
}

// y_i = sum_j a_ij x_j in j order; 64 rows per block, 64 x 64 tiles through LDS
__global__ __launch_bounds__(64) void k_gemv(const double* __restrict__ ah, const double* __restrict__ al, long rows,
                                             long cols, const double* __restrict__ xh, const double* __restrict__ xl,
                                             double* __restrict__ yh, double* __restrict__ yl) {
  __shared__ double th[64][65], tl[64][65];
  __shared__ double sxh[64], sxl[64];
  const int t = threadIdx.x;
  const long r0 = (long)blockIdx.x * 64;
  dd s = mkdd(0.0);
  for (long c0 = 0; c0 < cols; c0 += 64) {
    const int w = (int)min(64L, cols - c0);
    for (int r = 0; r < 64; ++r) {  // row r of the tile: its w contiguous entries
      const long gr = r0 + r;
      const bool ok = gr < rows && t < w;
      th[r][t] = ok ? ah[gr * cols + c0 + t] : 0.0;
      tl[r][t] = ok ? al[gr * cols + c0 + t] : 0.0;
    }
    sxh[t] = t < w ? xh[c0 + t] : 0.0;
    sxl[t] = t < w ? xl[c0 + t] : 0.0;
    __syncthreads();
    for (int k = 0; k < w; ++k) s = dd_add(s, dd_mul(mkdd(th[t][k], tl[t][k]), mkdd(sxh[k], sxl[k])));
    __syncthreads();
  }
  if (r0 + t < rows) {
    yh[r0 + t] = s.hi;
    yl[r0 + t] = s.lo;
  }
}

// Neumaier step (lib.rs:10-17 / 29-36)
__device__ __forceinline__ void neu_f64(double& sum, double& comp, double v) {
  const double t = sum + v;
  if (fabs(sum) >= fabs(v)) comp += (sum - t) + v;
  else comp += (v - t) + sum;
  sum = t;
}
__device__ __forceinline__ void neu_dd(dd& sum, dd& comp, dd v) {
  const dd t = dd_add(sum, v);
  if (fabs(dd_abs(sum).hi) >= fabs(dd_abs(v).hi)) comp = dd_add(comp, dd_add(dd_sub(sum, t), v));
  else comp = dd_add(comp, dd_add(dd_sub(v, t), sum));
  sum = t;
}
// chunk c of `chunk` values -> its Neumaier sum (chunk = n: the serial sum on one thread)
__global__ void k_neu_f64(const double* __restrict__ v, long n, long chunk, double* __restrict__ part) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * chunk;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + chunk);
  double sum = 0.0, comp = 0.0;
  for (long i = c0; i < c1; ++i) neu_f64(sum, comp, v[i]);
  part[c] = sum + comp;
}
__global__ void k_neu_dd(const double* __restrict__ vh, const double* __restrict__ vl, long n, long chunk,
                         double* __restrict__ ph, double* __restrict__ pl) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * chunk;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + chunk);
  dd sum = mkdd(0.0), comp = mkdd(0.0);
  for (long i = c0; i < c1; ++i) neu_dd(sum, comp, mkdd(vh[i], vl[i]));
  const dd r = dd_add(sum, comp);
  ph[c] = r.hi;
  pl[c] = r.lo;
}

unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int dot_dev(const double* ah, const double* al, const double* bh, const double* bl, long n, double* out_h,
            double* out_l, hipStream_t s) {
  const long np = std::max(1L, (n + DOT_CHUNK - 1) / DOT_CHUNK);
  DevBuf part;
  int rc;
  if ((rc = part.alloc(2 * np * sizeof(double)))) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_dot_chunks, dim3(nblk(np, 64)), dim3(64), 0, s, ah, al, bh, bl, n, part.d(), part.d() + np);
    HP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, s, part.d(), part.d() + np, np, n, out_h, out_l);
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

int gemv_dev(const double* ah, const double* al, long rows, long cols, const double* xh, const double* xl, double* yh,
             double* yl, hipStream_t s) {
  if (rows == 0) return HP_OK;
  hipLaunchKernelGGL(k_gemv, dim3(nblk(rows, 64)), dim3(64), 0, s, ah, al, rows, cols, xh, xl, yh, yl);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace hp

using namespace hp;

namespace {
int neumaier(const double* v, size_t n, int limbs, bool par, double* out) {
  if (!v || !out) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  out[0] = 0.0;
  if (limbs == 2) out[1] = 0.0;
  if (n == 0) return HP_OK;
  hipStream_t s = hipStreamPerThread;
  const bool chunked = par && (long)n >= NEU_CHUNK;  // lib.rs:43-66; else the serial sum
  const long chunk = chunked ? NEU_CHUNK : (long)n;
  const long np = ((long)n + chunk - 1) / chunk;
  DevBuf dv, part, fin;
  int rc;
  if ((rc = dv.alloc(limbs * n * sizeof(double))) || (rc = part.alloc(limbs * np * sizeof(double))) ||
      (rc = fin.alloc(limbs * sizeof(double))))
    return rc;
  if ((rc = upload_planes(v, n, limbs, dv.d(), s))) return rc;
  if (limbs == 1) {
    hipLaunchKernelGGL(k_neu_f64, dim3(nblk(np, 64)), dim3(64), 0, s, dv.d(), (long)n, chunk, part.d());
    HP_LAUNCH_CHECK();
    if (chunked) {  // the Neumaier sum of the partials, in order
      hipLaunchKernelGGL(k_neu_f64, dim3(1), dim3(64), 0, s, part.d(), np, np, fin.d());
      HP_LAUNCH_CHECK();
    } else {
      HP_HIP(hipMemcpyAsync(fin.p, part.p, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    HP_HIP(hipMemcpyAsync(out, fin.p, sizeof(double), hipMemcpyDeviceToHost, s));
  } else {
    hipLaunchKernelGGL(k_neu_dd, dim3(nblk(np, 64)), dim3(64), 0, s, dv.d(), dv.d() + n, (long)n, chunk, part.d(),
                       part.d() + np);
    HP_LAUNCH_CHECK();
    if (chunked) {
      hipLaunchKernelGGL(k_neu_dd, dim3(1), dim3(64), 0, s, part.d(), part.d() + np, np, np, fin.d(), fin.d() + 1);
      HP_LAUNCH_CHECK();
    } else {
      HP_HIP(hipMemcpyAsync(fin.p, part.p, sizeof(double), hipMemcpyDeviceToDevice, s));
      HP_HIP(hipMemcpyAsync(fin.d() + 1, part.d() + np, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    HP_HIP(hipMemcpyAsync(out, fin.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}
}  // namespace

extern "C" {

int hp_dot(const double* a, const double* b, size_t n, double* out_dd) {
  if (!out_dd || (n && (!a || !b))) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  DevBuf da, db, r;
  int rc;
  if ((rc = da.alloc(2 * n * sizeof(double))) || (rc = db.alloc(2 * n * sizeof(double))) || (rc = r.alloc(16)))
    return rc;
  if (n && ((rc = upload_planes(a, n, 2, da.d(), s)) || (rc = upload_planes(b, n, 2, db.d(), s)))) return rc;
  if ((rc = dot_dev(da.d(), da.d() + n, db.d(), db.d() + n, (long)n, r.d(), r.d() + 1, s))) return rc;
  HP_HIP(hipMemcpy(out_dd, r.p, 16, hipMemcpyDeviceToHost));
  return HP_OK;
}

int hp_dev_dot(const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo, size_t n,
               double* out_hi, double* out_lo, void* stream) {
  return dot_dev(a_hi, a_lo, b_hi, b_lo, (long)n, out_hi, out_lo, (hipStream_t)stream);
}

int hp_gemv(const double* a, size_t rows, size_t cols, const double* x, double* y) {
  if (rows == 0) return HP_OK;
  if (!a || !x || !y) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  DevBuf da, dx, dy;
  int rc;
  if ((rc = da.alloc(2 * rows * cols * sizeof(double))) || (rc = dx.alloc(2 * cols * sizeof(double))) ||
      (rc = dy.alloc(2 * rows * sizeof(double))))
    return rc;
  if ((rc = upload_planes(a, rows * cols, 2, da.d(), s)) || (cols && (rc = upload_planes(x, cols, 2, dx.d(), s))))
    return rc;
  if ((rc = gemv_dev(da.d(), da.d() + rows * cols, (long)rows, (long)cols, dx.d(), dx.d() + cols, dy.d(),
                     dy.d() + rows, s)))
    return rc;
  return download_planes(dy.d(), rows, 2, y, s);
}

int hp_dev_gemv(const double* a_hi, const double* a_lo, size_t rows, size_t cols, const double* x_hi,
                const double* x_lo, double* y_hi, double* y_lo, void* stream) {
  return gemv_dev(a_hi, a_lo, (long)rows, (long)cols, x_hi, x_lo, y_hi, y_lo, (hipStream_t)stream);
}

int hp_neumaier_sum_f64(const double* v, size_t n, double* out) { return neumaier(v, n, 1, false, out); }
int hp_par_neumaier_sum_f64(const double* v, size_t n, double* out) { return neumaier(v, n, 1, true, out); }
int hp_neumaier_sum(const double* v, size_t n, double* out_dd) { return neumaier(v, n, 2, false, out_dd); }
int hp_par_neumaier_sum(const double* v, size_t n, double* out_dd) { return neumaier(v, n, 2, true, out_dd); }

}  // extern "C"
