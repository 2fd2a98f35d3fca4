// k_batch.hip -- many dd SPD factor + solve instances of uneven size in one launch.
//
// A persistent kernel: each workgroup repeatedly takes the next instance from an atomic
// queue ordered by n^3, largest first (longest-processing-time order), copies S into its
// output L, factors it with cholesky_blocked(nb = 64) (crates/quad/src/linalg.rs:238-293,
// same per-element operation order, so each factor is bit-identical to the unbatched
// one), then solves L L^T X = B for the instance's right-hand sides: forward substitution
// in the reference order (linalg.rs:99-108) and a right-looking blocked backward solve.
// The whole batch (queue reset + persistent kernel) is captured once in a hipGraph and
// replayed; the queue is reset by a kernel, not a memset node.
//
// Layout: matrices packed one after another (instance i at sum_{j<i} n_j^2), hi plane then
// lo plane of the total; right-hand sides packed the same way, instance i's r-th vector at
// rhs_off_i + r*n_i.  Per-instance status: 0 ok, p+1 = Cholesky breakdown at pivot p.
#include <algorithm>
#include <numeric>
#include <vector>

#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int NB = 64;          // block size of the factorisation (cholesky_blocked default)
constexpr int NT = 512;         // threads per workgroup (8 wavefronts)
constexpr int MAXRHS = 8;

struct BatchDesc {
  int n;
  int pad;
  long long mat_off;
  long long rhs_off;
};

__device__ __forceinline__ double rlane(double v, int lane) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ dd rlane(dd v, int lane) { return mkdd(rlane(v.hi, lane), rlane(v.lo, lane)); }

// LDS, reused between phases
struct DiagLds {
  double Lh[NB][NB + 1], Ll[NB][NB + 1];
  double Sh[NB][NB + 1], Sl[NB][NB + 1];
};
struct TileLds {
  double As[2][2][16][64], Bs[2][2][16][64];  // [half][plane][k][row]
};
union BatchLds {
  DiagLds d;
  TileLds t;
};

__device__ __forceinline__ dd ldm(const double* h, const double* l, long o) { return mkdd(h[o], l[o]); }
__device__ __forceinline__ void stm(double* h, double* l, long o, dd v) { h[o] = v.hi; l[o] = v.lo; }

__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

__global__ __launch_bounds__(NT) void k_batch_chol_solve(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                         int count, int* __restrict__ counter,
                                                         const double* __restrict__ sh, const double* __restrict__ sl,
                                                         double* __restrict__ lh, double* __restrict__ ll,
                                                         const double* __restrict__ bh, const double* __restrict__ bl,
                                                         double* __restrict__ xh, double* __restrict__ xl, int nrhs,
                                                         int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  BatchLds& lds = *reinterpret_cast<BatchLds*>(smem_raw);
  __shared__ int s_idx, s_fail;
  __shared__ double s_inv[2];
  __shared__ double Yh[MAXRHS][NB], Yl[MAXRHS][NB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  for (;;) {
    if (tid == 0) s_idx = atomicAdd(counter, 1);
    __syncthreads();
    const int q = s_idx;
    __syncthreads();
    if (q >= count) break;
    const int inst = order[q];
    const long n = desc[inst].n;
    const long mo = desc[inst].mat_off, ro = desc[inst].rhs_off;
    const double* Sh = sh + mo;
    const double* Sl = sl + mo;
    double* Ah = lh + mo;
    double* Al = ll + mo;
    // L = S (the factorisation works in place on the copy)
    for (long i = tid; i < n * n; i += NT) {
      Ah[i] = Sh[i];
      Al[i] = Sl[i];
    }
    if (tid == 0) s_fail = 0;
    __syncthreads();

    for (long jb0 = 0; jb0 < n; jb0 += NB) {
      const int jb = (int)min((long)NB, n - jb0);
      // ---- diagonal block: unblocked algorithm on the block (linalg.rs:250-255)
      for (int idx = tid; idx < jb * jb; idx += NT) {
        int i = idx / jb, c = idx % jb;
        if (c <= i) {
          long o = (jb0 + i) * n + jb0 + c;
          lds.d.Lh[i][c] = Ah[o];
          lds.d.Ll[i][c] = Al[o];
          lds.d.Sh[i][c] = 0.0;
          lds.d.Sl[i][c] = 0.0;
        }
      }
      __syncthreads();
      for (int j = 0; j < jb; ++j) {
        if (tid == 0) {
          dd diag = dd_sub(mkdd(lds.d.Lh[j][j], lds.d.Ll[j][j]), mkdd(lds.d.Sh[j][j], lds.d.Sl[j][j]));
          if (diag.hi <= 0.0) {
            s_fail = 1;
            status[inst] = (int)(jb0 + j) + 1;
          } else {
            dd ljj = dd_sqrt(diag);
            lds.d.Lh[j][j] = ljj.hi;
            lds.d.Ll[j][j] = ljj.lo;
            dd inv = dd_recip(ljj);
            s_inv[0] = inv.hi;
            s_inv[1] = inv.lo;
          }
        }
        __syncthreads();
        if (s_fail) break;
        const dd inv = mkdd(s_inv[0], s_inv[1]);
        for (int i = j + 1 + tid; i < jb; i += NT) {
          dd v = dd_mul(dd_sub(mkdd(lds.d.Lh[i][j], lds.d.Ll[i][j]), mkdd(lds.d.Sh[i][j], lds.d.Sl[i][j])), inv);
          lds.d.Lh[i][j] = v.hi;
          lds.d.Ll[i][j] = v.lo;
        }
        __syncthreads();
        const int m = jb - j - 1;
        for (int idx = tid; idx < m * m; idx += NT) {
          int ii = idx / m, cc = idx % m;
          if (cc <= ii) {
            int i = j + 1 + ii, c = j + 1 + cc;
            dd v = dd_add(mkdd(lds.d.Sh[i][c], lds.d.Sl[i][c]),
                          dd_mul(mkdd(lds.d.Lh[i][j], lds.d.Ll[i][j]), mkdd(lds.d.Lh[c][j], lds.d.Ll[c][j])));
            lds.d.Sh[i][c] = v.hi;
            lds.d.Sl[i][c] = v.lo;
          }
        }
        __syncthreads();
      }
      if (s_fail) break;
      for (int idx = tid; idx < jb * jb; idx += NT) {
        int i = idx / jb, c = idx % jb;
        if (c <= i) {
          long o = (jb0 + i) * n + jb0 + c;
          Ah[o] = lds.d.Lh[i][c];
          Al[o] = lds.d.Ll[i][c];
        }
      }
      // ---- trsm_right_lt on the rows below (linalg.rs:224-235): one wavefront per row
      {
        dd invk = mkdd(0.0);
        if (lane < jb) invk = dd_recip(mkdd(lds.d.Lh[lane][lane], lds.d.Ll[lane][lane]));
        for (long row = jb0 + jb + wv; row < n; row += NT / 64) {
          dd t = mkdd(0.0), fin = mkdd(0.0);
          if (lane < jb) t = ldm(Ah, Al, row * n + jb0 + lane);
          for (int k = 0; k < jb; ++k) {
            dd bk = dd_mul(t, invk);
            if (lane == k) fin = bk;
            dd b = rlane(bk, k);
            if (lane > k && lane < jb) t = dd_sub(t, dd_mul(b, mkdd(lds.d.Lh[lane][k], lds.d.Ll[lane][k])));
          }
          if (lane < jb) stm(Ah, Al, row * n + jb0 + lane, fin);
        }
      }
      __syncthreads();
      // ---- trailing update (linalg.rs:268-279): s = sum_k p_ik p_ck from zero, a -= s.
      // 64 x 64 tiles of the lower triangle; each 256-thread half of the workgroup owns one
      const long t0 = jb0 + jb, rem = n - t0;
      if (rem > 0) {
        const long T = (rem + 63) / 64;
        const long ntiles = T * (T + 1) / 2;
        const int half = tid >> 8, ht = tid & 255, tx = ht & 15, ty = ht >> 4;
        for (long base = 0; base < ntiles; base += 2) {
          const long tile = base + half;
          long ti = 0, tc = 0;
          if (tile < ntiles) {  // tile -> (ti, tc), tc <= ti, row-major over the triangle
            ti = (long)((__builtin_sqrt(8.0 * (double)tile + 1.0) - 1.0) / 2.0);
            while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
            while (ti * (ti + 1) / 2 > tile) --ti;
            tc = tile - ti * (ti + 1) / 2;
          }
          const bool active = tile < ntiles;
          const long i0 = t0 + ti * 64, c0 = t0 + tc * 64;
          dd acc[4][4];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = mkdd(0.0);
          for (int k0 = 0; k0 < jb; k0 += 16) {
            {
              int r = ht >> 2, kc = (ht & 3) * 4;
#pragma unroll
              for (int qq = 0; qq < 4; ++qq) {
                int kk = k0 + kc + qq;
                long gi = i0 + r, gc = c0 + r;
                dd va = (active && gi < n && kk < jb) ? ldm(Ah, Al, gi * n + jb0 + kk) : mkdd(0.0);
                dd vb = (active && gc < n && kk < jb) ? ldm(Ah, Al, gc * n + jb0 + kk) : mkdd(0.0);
                lds.t.As[half][0][kc + qq][r] = va.hi;
                lds.t.As[half][1][kc + qq][r] = va.lo;
                lds.t.Bs[half][0][kc + qq][r] = vb.hi;
                lds.t.Bs[half][1][kc + qq][r] = vb.lo;
              }
            }
            __syncthreads();
            const int kmax = min(16, jb - k0);
            if (active) {
              for (int kk = 0; kk < kmax; ++kk) {
                dd av[4], bv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  av[r] = mkdd(lds.t.As[half][0][kk][ty * 4 + r], lds.t.As[half][1][kk][ty * 4 + r]);
                  bv[r] = mkdd(lds.t.Bs[half][0][kk][tx * 4 + r], lds.t.Bs[half][1][kk][tx * 4 + r]);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                  for (int c = 0; c < 4; ++c) acc[r][c] = dd_add(acc[r][c], dd_mul(av[r], bv[c]));
              }
            }
            __syncthreads();
          }
          if (active) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                long gi = i0 + ty * 4 + r, gc = c0 + tx * 4 + c;
                if (gi < n && gc <= gi) {
                  long o = gi * n + gc;
                  stm(Ah, Al, o, dd_sub(ldm(Ah, Al, o), acc[r][c]));
                }
              }
          }
        }
      }
      __syncthreads();
    }
    if (s_fail) {
      __syncthreads();
      continue;
    }
    if (tid == 0) status[inst] = 0;

    // ---- solves: forward in the reference order, backward right-looking
    const int R = nrhs;
    for (int idx = tid; idx < R * n; idx += NT) {
      xh[ro + idx] = bh[ro + idx];
      xl[ro + idx] = bl[ro + idx];
    }
    __syncthreads();
    for (long b0 = 0; b0 < n; b0 += NB) {
      const int bw = (int)min((long)NB, n - b0);
      if (wv < R) {  // wavefront r: diagonal block of right-hand side r, lane = row
        double* Xh = xh + ro + (long)wv * n;
        double* Xl = xl + ro + (long)wv * n;
        dd y = lane < bw ? ldm(Xh, Xl, b0 + lane) : mkdd(0.0);
        for (int j = 0; j < bw; ++j) {
          dd ljj = ldm(Ah, Al, (b0 + (lane < bw ? lane : 0)) * n + b0 + (lane < bw ? lane : 0));
          dd yj = dd_div(y, ljj);
          if (lane == j) y = yj;
          dd b = rlane(yj, j);
          if (lane > j && lane < bw) y = dd_sub(y, dd_mul(ldm(Ah, Al, (b0 + lane) * n + b0 + j), b));
        }
        if (lane < bw) {
          stm(Xh, Xl, b0 + lane, y);
          Yh[wv][lane] = y.hi;
          Yl[wv][lane] = y.lo;
        }
      }
      __syncthreads();
      for (long idx = tid; idx < (long)R * (n - b0 - bw); idx += NT) {
        const int r = (int)(idx / (n - b0 - bw));
        const long i = b0 + bw + idx % (n - b0 - bw);
        double* Xh = xh + ro + (long)r * n;
        double* Xl = xl + ro + (long)r * n;
        dd y = ldm(Xh, Xl, i);
        for (int j = 0; j < bw; ++j) y = dd_sub(y, dd_mul(ldm(Ah, Al, i * n + b0 + j), mkdd(Yh[r][j], Yl[r][j])));
        stm(Xh, Xl, i, y);
      }
      __syncthreads();
    }
    for (long b0 = ((n - 1) / NB) * NB; b0 >= 0; b0 -= NB) {
      const int bw = (int)min((long)NB, n - b0);
      if (wv < R) {
        double* Xh = xh + ro + (long)wv * n;
        double* Xl = xl + ro + (long)wv * n;
        dd y = lane < bw ? ldm(Xh, Xl, b0 + lane) : mkdd(0.0);
        for (int j = bw - 1; j >= 0; --j) {
          dd ljj = ldm(Ah, Al, (b0 + j) * n + b0 + j);
          dd xj = dd_div(y, ljj);
          if (lane == j) y = xj;
          dd b = rlane(xj, j);
          if (lane < j) y = dd_sub(y, dd_mul(ldm(Ah, Al, (b0 + j) * n + b0 + lane), b));
        }
        if (lane < bw) {
          stm(Xh, Xl, b0 + lane, y);
          Yh[wv][lane] = y.hi;
          Yl[wv][lane] = y.lo;
        }
      }
      __syncthreads();
      for (long idx = tid; idx < (long)R * b0; idx += NT) {
        const int r = (int)(idx / b0);
        const long i = idx % b0;
        double* Xh = xh + ro + (long)r * n;
        double* Xl = xl + ro + (long)r * n;
        dd y = ldm(Xh, Xl, i);
        for (int j = bw - 1; j >= 0; --j)
          y = dd_sub(y, dd_mul(ldm(Ah, Al, (b0 + j) * n + i), mkdd(Yh[r][j], Yl[r][j])));
        stm(Xh, Xl, i, y);
      }
      __syncthreads();
    }
  }
}

}  // namespace
}  // namespace hp

using namespace hp;

struct hp_batch {
  int count = 0, nrhs = 0;
  DevBuf desc, order, counter;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~hp_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {
int batch_prepare(hp_batch* b, const int* sizes, size_t count, int nrhs) {
  if (nrhs < 1 || nrhs > MAXRHS) {
    set_error("batch: nrhs must be in [1, %d]", MAXRHS);
    return HP_EINVAL;
  }
  std::vector<BatchDesc> d(count);
  long long mo = 0, ro = 0;
  for (size_t i = 0; i < count; ++i) {
    if (sizes[i] <= 0) {
      set_error("batch: instance %zu has size %d", i, sizes[i]);
      return HP_ESHAPE;
    }
    d[i].n = sizes[i];
    d[i].mat_off = mo;
    d[i].rhs_off = ro;
    mo += (long long)sizes[i] * sizes[i];
    ro += (long long)nrhs * sizes[i];
  }
  std::vector<int> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sizes[x] > sizes[y]; });
  int rc;
  if ((rc = b->desc.alloc(count * sizeof(BatchDesc))) || (rc = b->order.alloc(count * sizeof(int))) ||
      (rc = b->counter.alloc(sizeof(int))))
    return rc;
  HP_HIP(hipMemcpy(b->desc.p, d.data(), count * sizeof(BatchDesc), hipMemcpyHostToDevice));
  HP_HIP(hipMemcpy(b->order.p, ord.data(), count * sizeof(int), hipMemcpyHostToDevice));
  b->count = (int)count;
  b->nrhs = nrhs;
  static bool attr = false;
  if (!attr) {
    HP_HIP(hipFuncSetAttribute((const void*)k_batch_chol_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sizeof(BatchLds)));
    attr = true;
  }
  return HP_OK;
}

int batch_enqueue(hp_batch* b, const double* s_hi, const double* s_lo, double* l_hi, double* l_lo, const double* b_hi,
                  const double* b_lo, double* x_hi, double* x_lo, int* status, hipStream_t s) {
  int dev = 0, cus = 0;
  HP_HIP(hipGetDevice(&dev));
  HP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = std::max(1, std::min(b->count, cus));  // one resident workgroup per CU
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, (int*)b->counter.p);
  HP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_batch_chol_solve, dim3(grid), dim3(NT), sizeof(BatchLds), s, (const BatchDesc*)b->desc.p,
                     (const int*)b->order.p, b->count, (int*)b->counter.p, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi,
                     x_lo, b->nrhs, status);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
}  // namespace

extern "C" {

int hp_dev_batch_cholesky_solve(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream) {
  if (count == 0) return HP_OK;
  hp_batch b;
  int rc = batch_prepare(&b, sizes, count, nrhs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = batch_enqueue(&b, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s))) return rc;
  HP_HIP(hipStreamSynchronize(s));  // the descriptors are freed on return
  return HP_OK;
}

hp_batch* hp_batch_graph_create(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream) {
  hp_batch* b = new hp_batch();
  hipStream_t s = (hipStream_t)stream;
  hipStream_t cap = nullptr;
  if (count == 0 || batch_prepare(b, sizes, count, nrhs) != HP_OK) { delete b; return nullptr; }
  if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) { delete b; return nullptr; }
  bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
  ok = ok && batch_enqueue(b, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, cap) == HP_OK;
  ok = (hipStreamEndCapture(cap, &b->graph) == hipSuccess) && ok;
  ok = ok && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) == hipSuccess;
  (void)hipStreamDestroy(cap);
  (void)s;
  if (!ok) {
    set_error("batch graph capture failed");
    delete b;
    return nullptr;
  }
  return b;
}

int hp_batch_graph_launch(hp_batch* b, void* stream) {
  if (!b || !b->exec) {
    set_error("batch graph: null handle");
    return HP_EINVAL;
  }
  HP_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return HP_OK;
}

void hp_batch_graph_destroy(hp_batch* b) { delete b; }

}  // extern "C"
