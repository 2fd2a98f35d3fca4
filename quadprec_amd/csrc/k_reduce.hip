// k_reduce.hip -- dd dot / gemv (crates/quad/src/linalg.rs:7-40) and the Neumaier
// compensated sums (crates/compensated-sum/src/lib.rs:6-66) on the GPU, in the reference's
// summation orders:
//   * dot: below 1024 elements the serial dd sum; else one thread per 1024-element chunk
//     (serial inside the chunk, as dot_serial) and the chunk sums combined by rayon's
//     halving tree (split at len / 2 down to single chunks), reduced pairwise with DPP /
//     ds_swizzle across each wavefront;
//   * gemv: y_i = dot_serial(row_i, x), one thread per row; three loader wavefronts stage
//     64 x 32 tiles through double-buffered LDS while the fourth runs the row chains;
//   * par_neumaier_sum(_f64): one thread per 4096-element chunk, then the Neumaier sum of the
//     partials in order; neumaier_sum(_f64): the whole sum in order.  The in-order sums run
//     on one wavefront that loads 64 values at a time ahead of the chain.
#include <mutex>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr long DOT_CHUNK = 1024, NEU_CHUNK = 4096;

// one thread per 1024-element chunk, its serial dd sum; the loads of the next 8 elements
// are issued before their multiply-adds, so the dependent chain does not wait on memory
__global__ __launch_bounds__(64) void k_dot_chunks(const double* __restrict__ ah, const double* __restrict__ al,
                                                   const double* __restrict__ bh, const double* __restrict__ bl,
                                                   long n, double* __restrict__ ph, double* __restrict__ pl) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * DOT_CHUNK;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + DOT_CHUNK);
  dd s = mkdd(0.0);
  // the next 8 elements are loaded while the current 8 go through the chain (addresses
  // clamped into the chunk; elements past its end are never used)
  double x0[8], x1[8], y0[8], y1[8];
  auto fetch = [&](long i) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long q = min(i + e, c1 - 1);
      x0[e] = ah[q];
      x1[e] = al[q];
      y0[e] = bh[q];
      y1[e] = bl[q];
    }
  };
  fetch(c0);
  for (long i = c0; i < c1; i += 8) {
    double u0[8], u1[8], v0[8], v1[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      u0[e] = x0[e];
      u1[e] = x1[e];
      v0[e] = y0[e];
      v1[e] = y1[e];
    }
    if (i + 8 < c1) fetch(i + 8);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (i + e < c1) s = dd_add(s, dd_mul(mkdd(u0[e], u1[e]), mkdd(v0[e], v1[e])));
  }
  ph[c] = s.hi;
  pl[c] = s.lo;
}

__device__ __forceinline__ double lane_value(double v, int k) {  // v of lane k (k uniform)
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, k);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// ---------------------------------------------------------------- dot: rayon's halving tree
// The reference combines the chunk sums p_i with rayon's reduce(|| ZERO, |acc, x| acc + x);
// the tree taken here splits every range at len / 2 down to single chunks:
//   tree(lo, 1) = ZERO + p_lo,   tree(lo, len) = tree(lo, len/2) + tree(lo + len/2, len - len/2).
// It is laid out as a complete binary tree of depth D = ceil(log2 m): leaf slot g descends
// the D bits of g from (lo, len) = (0, m) (bit 1: lo += len/2, len -= len/2; bit 0: len =
// len/2) and holds chunk lo when one item is left, else nothing.  Each node is then left
// + right, an empty child passing the other through unchanged: the halving tree exactly.
// The complete tree is reduced pairwise, left operand always the lower slot: four slots per
// thread, the wavefront's lanes by DPP row shifts (distances 1, 2, 4, 8), ds_swizzle (16)
// and a lane shuffle (32), the four wavefronts through LDS.  One workgroup takes 1024 slots
// to one node; further passes of the same kernel take those nodes to the root.
constexpr int TREE_SLOTS = 1024;  // per workgroup: 256 threads x 4

__device__ __forceinline__ void tree_comb(dd& a, int& fa, dd b, int fb) {  // a <- a (+) b; a is the left node
  if (fa && fb) a = dd_add(a, b);
  else if (fb) a = b;
  fa = fa | fb;
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i<CTRL>((int)b), hi = (unsigned)dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double swz16_d(double v) {  // lane ^ 16 within 32 lanes (bitmask mode)
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)b, 0x401F);
  const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// after this, lane l with l % 64 == 0 holds the pairwise combination of the wavefront's 64 nodes
__device__ __forceinline__ void wave_tree(dd& v, int& f) {
  // distance d: lane l (l % 2d == 0) takes lane l + d as its right child.  row_shl:d
  // (DPP control 0x100 + d) reads lane l + d within a row of 16.
  tree_comb(v, f, mkdd(dpp_d<0x101>(v.hi), dpp_d<0x101>(v.lo)), dpp_i<0x101>(f));
  tree_comb(v, f, mkdd(dpp_d<0x102>(v.hi), dpp_d<0x102>(v.lo)), dpp_i<0x102>(f));
  tree_comb(v, f, mkdd(dpp_d<0x104>(v.hi), dpp_d<0x104>(v.lo)), dpp_i<0x104>(f));
  tree_comb(v, f, mkdd(dpp_d<0x108>(v.hi), dpp_d<0x108>(v.lo)), dpp_i<0x108>(f));
  tree_comb(v, f, mkdd(swz16_d(v.hi), swz16_d(v.lo)), __builtin_amdgcn_ds_swizzle(f, 0x401F));
  tree_comb(v, f, mkdd(__shfl_down(v.hi, 32, 64), __shfl_down(v.lo, 32, 64)), __shfl_down(f, 32, 64));
}

// LEAF: the slots are the tree's leaves over the m chunk sums in (ph, pl) (depth D);
// otherwise the first `count` slots are the nodes (ph, pl, pf) of the previous pass.
template <bool LEAF>
__global__ __launch_bounds__(256) void k_dot_tree(const double* __restrict__ ph, const double* __restrict__ pl,
                                                  const int* __restrict__ pf, long count, int D,
                                                  double* __restrict__ oh, double* __restrict__ ol,
                                                  int* __restrict__ of) {
  __shared__ double wh[4], wl[4];
  __shared__ int wf[4];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  dd x[4];
  int fx[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const long g = (long)blockIdx.x * TREE_SLOTS + 4 * t + e;
    x[e] = mkdd(0.0);
    fx[e] = 0;
    if (LEAF) {
      if (g < (1L << D)) {
        long lo = 0, len = count;
        for (int bit = D - 1; bit >= 0; --bit) {
          const long half = len >> 1;
          if ((g >> bit) & 1) { lo += half; len -= half; }
          else len = half;
        }
        if (len == 1) {
          x[e] = dd_add(mkdd(0.0), mkdd(ph[lo], pl[lo]));  // the leaf's fold: ZERO + p_lo
          fx[e] = 1;
        }
      }
    } else if (g < count) {
      x[e] = mkdd(ph[g], pl[g]);
      fx[e] = pf[g];
    }
  }
  tree_comb(x[0], fx[0], x[1], fx[1]);
  tree_comb(x[2], fx[2], x[3], fx[3]);
  tree_comb(x[0], fx[0], x[2], fx[2]);
  dd v = x[0];
  int f = fx[0];
  wave_tree(v, f);
  if (lane == 0) {
    wh[w] = v.hi;
    wl[w] = v.lo;
    wf[w] = f;
  }
  __syncthreads();
  if (t == 0) {
    dd a = mkdd(wh[0], wl[0]), b = mkdd(wh[2], wl[2]);
    int fa = wf[0], fb = wf[2];
    tree_comb(a, fa, mkdd(wh[1], wl[1]), wf[1]);
    tree_comb(b, fb, mkdd(wh[3], wl[3]), wf[3]);
    tree_comb(a, fa, b, fb);
    oh[blockIdx.x] = a.hi;
    ol[blockIdx.x] = a.lo;
    if (of) of[blockIdx.x] = fa;
  }
}

// n < 1024: the serial dot of the single chunk (k_dot_chunks' p_0), ZERO for n = 0
__global__ void k_dot_small(const double* ph, const double* pl, long n, double* oh, double* ol) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    oh[0] = n ? ph[0] : 0.0;
    ol[0] = n ? pl[0] : 0.0;
  }
}

// y_i = sum_j a_ij x_j in j order.  A block takes 64 rows: wavefronts 1-3 stage 64 x 32
// tiles of A (and x) into double-buffered LDS with coalesced loads, one tile ahead, while
// wavefront 0 runs the 64 serial row chains over the current tile.
constexpr int GV_T = 32;
__global__ __launch_bounds__(256) void k_gemv(const double* __restrict__ ah, const double* __restrict__ al, long rows,
                                              long cols, const double* __restrict__ xh, const double* __restrict__ xl,
                                              double* __restrict__ yh, double* __restrict__ yl) {
  __shared__ double th[2][64][GV_T + 1], tl[2][64][GV_T + 1];
  __shared__ double sxh[2][GV_T], sxl[2][GV_T];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const long r0 = (long)blockIdx.x * 64;
  const long ntile = (cols + GV_T - 1) / GV_T;
  auto load = [&](long tile, int buf) {  // wavefronts 1-3: every load issued before the LDS stores
    const long c0 = tile * GV_T;
    constexpr int NL = (64 * GV_T + 191) / 192;
    double vh[NL], vl[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = t - 64 + 192 * i;
      const long gr = r0 + e / GV_T, gc = c0 + e % GV_T;
      const bool ok = e < 64 * GV_T && gr < rows && gc < cols;
      vh[i] = ok ? ah[gr * cols + gc] : 0.0;
      vl[i] = ok ? al[gr * cols + gc] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = t - 64 + 192 * i;
      if (e < 64 * GV_T) {
        th[buf][e / GV_T][e % GV_T] = vh[i];
        tl[buf][e / GV_T][e % GV_T] = vl[i];
      }
    }
    for (int e = t - 64; e < GV_T; e += 192) {
      const long gc = c0 + e;
      sxh[buf][e] = gc < cols ? xh[gc] : 0.0;
      sxl[buf][e] = gc < cols ? xl[gc] : 0.0;
    }
  };
  if (w > 0 && ntile > 0) load(0, 0);
  __syncthreads();
  dd s = mkdd(0.0);
  for (long tile = 0; tile < ntile; ++tile) {
    const int buf = (int)(tile & 1);
    if (w > 0) {
      if (tile + 1 < ntile) load(tile + 1, buf ^ 1);
    } else {
      const long rem = cols - tile * GV_T;
      const int wc = rem < GV_T ? (int)rem : GV_T;
#pragma unroll
      for (int k = 0; k < GV_T; ++k)  // unrolled: the LDS reads go ahead of the chain
        if (k < wc) s = dd_add(s, dd_mul(mkdd(th[buf][lane][k], tl[buf][lane][k]), mkdd(sxh[buf][k], sxl[buf][k])));
    }
    __syncthreads();  // tile read by wavefront 0 and the next one written before either is reused
  }
  if (w == 0 && r0 + lane < rows) {
    yh[r0 + lane] = s.hi;
    yl[r0 + lane] = s.lo;
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
// chunk c of NEU_CHUNK values -> its Neumaier sum, one thread per chunk, 8 loads ahead
__global__ __launch_bounds__(64) void k_neu_f64(const double* __restrict__ v, long n, double* __restrict__ part) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * NEU_CHUNK;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + NEU_CHUNK);
  double sum = 0.0, comp = 0.0;
  for (long i = c0; i < c1; i += 8) {
    double x[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = i + e < c1 ? v[i + e] : 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (i + e < c1) neu_f64(sum, comp, x[e]);
  }
  part[c] = sum + comp;
}
__global__ __launch_bounds__(64) void k_neu_dd(const double* __restrict__ vh, const double* __restrict__ vl, long n,
                                               double* __restrict__ ph, double* __restrict__ pl) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long c0 = c * NEU_CHUNK;
  if (c0 >= n) return;
  const long c1 = min(n, c0 + NEU_CHUNK);
  dd sum = mkdd(0.0), comp = mkdd(0.0);
  for (long i = c0; i < c1; i += 8) {
    double x0[8], x1[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      x0[e] = i + e < c1 ? vh[i + e] : 0.0;
      x1[e] = i + e < c1 ? vl[i + e] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (i + e < c1) neu_dd(sum, comp, mkdd(x0[e], x1[e]));
  }
  const dd r = dd_add(sum, comp);
  ph[c] = r.hi;
  pl[c] = r.lo;
}
// the whole Neumaier sum of n values in order (the serial forms, and the sum of the
// partials): one wavefront loads 64 values at a time, the next 64 while these are summed,
// and runs the sum, the same in every lane, on values read across lanes
__global__ __launch_bounds__(64) void k_neu_f64_wave(const double* __restrict__ v, long n, double* __restrict__ out) {
  const int lane = threadIdx.x;
  double sum = 0.0, comp = 0.0;
  double nv = lane < n ? v[lane] : 0.0;
  for (long c0 = 0; c0 < n; c0 += 64) {
    const double x = nv;
    const long nx = c0 + 64 + lane;
    nv = nx < n ? v[nx] : 0.0;
    const long cnt = n - c0 < 64 ? n - c0 : 64;
#pragma unroll
    for (int k = 0; k < 64; ++k)
      if (k < cnt) neu_f64(sum, comp, lane_value(x, k));
  }
  if (lane == 0) out[0] = sum + comp;
}
__global__ __launch_bounds__(64) void k_neu_dd_wave(const double* __restrict__ vh, const double* __restrict__ vl,
                                                    long n, double* __restrict__ oh, double* __restrict__ ol) {
  const int lane = threadIdx.x;
  dd sum = mkdd(0.0), comp = mkdd(0.0);
  double nh = lane < n ? vh[lane] : 0.0, nl = lane < n ? vl[lane] : 0.0;
  for (long c0 = 0; c0 < n; c0 += 64) {
    const double h = nh, l = nl;
    const long nx = c0 + 64 + lane;
    nh = nx < n ? vh[nx] : 0.0;
    nl = nx < n ? vl[nx] : 0.0;
    const long cnt = n - c0 < 64 ? n - c0 : 64;
#pragma unroll
    for (int k = 0; k < 64; ++k)
      if (k < cnt) neu_dd(sum, comp, mkdd(lane_value(h, k), lane_value(l, k)));
  }
  if (lane == 0) {
    const dd r = dd_add(sum, comp);
    oh[0] = r.hi;
    ol[0] = r.lo;
  }
}

unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// scratch of dot_dev: the m chunk sums (2 m doubles), then the tree passes' nodes (values
// and flags, at most 2 (1 << D) / 1024 over all passes); dot_scratch_bytes(n) bytes
constexpr long dot_tree_nodes(long m) {
  int D = 0;
  while ((1L << D) < m) ++D;
  return 2 * ((1L << D) / TREE_SLOTS > 1 ? (1L << D) / TREE_SLOTS : 1);
}
constexpr size_t dot_scratch_bytes(long n) {
  const long m = (n + DOT_CHUNK - 1) / DOT_CHUNK > 1 ? (n + DOT_CHUNK - 1) / DOT_CHUNK : 1;
  const long nodes = dot_tree_nodes(m);
  return (size_t)(2 * m + 2 * nodes) * sizeof(double) + (size_t)nodes * sizeof(int);
}
int dot_dev_scratch(const double* ah, const double* al, const double* bh, const double* bl, long n, double* out_h,
                    double* out_l, void* scr, hipStream_t s) {
  const long m = std::max(1L, (n + DOT_CHUNK - 1) / DOT_CHUNK);
  double* ph = static_cast<double*>(scr);
  double* pl = ph + m;
  if (n > 0) {
    hipLaunchKernelGGL(k_dot_chunks, dim3(nblk(m, 64)), dim3(64), 0, s, ah, al, bh, bl, n, ph, pl);
    HP_LAUNCH_CHECK();
  }
  if (n < DOT_CHUNK) {
    hipLaunchKernelGGL(k_dot_small, dim3(1), dim3(64), 0, s, ph, pl, n, out_h, out_l);
    HP_LAUNCH_CHECK();
    return HP_OK;
  }
  int D = 0;
  while ((1L << D) < m) ++D;
  const long cap = dot_tree_nodes(m);
  double* nh = pl + m;  // node values and flags of all passes, pass after pass
  double* nl = nh + cap;
  int* nf = reinterpret_cast<int*>(nl + cap);
  long count = 1L << D, off = 0;  // slots of this pass; where its nodes go
  const double *inh = ph, *inl = pl;
  const int* inf = nullptr;
  long in_count = m;
  for (bool leaf = true;; leaf = false) {
    const long grid = (count + TREE_SLOTS - 1) / TREE_SLOTS;
    const bool last = grid == 1;
    double* dh = last ? out_h : nh + off;
    double* dl = last ? out_l : nl + off;
    int* df = last ? nullptr : nf + off;
    if (leaf)
      hipLaunchKernelGGL(k_dot_tree<true>, dim3((unsigned)grid), dim3(256), 0, s, inh, inl, inf, in_count, D, dh, dl, df);
    else
      hipLaunchKernelGGL(k_dot_tree<false>, dim3((unsigned)grid), dim3(256), 0, s, inh, inl, inf, in_count, D, dh,
                         dl, df);
    HP_LAUNCH_CHECK();
    if (last) break;
    inh = dh;
    inl = dl;
    inf = df;
    in_count = count = grid;
    off += grid;
  }
  return HP_OK;
}
int dot_dev(const double* ah, const double* al, const double* bh, const double* bl, long n, double* out_h,
            double* out_l, hipStream_t s) {
  DevBuf scr;
  int rc;
  if ((rc = scr.alloc(dot_scratch_bytes(n)))) return rc;
  if ((rc = dot_dev_scratch(ah, al, bh, bl, n, out_h, out_l, scr.p, s))) return rc;
  HP_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return HP_OK;
}

// ---------------------------------------------------------------- allocation-free dot
// hp_dev_dot_noalloc runs dot_dev's kernels with their scratch (the chunk sums and the tree
// nodes) in `scratch`, a file-scope device array sized for n <= NOALLOC_DOT_MAX_N, so the
// call allocates nothing.  In host code `scratch` names a host-side stand-in: its device
// address, per device, comes from hipGetSymbolAddress on the first call on that device and
// is the only form passed to kernels.  Nothing in it must start at zero (every chunk sum and
// node is written before it is read).  All calls on one device share the array: calls that
// overlap in time on one device (two streams, two threads) race; calls in one stream are
// ordered.
constexpr long NOALLOC_DOT_MAX_N = 1L << 26;
constexpr size_t NOALLOC_DOT_SCRATCH = dot_scratch_bytes(NOALLOC_DOT_MAX_N);  // ~1.0 MiB
static __device__ __align__(256) unsigned char scratch[NOALLOC_DOT_SCRATCH];

std::mutex g_scratch_mu;
void* g_scratch_dev[64] = {};  // device address of `scratch` per device ordinal (under g_scratch_mu)

int noalloc_scratch(void** out) {
  int dev = 0;
  HP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("hp_dev_dot_noalloc: device ordinal %d above 63", dev);
    return HP_EDEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return HP_OK;
}

int gemv_dev(const double* ah, const double* al, long rows, long cols, const double* xh, const double* xl, double* yh,
             double* yl, hipStream_t s) {
  if (rows == 0) return HP_OK;
  hipLaunchKernelGGL(k_gemv, dim3(nblk(rows, 64)), dim3(256), 0, s, ah, al, rows, cols, xh, xl, yh, yl);
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
  const long np = chunked ? ((long)n + NEU_CHUNK - 1) / NEU_CHUNK : 0;
  DevBuf dv, part, fin;
  int rc;
  if ((rc = dv.alloc(limbs * n * sizeof(double))) || (rc = part.alloc(limbs * std::max(np, 1L) * sizeof(double))) ||
      (rc = fin.alloc(limbs * sizeof(double))))
    return rc;
  if ((rc = upload_planes(v, n, limbs, dv.d(), s))) return rc;
  if (limbs == 1) {
    if (chunked) {  // the chunk sums, then the Neumaier sum of the partials in order
      hipLaunchKernelGGL(k_neu_f64, dim3(nblk(np, 64)), dim3(64), 0, s, dv.d(), (long)n, part.d());
      HP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_neu_f64_wave, dim3(1), dim3(64), 0, s, part.d(), np, fin.d());
    } else {
      hipLaunchKernelGGL(k_neu_f64_wave, dim3(1), dim3(64), 0, s, dv.d(), (long)n, fin.d());
    }
    HP_LAUNCH_CHECK();
    HP_HIP(hipMemcpyAsync(out, fin.p, sizeof(double), hipMemcpyDeviceToHost, s));
  } else {
    if (chunked) {
      hipLaunchKernelGGL(k_neu_dd, dim3(nblk(np, 64)), dim3(64), 0, s, dv.d(), dv.d() + n, (long)n, part.d(),
                         part.d() + np);
      HP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_neu_dd_wave, dim3(1), dim3(64), 0, s, part.d(), part.d() + np, np, fin.d(), fin.d() + 1);
    } else {
      hipLaunchKernelGGL(k_neu_dd_wave, dim3(1), dim3(64), 0, s, dv.d(), dv.d() + n, (long)n, fin.d(), fin.d() + 1);
    }
    HP_LAUNCH_CHECK();
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

int hp_dev_dot_noalloc(const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo, size_t n,
                       double* out_hi, double* out_lo, void* stream) {
  if (n > (size_t)NOALLOC_DOT_MAX_N) {
    set_error("hp_dev_dot_noalloc: n = %zu above the scratch's maximum %ld (use hp_dev_dot)", n, NOALLOC_DOT_MAX_N);
    return HP_EINVAL;
  }
  if (!out_hi || !out_lo || (n && (!a_hi || !a_lo || !b_hi || !b_lo))) {
    set_error("hp_dev_dot_noalloc: null pointer");
    return HP_EINVAL;
  }
  void* scr = nullptr;
  int rc = noalloc_scratch(&scr);
  if (rc) return rc;
  return dot_dev_scratch(a_hi, a_lo, b_hi, b_lo, (long)n, out_hi, out_lo, scr, (hipStream_t)stream);
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
