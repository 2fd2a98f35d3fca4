// k_adjoint.hip -- adjoints of the overlap build and its Cholesky factor.
//
//   F1 = log det S = 2 sum_i ln L_ii,      F2 = b^T S^-1 b = b^T x,      S = L L^T,
// S the normalised s-type overlap matrix (hp_overlap.hpp), L its dd cholesky_blocked
// factor (crates/quad/src/linalg.rs:238-293) or the split-dd factor, x refined with
// quad-double residuals.  The reference has no differentiation; what is differentiated is
// its overlap entry and its factor.  Gradients over the 4n parameters (3n centre
// coordinates, Gaussian-major, then n exponents): S_uu = 1 for every parameter value and
// S_uv depends on Gaussian u only through row / column u, so with G = dF/dS symmetric
//   dF/dtheta_u = 2 sum_{v != u} G_uv dS_uv/dtheta_u,      G = S^-1 (F1),  G = -x x^T (F2)
//   dS_uv/dR_uc = -2 p (R_uc - R_vc) S_uv,                 p = a_u a_v / g,  g = a_u + a_v
//   dS_uv/da_u  = S_uv (3/4 (1/a_u - 2/g) - |R_u - R_v|^2 a_v^2 / g^2).
//
// Forward keeps the factor L and x (what a torch.autograd.Function would save).
// Backward recomputes, never stores, the rest:
//   X = L^-1   inverses of the 64 x 64 diagonal blocks, then recursive doubling: at block
//              size m every pair of neighbouring blocks is joined by T = L21 X11,
//              X21 = -X22 T (all pairs of a level in one batched launch on the plain dd
//              GEMM; the large levels optionally on the INT8 split GEMM)
//   W = X^T X  = S^-1, lower triangle (plain: tiles above the diagonal skipped and each
//              tile's sum started at its first row, where X's zeros end), then mirrored
//   S_uv       recomputed per pair with the same dd operations as the build
// and contracts row u of W and x against the pair factors in one workgroup per row.
// Work at n: X about n^3/3 dd MACs, W n^3/6 (plain) or n^3/2 (split, lower tiles), the
// contraction n^2 overlap entries (~1.6k FP64 operations each).
#include <chrono>
#include <cstring>

#include "hp_chol.hpp"
#include "hp_internal.hpp"
#include "hp_overlap.hpp"

namespace hp {

template <class A>
int chol_exact(Mat a, long n, long nb, bool unblocked, Mat S, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
int chol_split(Mat a, long n, long nb, Mat S, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
int solve_with_dd_factor(Mat Lm, const double* sh, const double* sl, const double* bh, const double* bl, long n,
                         double tol, int max_iter, double* xh, double* xl, int* iters, double* last, hipStream_t s);
int diag_block_inverses_dd(Mat l, long n, double* ih, double* il, hipStream_t s);
int launch_overlap_s(const double* centers, const double* alpha, size_t n, double* shi, double* slo, hipStream_t s);

int g_adj_split_min = 512;  // smallest block size whose joining GEMMs use the split GEMM

namespace {

constexpr int BW = 64;

// X(b0 + i, b0 + c) = inv(L_bb)[i][c], c <= i, from the compact [block][i][c] layout
__global__ __launch_bounds__(256) void k_scatter_diag_inv(const double* __restrict__ ih, const double* __restrict__ il,
                                                          double* __restrict__ xh, double* __restrict__ xl, long n) {
  const long blk = blockIdx.x, b0 = blk * BW;
  const int bw = (int)min((long)BW, n - b0);
  for (int e = threadIdx.x; e < BW * BW; e += 256) {
    const int i = e / BW, c = e % BW;
    if (i < bw && c <= i) {
      const size_t src = (size_t)blk * BW * BW + e, dst = (size_t)(b0 + i) * n + b0 + c;
      xh[dst] = ih[src];
      xl[dst] = il[src];
    }
  }
}

// upper triangle of W from its lower triangle, 32 x 32 tiles transposed through LDS.  Source
// tiles are on or below the diagonal and only strictly-upper elements are written.
__global__ __launch_bounds__(256) void k_mirror_lower(double* __restrict__ wh, double* __restrict__ wl, long n) {
  __shared__ double th[32][33], tl[32][33];
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bj > bi) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const long i = (long)bi * 32 + r, j = (long)bj * 32 + tx;
    if (i < n && j < n) {
      th[r][tx] = wh[i * n + j];
      tl[r][tx] = wl[i * n + j];
    }
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const long i = (long)bj * 32 + r, j = (long)bi * 32 + tx;  // W(i, j) = W(j, i) = tile[tx][r]
    if (i < n && j < n && j > i) {
      wh[i * n + j] = th[tx][r];
      wl[i * n + j] = tl[tx][r];
    }
  }
}

// One 256-thread workgroup per row u.  Per pair (u, v): S_uv with the build's dd operations,
// the factors t_c = p (R_uc - R_vc) S_uv and t_3 = dS_uv/da_u, accumulated against W_uv
// (when W is given) and x_v; the 8 dd sums are reduced through LDS and scaled:
//   g1 = dF1: centres -4 sum W t_c, exponent 2 sum W t_3
//   g2 = dF2: centres 4 x_u sum x t_c, exponent -2 x_u sum x t_3
__global__ __launch_bounds__(256) void k_adj_contract(const double* __restrict__ centers,
                                                      const double* __restrict__ alpha, const double* __restrict__ wh,
                                                      const double* __restrict__ wl, const double* __restrict__ xh,
                                                      const double* __restrict__ xl, long n, double* __restrict__ g1h,
                                                      double* __restrict__ g1l, double* __restrict__ g2h,
                                                      double* __restrict__ g2l) {
  __shared__ double red[16][256];
  const long u = blockIdx.x;
  const int t = threadIdx.x;
  const bool use_w = wh != nullptr;
  const double au = alpha[u];
  const double cu[3] = {centers[3 * u], centers[3 * u + 1], centers[3 * u + 2]};
  const dd a = mkdd(au);
  const dd inv_a = dd_recip(a);
  dd s1[4], s2[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) s1[c] = s2[c] = mkdd(0.0);
  for (long v = t; v < n; v += 256) {
    if (v == u) continue;
    const double av = alpha[v];
    const double cv[3] = {centers[3 * v], centers[3 * v + 1], centers[3 * v + 2]};
    // S_uv exactly as overlap_s_entry, sharing g, ab, p, d, r2 with the factors
    const dd b = mkdd(av);
    const dd g = dd_add(a, b);
    const dd ab = dd_mul(a, b);
    const dd p = dd_div(ab, g);
    dd d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) two_sum(cu[c], -cv[c], d[c].hi, d[c].lo);
    const dd r2 = dd_add(dd_add(dd_mul(d[0], d[0]), dd_mul(d[1], d[1])), dd_mul(d[2], d[2]));
    const dd arg = dd_neg(dd_mul(p, r2));
    if (arg.hi < -708.0) continue;  // S_uv = 0: every term of this pair is an exact zero
    const dd ratio = dd_div(dd_mul_f(ab, 4.0), dd_mul(g, g));
    const dd tq = dd_sqrt(ratio);
    const dd pref = dd_mul(tq, dd_sqrt(tq));
    const dd s = dd_mul(pref, dd_exp(arg));
    const dd ps = dd_mul(p, s);
    dd tt[4];
#pragma unroll
    for (int c = 0; c < 3; ++c) tt[c] = dd_mul(ps, d[c]);
    const dd bg = dd_div(b, g);
    const dd fa = dd_sub(dd_mul_f(dd_sub(inv_a, dd_div(mkdd(2.0), g)), 0.75), dd_mul(r2, dd_mul(bg, bg)));
    tt[3] = dd_mul(s, fa);
    if (use_w) {
      const dd w = mkdd(wh[u * n + v], wl[u * n + v]);
#pragma unroll
      for (int c = 0; c < 4; ++c) s1[c] = dd_add(s1[c], dd_mul(w, tt[c]));
    }
    const dd xv = mkdd(xh[v], xl[v]);
#pragma unroll
    for (int c = 0; c < 4; ++c) s2[c] = dd_add(s2[c], dd_mul(xv, tt[c]));
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    red[2 * c][t] = s1[c].hi;
    red[2 * c + 1][t] = s1[c].lo;
    red[8 + 2 * c][t] = s2[c].hi;
    red[8 + 2 * c + 1][t] = s2[c].lo;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        dd x = dd_add(mkdd(red[2 * q][t], red[2 * q + 1][t]), mkdd(red[2 * q][t + w], red[2 * q + 1][t + w]));
        red[2 * q][t] = x.hi;
        red[2 * q + 1][t] = x.lo;
      }
    }
    __syncthreads();
  }
  if (t != 0) return;
  const dd xu = mkdd(xh[u], xl[u]);
  for (int c = 0; c < 4; ++c) {
    const dd r1 = mkdd(red[2 * c][0], red[2 * c + 1][0]);
    const dd r2v = mkdd(red[8 + 2 * c][0], red[8 + 2 * c + 1][0]);
    const long o = c < 3 ? 3 * u + c : 3 * n + u;
    if (use_w) {
      const dd v1 = dd_mul_f(r1, c < 3 ? -4.0 : 2.0);
      g1h[o] = v1.hi;
      g1l[o] = v1.lo;
    }
    if (g2h) {
      const dd v2 = dd_mul(dd_mul_f(r2v, c < 3 ? 4.0 : -2.0), xu);
      g2h[o] = v2.hi;
      g2l[o] = v2.lo;
    }
  }
}

// log det S = sum_i 2 ln L_ii and b^T x, one workgroup, tree-reduced in a fixed order
__global__ __launch_bounds__(1024) void k_logdet_quad(const double* __restrict__ lh, const double* __restrict__ ll,
                                                      const double* __restrict__ bh, const double* __restrict__ bl,
                                                      const double* __restrict__ xh, const double* __restrict__ xl,
                                                      long n, double* __restrict__ out) {
  __shared__ double red[4][1024];
  const int t = threadIdx.x;
  dd s1 = mkdd(0.0), s2 = mkdd(0.0);
  for (long i = t; i < n; i += 1024) {
    s1 = dd_add(s1, dd_mul_f(dd_ln(mkdd(lh[i * n + i], ll[i * n + i])), 2.0));
    s2 = dd_add(s2, dd_mul(mkdd(bh[i], bl[i]), mkdd(xh[i], xl[i])));
  }
  red[0][t] = s1.hi;
  red[1][t] = s1.lo;
  red[2][t] = s2.hi;
  red[3][t] = s2.lo;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      dd a1 = dd_add(mkdd(red[0][t], red[1][t]), mkdd(red[0][t + w], red[1][t + w]));
      dd a2 = dd_add(mkdd(red[2][t], red[3][t]), mkdd(red[2][t + w], red[3][t + w]));
      red[0][t] = a1.hi;
      red[1][t] = a1.lo;
      red[2][t] = a2.hi;
      red[3][t] = a2.lo;
    }
    __syncthreads();
  }
  if (t < 4) out[t] = red[t][0];
}

// vector-Jacobian product: out = c1 g1 + c2 g2 over the 4n parameters (g1 may be absent),
// scattered into centre (3n) and exponent (n) gradients; gb = 2 c2 x (when requested)
__global__ __launch_bounds__(256) void k_adj_combine(const double* __restrict__ g1h, const double* __restrict__ g1l,
                                                     const double* __restrict__ g2h, const double* __restrict__ g2l,
                                                     dd c1, dd c2, long n, double* __restrict__ gch,
                                                     double* __restrict__ gcl, double* __restrict__ gah,
                                                     double* __restrict__ gal, const double* __restrict__ xh,
                                                     const double* __restrict__ xl, double* __restrict__ gbh,
                                                     double* __restrict__ gbl) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < 4 * n) {
    dd v = dd_mul(c2, mkdd(g2h[i], g2l[i]));
    if (g1h) v = dd_add(dd_mul(c1, mkdd(g1h[i], g1l[i])), v);
    if (i < 3 * n) {
      gch[i] = v.hi;
      gcl[i] = v.lo;
    } else {
      gah[i - 3 * n] = v.hi;
      gal[i - 3 * n] = v.lo;
    }
  }
  if (gbh && i < n) {
    dd v = dd_mul_f(dd_mul(c2, mkdd(xh[i], xl[i])), 2.0);
    gbh[i] = v.hi;
    gbl[i] = v.lo;
  }
}

unsigned nblocks(long count, int per) { return (unsigned)std::max<long>(1, (count + per - 1) / per); }

// join `batch` neighbouring block pairs of a level on the plain GEMM: pair z has its first
// block at r0 + 2 m (n + 1) z (size m) and its second at + m (size m2)
int join_plain(Mat L, double* xh, double* xl, double* th, double* tl, long n, long r0, long m, long m2, int batch,
               hipStream_t s) {
  const long st = 2 * m * (n + 1);
  const long oL21 = (r0 + m) * n + r0, oX11 = r0 * n + r0, oX22 = (r0 + m) * n + r0 + m;
  int rc = launch_gemm_dd_plain_ex(0, L.hi + oL21, L.lo + oL21, n, xh + oX11, xl + oX11, n, th, tl, m, m2, m, m,
                                   false, false, batch, st, st, m * m, s);
  if (rc) return rc;
  return launch_gemm_dd_plain_ex(0, xh + oX22, xl + oX22, n, th, tl, m, xh + oL21, xl + oL21, n, m2, m, m2, true,
                                 false, batch, st, m * m, st, s);
}

// the same join for one pair on the INT8 split GEMM (X21 is zero, so C -= X22 T gives -X22 T)
int join_split(Mat L, double* xh, double* xl, double* th, double* tl, long n, long r0, long m, long m2,
               hipStream_t s) {
  const long oL21 = (r0 + m) * n + r0, oX11 = r0 * n + r0, oX22 = (r0 + m) * n + r0 + m;
  SplitOperand A{L.hi + oL21, L.lo + oL21, n, true};
  SplitOperand B{xh + oX11, xl + oX11, n, false};  // B^T(j, p) = X11(p, j)
  int rc = split_gemm_general(A, B, th, tl, (size_t)m, (size_t)m2, (size_t)m, (size_t)m, 0, s, SplitOpts{});
  if (rc) return rc;
  SplitOperand A2{xh + oX22, xl + oX22, n, true};
  SplitOperand B2{th, tl, m, false};  // B^T(j, p) = T(p, j)
  SplitOpts o;
  o.subtract = true;
  return split_gemm_general(A2, B2, xh + oL21, xl + oL21, (size_t)n, (size_t)m2, (size_t)m, (size_t)m2, 0, s, o);
}

// X = L^-1 (lower triangular, zero above the diagonal)
int tri_inverse(Mat L, long n, double* xh, double* xl, double* th, double* tl, int gemm_mode, hipStream_t s) {
  const size_t nn = (size_t)n * n;
  HP_HIP(hipMemsetAsync(xh, 0, nn * sizeof(double), s));
  HP_HIP(hipMemsetAsync(xl, 0, nn * sizeof(double), s));
  const long nblk = (n + BW - 1) / BW;
  DevBuf inv;
  int rc;
  if ((rc = inv.alloc((size_t)nblk * BW * BW * 2 * sizeof(double)))) return rc;
  double* ih = inv.d();
  double* il = ih + (size_t)nblk * BW * BW;
  if ((rc = diag_block_inverses_dd(L, n, ih, il, s))) return rc;
  hipLaunchKernelGGL(k_scatter_diag_inv, dim3((unsigned)nblk), dim3(256), 0, s, ih, il, xh, xl, n);
  HP_LAUNCH_CHECK();
  for (long m = BW; m < n; m *= 2) {
    const long full = n / (2 * m);
    const long r0r = full * 2 * m, m2r = n - r0r - m;
    const bool split = gemm_mode == HP_GEMM_SPLIT_I8 && m >= g_adj_split_min;
    if (full > 0) {
      if (!split) {
        if ((rc = join_plain(L, xh, xl, th, tl, n, 0, m, m, (int)full, s))) return rc;
      } else {
        for (long z = 0; z < full; ++z)
          if ((rc = join_split(L, xh, xl, th, tl, n, 2 * m * z, m, m, s))) return rc;
      }
    }
    if (m2r > 0) {
      rc = split ? join_split(L, xh, xl, th, tl, n, r0r, m, m2r, s) : join_plain(L, xh, xl, th, tl, n, r0r, m, m2r, 1, s);
      if (rc) return rc;
    }
  }
  HP_HIP(hipStreamSynchronize(s));  // inv is freed on return
  return HP_OK;
}

}  // namespace
}  // namespace hp

using namespace hp;

// Saved state of one forward evaluation: inputs, the factor L and x = S^-1 b, plus the
// backward's work buffers (allocated on first use, reused by later backward calls).
struct hp_adjoint {
  long n = 0;
  int factor_kind = 0;
  DevBuf centers, alpha, b, L, x;
  DevBuf X, W, T, G;  // backward: L^-1, S^-1, join products, the two 4n gradients
  hp_adjoint_info info{};
  hipEvent_t ev[4] = {};  // backward phase boundaries (phase timing on)
  ~hp_adjoint() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace hp {

int adjoint_forward(const double* centers, const double* alpha, const double* bh, const double* bl, size_t n_,
                    int factor_kind, hp_adjoint** out, hp_adjoint_info* info, hipStream_t s) {
  using clk = std::chrono::steady_clock;
  *out = nullptr;
  if (factor_kind != HP_FACTOR_DD && factor_kind != HP_FACTOR_DD_SPLIT) {
    set_error("overlap adjoint: factor kind %d is not a dd factor", factor_kind);
    return HP_EINVAL;
  }
  if (n_ == 0 || n_ > 65535) {
    set_error("overlap adjoint: n = %zu outside [1, 65535]", n_);
    return HP_ESHAPE;
  }
  const long n = (long)n_;
  const size_t nn = (size_t)n * n;
  auto t0 = clk::now();
  hp_adjoint* c = new hp_adjoint;
  c->n = n;
  c->factor_kind = factor_kind;
  DevBuf S, acc, err, inv, vals;
  int rc;
  if ((rc = c->centers.alloc(3 * n * sizeof(double))) || (rc = c->alpha.alloc(n * sizeof(double))) ||
      (rc = c->b.alloc(2 * n * sizeof(double))) || (rc = c->L.alloc(2 * nn * sizeof(double))) ||
      (rc = c->x.alloc(2 * n * sizeof(double))) || (rc = S.alloc(2 * nn * sizeof(double))) ||
      (rc = acc.alloc(2 * nn * sizeof(double))) || (rc = err.alloc(sizeof(ErrRec))) ||
      (rc = inv.alloc(2 * 64 * sizeof(double))) || (rc = vals.alloc(4 * sizeof(double)))) {
    delete c;
    return rc;
  }
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(s);
    delete c;
    return code;
  };
#define ADJ_HIP(call)                                                                                      \
  do {                                                                                                     \
    hipError_t _e = (call);                                                                                \
    if (_e != hipSuccess) {                                                                                \
      set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__);                \
      return fail(HP_EDEVICE);                                                                             \
    }                                                                                                      \
  } while (0)
  ADJ_HIP(hipMemcpyAsync(c->centers.p, centers, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  ADJ_HIP(hipMemcpyAsync(c->alpha.p, alpha, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  ADJ_HIP(hipMemcpyAsync(c->b.d(), bh, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  ADJ_HIP(hipMemcpyAsync(c->b.d() + n, bl, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  double* shi = S.d();
  double* slo = S.d() + nn;
  if ((rc = launch_overlap_s(c->centers.d(), c->alpha.d(), n, shi, slo, s))) return fail(rc);
  Mat Lm{c->L.d(), c->L.d() + nn, n};
  Mat Am{acc.d(), acc.d() + nn, n};
  ADJ_HIP(hipMemcpyAsync(Lm.hi, shi, 2 * nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  ADJ_HIP(hipMemsetAsync(err.p, 0, sizeof(ErrRec), s));
  ErrRec* er = (ErrRec*)err.p;
  rc = factor_kind == HP_FACTOR_DD ? chol_exact<ArDD>(Lm, n, 64, false, Am, er, inv.d(), inv.d() + 64, s)
                                   : chol_split(Lm, n, g_split_panel.load(), Am, er, inv.d(), inv.d() + 64, s);
  if (rc) return fail(rc);
  ErrRec e;
  ADJ_HIP(hipMemcpyAsync(&e, er, sizeof(e), hipMemcpyDeviceToHost, s));
  ADJ_HIP(hipStreamSynchronize(s));
  std::memset(info, 0, sizeof(*info));
  if (e.flag) {
    set_error("Cholesky failed at pivot %lld (value=%.4e)", e.pivot, e.value);
    info->pivot = (size_t)e.pivot;
    return fail(HP_ENOTSPD);
  }
  int it = 0;
  double last = 0;
  if ((rc = solve_with_dd_factor(Lm, shi, slo, c->b.d(), c->b.d() + n, n, 1e-32, 6, c->x.d(), c->x.d() + n, &it,
                                 &last, s)))
    return fail(rc);
  hipLaunchKernelGGL(k_logdet_quad, dim3(1), dim3(1024), 0, s, Lm.hi, Lm.lo, c->b.d(), c->b.d() + n, c->x.d(),
                     c->x.d() + n, n, vals.d());
  if (hipGetLastError() != hipSuccess) {
    set_error("overlap adjoint: launch failed");
    return fail(HP_EDEVICE);
  }
  double v[4];
  ADJ_HIP(hipMemcpyAsync(v, vals.p, sizeof(v), hipMemcpyDeviceToHost, s));
  ADJ_HIP(hipStreamSynchronize(s));
#undef ADJ_HIP
  info->logdet[0] = v[0];
  info->logdet[1] = v[1];
  info->quad[0] = v[2];
  info->quad[1] = v[3];
  info->refine_iters = it;
  info->last_correction = last;
  info->forward_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  c->info = *info;
  *out = c;
  return HP_OK;
}

// both gradients (4n dd each; g1 = d log det S, g2 = d b^T S^-1 b), g1 skipped when g1h is null
int adjoint_gradients(hp_adjoint* c, int gemm_mode, double* g1h, double* g1l, double* g2h, double* g2l,
                      hipStream_t s) {
  if (gemm_mode != HP_GEMM_PLAIN && gemm_mode != HP_GEMM_SPLIT_I8) {
    set_error("overlap adjoint: unknown gemm mode %d", gemm_mode);
    return HP_EINVAL;
  }
  const long n = c->n;
  const size_t nn = (size_t)n * n;
  const bool timing = phase_timer().on;
  if (timing)
    for (hipEvent_t& e : c->ev)
      if (!e) HP_HIP(hipEventCreate(&e));
  if (timing) HP_HIP(hipEventRecord(c->ev[0], s));
  int rc;
  double *wh = nullptr, *wl = nullptr;
  if (g1h) {
    if (!c->X.p && ((rc = c->X.alloc(2 * nn * sizeof(double))) || (rc = c->W.alloc(2 * nn * sizeof(double))) ||
                    (rc = c->T.alloc(2 * (nn / 2 + 64 * (size_t)n) * sizeof(double)))))
      return rc;
    const size_t tsz = nn / 2 + 64 * (size_t)n;
    Mat Lm{c->L.d(), c->L.d() + nn, n};
    double *xh = c->X.d(), *xl = c->X.d() + nn;
    wh = c->W.d();
    wl = c->W.d() + nn;
    if ((rc = tri_inverse(Lm, n, xh, xl, c->T.d(), c->T.d() + tsz, gemm_mode, s))) return rc;
    if (timing) HP_HIP(hipEventRecord(c->ev[1], s));
    if (gemm_mode == HP_GEMM_PLAIN) {
      rc = launch_gemm_dd_plain_ex(1, xh, xl, n, xh, xl, n, wh, wl, n, n, n, n, false, true, 1, 0, 0, 0, s);
    } else {
      SplitOperand A{xh, xl, n, false};  // A(i, p) = X(p, i)
      SplitOpts o;
      o.lower_only = true;
      o.same_operand = true;
      rc = split_gemm_general(A, A, wh, wl, (size_t)n, (size_t)n, (size_t)n, (size_t)n, 0, s, o);
    }
    if (rc) return rc;
    const unsigned t32 = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(k_mirror_lower, dim3(t32, t32), dim3(256), 0, s, wh, wl, n);
    HP_LAUNCH_CHECK();
  } else if (timing) {
    HP_HIP(hipEventRecord(c->ev[1], s));
  }
  if (timing) HP_HIP(hipEventRecord(c->ev[2], s));
  hipLaunchKernelGGL(k_adj_contract, dim3((unsigned)n), dim3(256), 0, s, c->centers.d(), c->alpha.d(), wh, wl,
                     c->x.d(), c->x.d() + n, n, g1h, g1l, g2h, g2l);
  HP_LAUNCH_CHECK();
  if (timing) {
    HP_HIP(hipEventRecord(c->ev[3], s));
    HP_HIP(hipEventSynchronize(c->ev[3]));
    float t[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < 3; ++i) HP_HIP(hipEventElapsedTime(&t[i], c->ev[i], c->ev[i + 1]));
    c->info.inverse_ms = t[0];
    c->info.gram_ms = t[1];
    c->info.contract_ms = t[2];
  }
  return HP_OK;
}

}  // namespace hp

extern "C" {

int hp_dev_overlap_forward(const double* centers, const double* alpha, const double* b_hi, const double* b_lo,
                           size_t n, int factor_kind, hp_adjoint** ctx, hp_adjoint_info* info, void* stream) {
  if (!centers || !alpha || !b_hi || !b_lo || !ctx || !info) {
    set_error("overlap forward: null argument");
    return HP_EINVAL;
  }
  return adjoint_forward(centers, alpha, b_hi, b_lo, n, factor_kind, ctx, info, (hipStream_t)stream);
}

int hp_dev_overlap_gradients(hp_adjoint* ctx, int gemm_mode, double* gld_hi, double* gld_lo, double* gq_hi,
                             double* gq_lo, void* stream) {
  if (!ctx || !gq_hi || !gq_lo || (!gld_hi) != (!gld_lo)) {
    set_error("overlap gradients: null argument");
    return HP_EINVAL;
  }
  return adjoint_gradients(ctx, gemm_mode, gld_hi, gld_lo, gq_hi, gq_lo, (hipStream_t)stream);
}

int hp_dev_overlap_backward(hp_adjoint* ctx, const double* cot, int gemm_mode, double* gc_hi, double* gc_lo,
                            double* ga_hi, double* ga_lo, double* gb_hi, double* gb_lo, void* stream) {
  if (!ctx || !cot || !gc_hi || !gc_lo || !ga_hi || !ga_lo || (!gb_hi) != (!gb_lo)) {
    set_error("overlap backward: null argument");
    return HP_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const long n = ctx->n;
  int rc;
  if (!ctx->G.p && (rc = ctx->G.alloc(16 * n * sizeof(double)))) return rc;
  double* g1 = ctx->G.d();
  double* g2 = g1 + 8 * n;
  const dd c1 = mkdd(cot[0], cot[1]), c2 = mkdd(cot[2], cot[3]);
  const bool need1 = c1.hi != 0.0 || c1.lo != 0.0;  // b^T S^-1 b alone needs no S^-1
  if ((rc = adjoint_gradients(ctx, gemm_mode, need1 ? g1 : nullptr, need1 ? g1 + 4 * n : nullptr, g2, g2 + 4 * n, s)))
    return rc;
  hipLaunchKernelGGL(k_adj_combine, dim3(nblocks(4 * n, 256)), dim3(256), 0, s, need1 ? g1 : nullptr,
                     need1 ? g1 + 4 * n : nullptr, g2, g2 + 4 * n, c1, c2, n, gc_hi, gc_lo, ga_hi, ga_lo,
                     ctx->x.d(), ctx->x.d() + n, gb_hi, gb_lo);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

int hp_adjoint_get_info(const hp_adjoint* ctx, hp_adjoint_info* info) {
  if (!ctx || !info) {
    set_error("adjoint info: null argument");
    return HP_EINVAL;
  }
  *info = ctx->info;
  return HP_OK;
}

void hp_adjoint_destroy(hp_adjoint* ctx) {
  if (!ctx) return;
  (void)hipDeviceSynchronize();
  delete ctx;
}

int hp_overlap_adjoint(const double* centers, const double* alpha, const double* b_dd, size_t n, int factor_kind,
                       int gemm_mode, hp_adjoint_info* info, double* g_ld_dd, double* g_q_dd) {
  if (!centers || !alpha || !b_dd || !info || !g_ld_dd || !g_q_dd) {
    set_error("overlap adjoint: null argument");
    return HP_EINVAL;
  }
  if (n == 0 || n > 65535) {
    set_error("overlap adjoint: n = %zu outside [1, 65535]", n);
    return HP_ESHAPE;
  }
  hipStream_t s = nullptr;
  DevBuf dc, da, db, dg;
  int rc;
  if ((rc = dc.alloc(3 * n * sizeof(double))) || (rc = da.alloc(n * sizeof(double))) ||
      (rc = db.alloc(2 * n * sizeof(double))) || (rc = dg.alloc(16 * n * sizeof(double))))
    return rc;
  HP_HIP(hipMemcpy(dc.p, centers, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  HP_HIP(hipMemcpy(da.p, alpha, n * sizeof(double), hipMemcpyHostToDevice));
  if ((rc = upload_planes(b_dd, n, 2, db.d(), s))) return rc;
  hp_adjoint* ctx = nullptr;
  if ((rc = adjoint_forward(dc.d(), da.d(), db.d(), db.d() + n, n, factor_kind, &ctx, info, s))) return rc;
  rc = adjoint_gradients(ctx, gemm_mode, dg.d(), dg.d() + 4 * n, dg.d() + 8 * n, dg.d() + 12 * n, s);
  if (!rc) rc = download_planes(dg.d(), 4 * n, 2, g_ld_dd, s);
  if (!rc) rc = download_planes(dg.d() + 8 * n, 4 * n, 2, g_q_dd, s);
  if (!rc && hipStreamSynchronize(s) != hipSuccess) {
    set_error("overlap adjoint: synchronisation failed");
    rc = HP_EDEVICE;
  }
  hp_adjoint_destroy(ctx);
  return rc;
}

}  // extern "C"
