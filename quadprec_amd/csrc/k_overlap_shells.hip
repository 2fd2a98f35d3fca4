// k_overlap_shells.hip -- assemble_overlap_mixed over contracted s/p/d/f shells
// (crates/overlap/src/assemble.rs:35-85, screen.rs:57-101, integral.rs:65-346).
//
// One thread per element (a, b) of the n_ao x n_ao matrix whose shell pair (s_a, s_b) has
// s_a >= s_b; it screens the pair (every thread of a pair reaches the same tier), forms its
// Cartesian component of the contracted block in the tier's arithmetic and writes it with
// its mirror.  The first element of each pair counts the tier.  Per element the operations
// are the reference's, in its order: the primitive-pair sums run p outer, q inner; the 1D
// Obara-Saika tables are built as integral.rs builds them (statically unrolled 4 x 4 so they
// stay in registers).  libm exp (glibc on the host, the device library here) is the one
// operation whose last bit may differ.
#include <cmath>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr double PI_F64 = 3.141592653589793;

struct ShellRec {
  double c[3];
  double amin, amax;  // alpha_min / alpha_max over the contraction (basis.rs:40-48)
  long p0, p1;        // primitives
  int l, off;         // angular momentum, first AO
};

__device__ __forceinline__ double dist2(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

// screen_shell_pair: screen.rs:57-101 (0 F64, 1 F128, 2 Skip); prm = {diffuse, ratio,
// -ln(eps_skip), safe_distance}
__device__ int screen(const ShellRec& a, const ShellRec& b, const double* prm) {
  const double r2 = dist2(a.c, b.c);
  const double alpha_sum = a.amin + b.amin;
  const double gamma = a.amin * b.amin / alpha_sum;
  if (gamma * r2 > prm[2]) return 2;
  const bool ad = a.amin < prm[0], bd = b.amin < prm[0];
  if (!ad && !bd) return 0;
  const double ratio = fmax(a.amax / b.amin, b.amax / a.amin);
  const double r = sqrt(r2);
  if (ratio > prm[1] && r < prm[3]) return 1;
  if (ad && bd && r < prm[3]) return 1;
  return 0;
}

// cart_components (integral.rs:220-231): the u-th (lx, ly, lz) of angular momentum l
__device__ __forceinline__ void cart(int l, int u, int& lx, int& ly, int& lz) {
  int n = 0;
  lx = ly = lz = 0;
#pragma unroll
  for (int x = 3; x >= 0; --x)
#pragma unroll
    for (int y = 3; y >= 0; --y) {
      if (x > l || y > l - x) continue;
      if (n == u) {
        lx = x;
        ly = y;
        lz = l - x - y;
      }
      ++n;
    }
}

// entry (ei, ej) of obara_saika_1d_f64 (integral.rs:158-185)
__device__ double os1d_f64(int la, int lb, double pa, double pb, double gamma, int ei, int ej) {
  double s[4][4];
  const double oo2g = 0.5 / gamma;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) s[i][j] = 0.0;
  s[0][0] = 1.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < la) s[i + 1][0] = pa * s[i][0] + oo2g * (double)i * (i > 0 ? s[i > 0 ? i - 1 : 0][0] : 0.0);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j >= lb) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i > la) continue;
      const double tim1 = oo2g * (double)i * (i > 0 ? s[i > 0 ? i - 1 : 0][j] : 0.0);
      const double tjm1 = oo2g * (double)j * (j > 0 ? s[i][j > 0 ? j - 1 : 0] : 0.0);
      s[i][j + 1] = pb * s[i][j] + tim1 + tjm1;
    }
  }
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i == ei && j == ej) r = s[i][j];
  return r;
}

// entry (ei, ej) of obara_saika_1d_f128 (integral.rs:187-216)
__device__ dd os1d_dd(int la, int lb, dd pa, dd pb, dd gamma, int ei, int ej) {
  dd s[4][4];
  const dd oo2g = dd_div(mkdd(0.5), gamma);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) s[i][j] = mkdd(0.0);
  s[0][0] = mkdd(1.0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i >= la) continue;
    const dd prev = i > 0 ? s[i > 0 ? i - 1 : 0][0] : mkdd(0.0);
    s[i + 1][0] = dd_add(dd_mul(pa, s[i][0]), dd_mul(dd_mul(oo2g, mkdd((double)i)), prev));
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j >= lb) continue;
    const dd fj = mkdd((double)j);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i > la) continue;
      const dd fi = mkdd((double)i);
      const dd im1 = i > 0 ? s[i > 0 ? i - 1 : 0][j] : mkdd(0.0);
      const dd jm1 = j > 0 ? s[i][j > 0 ? j - 1 : 0] : mkdd(0.0);
      s[i][j + 1] = dd_add(dd_add(dd_mul(pb, s[i][j]), dd_mul(dd_mul(oo2g, fi), im1)), dd_mul(dd_mul(oo2g, fj), jm1));
    }
  }
  dd r = mkdd(0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i == ei && j == ej) r = s[i][j];
  return r;
}

// overlap_primitive_f64 (integral.rs:15-25)
__device__ double prim_f64(double ai, const double* ci, double aj, const double* cj) {
  const double gamma = ai + aj;
  const double mu = ai * aj / gamma;
  const double r2 = dist2(ci, cj);
  const double sq = sqrt(PI_F64 / gamma);
  const double prefactor = (sq * sq) * sq;  // powi(3)
  return prefactor * exp(-mu * r2);
}
// overlap_primitive_f128 (integral.rs:29-56): f64 pi, f64 centre differences, f64 exp
__device__ dd prim_dd(double ai, const double* ci, double aj, const double* cj) {
  const dd a = mkdd(ai), b = mkdd(aj);
  const dd gamma = dd_add(a, b);
  const dd mu = dd_div(dd_mul(a, b), gamma);
  const dd dx = mkdd(ci[0] - cj[0]), dy = mkdd(ci[1] - cj[1]), dz = mkdd(ci[2] - cj[2]);
  const dd r2 = dd_add(dd_add(dd_mul(dx, dx), dd_mul(dy, dy)), dd_mul(dz, dz));
  const dd ratio = dd_div(mkdd(PI_F64), gamma);
  const dd prefactor = dd_mul(ratio, dd_sqrt(ratio));
  const dd nmr2 = dd_mul(dd_sub(mkdd(0.0), mu), r2);
  const dd exp_val = dd_mul(mkdd(exp(nmr2.hi)), dd_add(mkdd(1.0), mkdd(nmr2.lo)));
  return dd_mul(prefactor, exp_val);
}

// component (u, v) of the contracted block of shell pair (A, B) in the given tier
__device__ double element(const ShellRec& A, const ShellRec& B, int u, int v, int tier,
                          const double* __restrict__ alpha, const double* __restrict__ coeff) {
  if (tier == 0) {
    if (A.l == 0 && B.l == 0) {  // integral.rs:99-110
      double sum = 0.0;
      for (long p = A.p0; p < A.p1; ++p)
        for (long q = B.p0; q < B.p1; ++q) sum += coeff[p] * coeff[q] * prim_f64(alpha[p], A.c, alpha[q], B.c);
      return sum;
    }
    int ux, uy, uz, vx, vy, vz;
    cart(A.l, u, ux, uy, uz);
    cart(B.l, v, vx, vy, vz);
    double acc = 0.0;
    for (long p = A.p0; p < A.p1; ++p)  // integral.rs:233-282
      for (long q = B.p0; q < B.p1; ++q) {
        const double ai = alpha[p], aj = alpha[q];
        const double gamma = ai + aj;
        const double mu = ai * aj / gamma;
        const double r2 = dist2(A.c, B.c);
        const double sq = sqrt(PI_F64 / gamma);
        const double prefactor = (sq * sq) * sq * exp(-mu * r2);
        double pa[3], pb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double pk = (ai * A.c[k] + aj * B.c[k]) / gamma;
          pa[k] = pk - A.c[k];
          pb[k] = pk - B.c[k];
        }
        const double sx = os1d_f64(A.l, B.l, pa[0], pb[0], gamma, ux, vx);
        const double sy = os1d_f64(A.l, B.l, pa[1], pb[1], gamma, uy, vy);
        const double sz = os1d_f64(A.l, B.l, pa[2], pb[2], gamma, uz, vz);
        acc += sx * sy * sz * prefactor * coeff[p] * coeff[q];
      }
    return acc;
  }
  if (A.l == 0 && B.l == 0) {  // integral.rs:119-136
    dd sum = mkdd(0.0);
    for (long p = A.p0; p < A.p1; ++p)
      for (long q = B.p0; q < B.p1; ++q)
        sum = dd_add(sum, dd_mul(dd_mul(mkdd(coeff[p]), mkdd(coeff[q])), prim_dd(alpha[p], A.c, alpha[q], B.c)));
    return sum.hi + sum.lo;
  }
  int ux, uy, uz, vx, vy, vz;
  cart(A.l, u, ux, uy, uz);
  cart(B.l, v, vx, vy, vz);
  dd acc = mkdd(0.0);
  for (long p = A.p0; p < A.p1; ++p)  // integral.rs:284-346
    for (long q = B.p0; q < B.p1; ++q) {
      const dd ai = mkdd(alpha[p]), aj = mkdd(alpha[q]);
      const dd gamma = dd_add(ai, aj);
      const dd mu = dd_div(dd_mul(ai, aj), gamma);
      const dd dx = mkdd(A.c[0] - B.c[0]), dy = mkdd(A.c[1] - B.c[1]), dz = mkdd(A.c[2] - B.c[2]);
      const dd r2 = dd_add(dd_add(dd_mul(dx, dx), dd_mul(dy, dy)), dd_mul(dz, dz));
      const dd ratio = dd_div(mkdd(PI_F64), gamma);
      const dd prefactor = dd_mul(ratio, dd_sqrt(ratio));
      const dd nmr2 = dd_mul(dd_sub(mkdd(0.0), mu), r2);
      const dd exp_val = dd_mul(mkdd(exp(nmr2.hi)), dd_add(mkdd(1.0), mkdd(nmr2.lo)));
      const dd total_pre = dd_mul(prefactor, exp_val);
      dd pa[3], pb[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const dd cik = mkdd(A.c[k]), cjk = mkdd(B.c[k]);
        const dd pk = dd_div(dd_add(dd_mul(ai, cik), dd_mul(aj, cjk)), gamma);
        pa[k] = dd_sub(pk, cik);
        pb[k] = dd_sub(pk, cjk);
      }
      const dd sx = os1d_dd(A.l, B.l, pa[0], pb[0], gamma, ux, vx);
      const dd sy = os1d_dd(A.l, B.l, pa[1], pb[1], gamma, uy, vy);
      const dd sz = os1d_dd(A.l, B.l, pa[2], pb[2], gamma, uz, vz);
      const dd cf = dd_mul(dd_mul(mkdd(coeff[p]), mkdd(coeff[q])), total_pre);
      acc = dd_add(acc, dd_mul(dd_mul(dd_mul(sx, sy), sz), cf));
    }
  return acc.hi + acc.lo;
}

__global__ __launch_bounds__(128) void k_overlap_shells(const ShellRec* __restrict__ sh, const int* __restrict__ ao_shell,
                                                        const double* __restrict__ alpha,
                                                        const double* __restrict__ coeff, long n_ao, const double* prm,
                                                        double* __restrict__ s, unsigned long long* __restrict__ counts) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_ao) return;
  for (long a = blockIdx.y; a < n_ao; a += gridDim.y) {
    const int sa = ao_shell[a], sb = ao_shell[b];
    if (sa < sb) continue;
    const ShellRec A = sh[sa], B = sh[sb];
    const int u = (int)(a - A.off), v = (int)(b - B.off);
    const int tier = screen(A, B, prm);
    if (u == 0 && v == 0) atomicAdd(&counts[tier], 1ull);
    if (tier == 2) continue;
    const double val = element(A, B, u, v, tier, alpha, coeff);
    s[a * n_ao + b] = val;
    if (sa != sb) s[b * n_ao + a] = val;
  }
}

}  // namespace
}  // namespace hp

using namespace hp;

extern "C" {

int hp_assemble_overlap_mixed(const double* centers, const int* l, const long* prim_off, const double* alpha,
                              const double* coeff, size_t n_shells, const double* params4, double* s, size_t n_ao,
                              size_t* counts) {
  if (!counts || !params4) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  counts[0] = counts[1] = counts[2] = 0;
  if (n_shells == 0) return n_ao == 0 ? HP_OK : HP_ESHAPE;
  if (!centers || !l || !prim_off || !alpha || !coeff || !s) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  std::vector<ShellRec> rec(n_shells);
  std::vector<int> ao_shell;
  long off = 0;
  for (size_t i = 0; i < n_shells; ++i) {
    if (l[i] < 0 || l[i] > 3 || prim_off[i + 1] < prim_off[i]) {
      set_error("assemble_overlap_mixed: shell %zu: l must be 0..3 (S, P, D, F) and primitives non-negative", i);
      return HP_EINVAL;
    }
    ShellRec& r = rec[i];
    for (int k = 0; k < 3; ++k) r.c[k] = centers[3 * i + k];
    r.amin = INFINITY;
    r.amax = 0.0;
    for (long p = prim_off[i]; p < prim_off[i + 1]; ++p) {
      r.amin = std::fmin(r.amin, alpha[p]);
      r.amax = std::fmax(r.amax, alpha[p]);
    }
    r.p0 = prim_off[i];
    r.p1 = prim_off[i + 1];
    r.l = l[i];
    r.off = (int)off;
    const int nf = (l[i] + 1) * (l[i] + 2) / 2;
    for (int k = 0; k < nf; ++k) ao_shell.push_back((int)i);
    off += nf;
  }
  if ((size_t)off != n_ao) {
    set_error("assemble_overlap_mixed: n_ao = %zu, the shells give %ld", n_ao, off);
    return HP_ESHAPE;
  }
  const long np = prim_off[n_shells];
  const double prm[4] = {params4[0], params4[1], -std::log(params4[2]), params4[3]};
  hipStream_t st = hipStreamPerThread;
  DevBuf dsh, dao, dal, dco, dprm, ds, dcnt;
  int rc;
  if ((rc = dsh.alloc(n_shells * sizeof(ShellRec))) || (rc = dao.alloc(n_ao * sizeof(int))) ||
      (rc = dal.alloc(np * sizeof(double))) || (rc = dco.alloc(np * sizeof(double))) || (rc = dprm.alloc(32)) ||
      (rc = ds.alloc(n_ao * n_ao * sizeof(double))) || (rc = dcnt.alloc(3 * sizeof(unsigned long long))))
    return rc;
  HP_HIP(hipMemcpyAsync(dsh.p, rec.data(), n_shells * sizeof(ShellRec), hipMemcpyHostToDevice, st));
  HP_HIP(hipMemcpyAsync(dao.p, ao_shell.data(), n_ao * sizeof(int), hipMemcpyHostToDevice, st));
  if (np > 0) {
    HP_HIP(hipMemcpyAsync(dal.p, alpha, np * sizeof(double), hipMemcpyHostToDevice, st));
    HP_HIP(hipMemcpyAsync(dco.p, coeff, np * sizeof(double), hipMemcpyHostToDevice, st));
  }
  HP_HIP(hipMemcpyAsync(dprm.p, prm, sizeof(prm), hipMemcpyHostToDevice, st));
  HP_HIP(hipMemsetAsync(ds.p, 0, n_ao * n_ao * sizeof(double), st));
  HP_HIP(hipMemsetAsync(dcnt.p, 0, 3 * sizeof(unsigned long long), st));
  const unsigned gx = (unsigned)((n_ao + 127) / 128), gy = (unsigned)std::min<size_t>(n_ao, 65535);
  hipLaunchKernelGGL(k_overlap_shells, dim3(gx, gy), dim3(128), 0, st, static_cast<const ShellRec*>(dsh.p),
                     static_cast<const int*>(dao.p), dal.d(), dco.d(), (long)n_ao, dprm.d(), ds.d(),
                     static_cast<unsigned long long*>(dcnt.p));
  HP_LAUNCH_CHECK();
  unsigned long long hc[3];
  HP_HIP(hipMemcpyAsync(s, ds.p, n_ao * n_ao * sizeof(double), hipMemcpyDeviceToHost, st));
  HP_HIP(hipMemcpyAsync(hc, dcnt.p, sizeof(hc), hipMemcpyDeviceToHost, st));
  HP_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 3; ++k) counts[k] = (size_t)hc[k];
  return HP_OK;
}

}  // extern "C"
