// api_core.hip -- C ABI plumbing: errors, AoS<->SoA staging, GEMM entry points and the
// deterministic input generators.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {
// interleaved (AoS) <-> limb planes (SoA) on the device, so host copies stay contiguous
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, size_t count, int limbs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride)
    for (int l = 0; l < limbs; ++l) soa[l * count + i] = aos[i * limbs + l];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, size_t count, int limbs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride)
    for (int l = 0; l < limbs; ++l) aos[i * limbs + l] = soa[l * count + i];
}
unsigned grid_for(size_t count) {
  size_t g = (count + 255) / 256;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return (unsigned)g;
}
}  // namespace

int upload_planes(const double* host_aos, size_t count, int limbs, double* dev_planes, hipStream_t s) {
  if (count == 0) return HP_OK;
  DevBuf tmp;
  int rc = tmp.alloc(count * limbs * sizeof(double));
  if (rc) return rc;
  HP_HIP(hipMemcpyAsync(tmp.d(), host_aos, count * limbs * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(count)), dim3(256), 0, s, tmp.d(), dev_planes, count, limbs);
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

int download_planes(const double* dev_planes, size_t count, int limbs, double* host_aos, hipStream_t s) {
  if (count == 0) return HP_OK;
  DevBuf tmp;
  int rc = tmp.alloc(count * limbs * sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(count)), dim3(256), 0, s, dev_planes, tmp.d(), count, limbs);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(host_aos, tmp.d(), count * limbs * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

static thread_local PhaseTimer g_timer;
PhaseTimer& phase_timer() { return g_timer; }
void PhaseTimer::mark(hipStream_t s) {
  if (!on || n > MAXP) return;
  if (on == 2) {
    if (n == 0) call_start = used;
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return;
      pool.push_back(e);
    }
    (void)hipEventRecord(pool[used++], s);
    ++n;
    return;
  }
  if (!ev[n]) (void)hipEventCreate(&ev[n]);
  (void)hipEventRecord(ev[n], s);
  ++n;
}
void PhaseTimer::finish() {
  if (on == 2) {
    if (n > 1) calls.emplace_back(call_start, n);
    return;
  }
  if (!on) return;
  for (int i = 0; i + 1 < n; ++i) {
    float t = 0.f;
    (void)hipEventSynchronize(ev[i + 1]);
    (void)hipEventElapsedTime(&t, ev[i], ev[i + 1]);
    ms[i] = t;
  }
}
int PhaseTimer::collect(double* sum, int max, int* ncalls) {
  int phases = 0;
  for (int i = 0; i < max; ++i) sum[i] = 0.0;
  if (!calls.empty()) {
    const auto& last = calls.back();
    (void)hipEventSynchronize(pool[last.first + last.second - 1]);
    phases = MAXP;
    for (const auto& c : calls) phases = std::min(phases, c.second - 1);
    for (const auto& c : calls)
      for (int i = 0; i < phases && i < max; ++i) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, pool[c.first + i], pool[c.first + i + 1]);
        sum[i] += t;
      }
  }
  if (ncalls) *ncalls = (int)calls.size();
  calls.clear();
  used = 0;
  return std::min(phases, max);
}

}  // namespace hp

using namespace hp;

namespace {
// limbs 2 with dd: the f128 kernel; otherwise MultiFloat<limbs> (MultiFloat<2> is not f128:
// its add and mul are the expansion algorithms)
int gemm_host(int trans_a, int limbs, bool dd, const double* a, const double* b, double* c, size_t m, size_t n,
              size_t k) {
  if (!a || !b || !c) {
    if (m * n == 0) return HP_OK;
    set_error("gemm: null pointer");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  DevBuf da, db, dc;
  int rc;
  if ((rc = da.alloc(m * k * limbs * sizeof(double))) || (rc = db.alloc(k * n * limbs * sizeof(double))) ||
      (rc = dc.alloc(m * n * limbs * sizeof(double))))
    return rc;
  if ((rc = upload_planes(a, m * k, limbs, da.d(), s)) || (rc = upload_planes(b, k * n, limbs, db.d(), s))) return rc;
  if (dd)
    rc = launch_gemm_dd_plain(trans_a, da.d(), da.d() + m * k, db.d(), db.d() + k * n, dc.d(), dc.d() + m * n, m, n,
                              k, s);
  else
    rc = launch_mfn_gemm(limbs, trans_a, da.d(), db.d(), dc.d(), m, n, k, s);
  if (rc) return rc;
  return download_planes(dc.d(), m * n, limbs, c, s);
}

// SplitMix64 (Steele, Lea, Flood 2014)
struct SplitMix64 {
  uint64_t x;
  uint64_t next() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};
struct Normal {
  SplitMix64 rng;
  bool have = false;
  double spare = 0.0;
  double next() {  // Box-Muller
    if (have) { have = false; return spare; }
    double u1 = 1.0 - rng.uniform();  // (0, 1]
    double u2 = rng.uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    double th = 6.283185307179586 * u2;
    spare = r * std::sin(th);
    have = true;
    return r * std::cos(th);
  }
};
}  // namespace

extern "C" {

const char* hp_last_error(void) { return hp::g_err; }

int hp_split_gemm_variant(size_t m, size_t n, size_t k, int slices) { return split_gemm_variant(m, n, k, slices); }
int hp_split_slices_for_bits(size_t k, int bits) { return split_slices_for_bits(bits, k); }
int hp_gemm_plan_get(int trans_a, size_t m, size_t n, size_t k, int bits, hp_gemm_plan* plan) {
  if (!plan) {
    set_error("hp_gemm_plan_get: null plan");
    return HP_EINVAL;
  }
  return gemm_plan_get(trans_a, m, n, k, bits, plan);
}
int hp_split_gemm_set_variant(size_t m, size_t n, size_t k, int slices, int variant) {
  return set_split_gemm_variant(m, n, k, slices, variant);
}

void hp_set_phase_timing(int on) {
  PhaseTimer& t = phase_timer();
  t.on = on == 2 ? 2 : (on != 0);
  t.calls.clear();
  t.used = 0;
}
int hp_collect_phase_times(double* ms_sum, int max, int* calls) {
  if (!ms_sum || max < 0) {
    set_error("hp_collect_phase_times: null pointer");
    return HP_EINVAL;
  }
  return phase_timer().collect(ms_sum, max, calls);
}
int hp_get_phase_times(double* ms, int max) {
  PhaseTimer& t = phase_timer();
  int cnt = t.n > 0 ? t.n - 1 : 0;
  for (int i = 0; i < cnt && i < max; ++i) ms[i] = t.ms[i];
  return cnt;
}
int hp_version(void) { return 200; }

int hp_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  return gemm_host(0, 2, true, a, b, c, m, n, k);
}
int hp_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  return gemm_host(1, 2, true, a, b, c, m, n, k);
}
int hp_mf_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  return gemm_host(0, 4, false, a, b, c, m, n, k);
}
int hp_mf_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  return gemm_host(1, 4, false, a, b, c, m, n, k);
}
int hp_mfn_gemm(int nl, const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  if (nl < 1 || nl > 8) return launch_mfn_gemm(nl, 0, a, b, c, m, n, k, nullptr);  // the error
  return gemm_host(0, nl, false, a, b, c, m, n, k);
}
int hp_mfn_gemm_atb(int nl, const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  if (nl < 1 || nl > 8) return launch_mfn_gemm(nl, 1, a, b, c, m, n, k, nullptr);
  return gemm_host(1, nl, false, a, b, c, m, n, k);
}

int hp_dev_gemm(int trans_a, const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo,
                double* c_hi, double* c_lo, size_t m, size_t n, size_t k, int mode, int slices, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  switch (mode) {
    case HP_GEMM_PLAIN:
      return launch_gemm_dd_plain(trans_a, a_hi, a_lo, b_hi, b_lo, c_hi, c_lo, m, n, k, s);
    case HP_GEMM_SPLIT_I8:
      return launch_gemm_split_i8(trans_a, a_hi, a_lo, b_hi, b_lo, c_hi, c_lo, m, n, k, slices, s);
    case HP_GEMM_AUTO:
      return gemm_auto(trans_a, a_hi, a_lo, b_hi, b_lo, c_hi, c_lo, m, n, k, slices, s);
    default:
      set_error("hp_dev_gemm: unknown mode %d", mode);
      return HP_EINVAL;
  }
}

int hp_dev_mf_gemm(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                   void* stream) {
  return launch_mf4_gemm(trans_a, a, b, c, m, n, k, (hipStream_t)stream);
}

void hp_gen_dd_normal(uint64_t seed, size_t count, double* out) {
  Normal nd{SplitMix64{seed}};
  SplitMix64 u{seed ^ 0xD1B54A32D192ED03ull};
  for (size_t i = 0; i < count; ++i) {
    double hi = nd.next();
    double lo = hi * ((2.0 * u.uniform() - 1.0) * 0x1.0p-53);
    double s, e;
    quick_two_sum(hi, lo, s, e);
    out[2 * i] = s;
    out[2 * i + 1] = e;
  }
}

void hp_gen_mf4_normal(uint64_t seed, size_t count, double* out) {
  Normal nd{SplitMix64{seed}};
  SplitMix64 u{seed ^ 0xD1B54A32D192ED03ull};
  for (size_t i = 0; i < count; ++i) {
    double l[4];
    l[0] = nd.next();
    for (int q = 1; q < 4; ++q) l[q] = l[q - 1] * ((2.0 * u.uniform() - 1.0) * 0x1.0p-53);
    mf_renormalize(l);
    for (int q = 0; q < 4; ++q) out[4 * i + q] = l[q];
  }
}

void hp_gen_overlap_basis(uint64_t seed, size_t n, double alpha_lo, double alpha_hi, double spacing, double jitter,
                          double* centers, double* alpha) {
  SplitMix64 u{seed};
  size_t L = 1;
  while (L * L * L < n) ++L;
  for (size_t i = 0; i < n; ++i) {
    size_t ix = i / (L * L), iy = (i / L) % L, iz = i % L;
    size_t g[3] = {ix, iy, iz};
    for (int c = 0; c < 3; ++c) centers[3 * i + c] = (double)g[c] * spacing + (2.0 * u.uniform() - 1.0) * jitter;
  }
  const double l0 = std::log10(alpha_lo), l1 = std::log10(alpha_hi);
  for (size_t i = 0; i < n; ++i) alpha[i] = std::pow(10.0, l0 + (l1 - l0) * u.uniform());
}

}  // extern "C"
