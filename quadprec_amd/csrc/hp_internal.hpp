// hp_internal.hpp -- shared host-side plumbing of the HIP library: status codes, the
// thread-local error message, HIP error checks and device scratch buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <utility>
#include <vector>

#include <cstddef>
#include <cstdio>
#include <string>

#include "../../include/hyperprec_hip.h"

namespace hp {

void set_error(const char* fmt, ...);

#define HP_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t _e = (call);                                                           \
    if (_e != hipSuccess) {                                                           \
      ::hp::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                      __LINE__);                                                      \
      return HP_EDEVICE;                                                              \
    }                                                                                 \
  } while (0)

#define HP_LAUNCH_CHECK()                                                             \
  do {                                                                                \
    hipError_t _e = hipGetLastError();                                                \
    if (_e != hipSuccess) {                                                           \
      ::hp::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),       \
                      __FILE__, __LINE__);                                            \
      return HP_EDEVICE;                                                              \
    }                                                                                 \
  } while (0)

// RAII device buffer used by the host (drop-in) entry points.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t nbytes) {
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
      return HP_ENOMEM;
    }
    bytes = nbytes;
    return HP_OK;
  }
  double* d() const { return static_cast<double*>(p); }
};

// Optional per-phase HIP-event timing of multi-kernel entry points (hp_set_phase_timing).
// Phase boundaries are recorded on the stream the kernels run on.
struct PhaseTimer {
  static constexpr int MAXP = 8;
  int on = 0;  // 0 off; 1: each call waits for its own events; 2: calls accumulate (collect)
  int n = 0;   // boundaries recorded by the current call
  hipEvent_t ev[MAXP + 1] = {};
  double ms[MAXP] = {};
  // mode 2: every call's boundaries in a growing pool of events, summed by collect()
  std::vector<hipEvent_t> pool;
  size_t used = 0, call_start = 0;
  std::vector<std::pair<size_t, int>> calls;  // (first boundary in the pool, boundaries)
  void mark(hipStream_t s);  // record the next boundary (no-op when off)
  void finish();             // mode 1: after synchronisation, elapsed times between boundaries
  int collect(double* sum, int max, int* ncalls);  // mode 2: per-phase sums since the last collect
};
PhaseTimer& phase_timer();

// AoS <-> SoA conversion for the host entry points (the host array is a Rust &[f128]
// layout: hi, lo interleaved).  `limbs` is 2 for dd, 4 for MF4.
int upload_planes(const double* host_aos, size_t count, int limbs, double* dev_planes, hipStream_t s);
int download_planes(const double* dev_planes, size_t count, int limbs, double* host_aos, hipStream_t s);

// kernels' launchers (device pointers, SoA planes)
int launch_gemm_dd_plain(int trans_a, const double* ahi, const double* alo, const double* bhi,
                         const double* blo, double* chi, double* clo, size_t m, size_t n, size_t k,
                         hipStream_t s);
// strided / negated / batched / triangular form (see k_gemm_dd.hip)
int launch_gemm_dd_plain_ex(int trans_a, const double* ahi, const double* alo, long lda, const double* bhi,
                            const double* blo, long ldb, double* chi, double* clo, long ldc, size_t m, size_t n,
                            size_t k, bool neg, bool tri, int batch, long sa, long sb, long sc, hipStream_t s);
int launch_gemm_split_i8(int trans_a, const double* ahi, const double* alo, const double* bhi, const double* blo,
                         double* chi, double* clo, size_t m, size_t n, size_t k, int R, hipStream_t s);
// operand of the split GEMM: logical element (row, p) at hi[row*ld + p] when contig, else
// hi[p*ld + row]; A is m x k, B is given as its transpose (n x k)
struct SplitOperand {
  const double* hi;
  const double* lo;
  long ld;
  bool contig;
};
struct SplitOpts {
  bool lower_only = false;    // C symmetric: only j <= i is formed (m == n)
  bool subtract = false;      // C -= A B^T instead of C = A B^T
  bool same_operand = false;  // B is A (one set of residues)
  int slot = 0;     // workspace 0 or 1: calls in flight on two streams at once use different slots
  int variant = 0;  // int8 kernel (a variant number of set_split_gemm_variant); 0: the tuned one
};
int split_gemm_general(const SplitOperand& A, const SplitOperand& B, double* chi, double* clo, size_t ldc, size_t m,
                       size_t n, size_t k, int R, hipStream_t s, const SplitOpts& opt);
int crt_bits_W(int R, size_t k);
int split_slices_for_bits(int bits, size_t k);
int plain_bits_for_k(size_t k);
int gemm_plan_get(int trans_a, size_t m, size_t n, size_t k, int bits, hp_gemm_plan* plan);
int gemm_auto(int trans_a, const double* ahi, const double* alo, const double* bhi, const double* blo, double* chi,
              double* clo, size_t m, size_t n, size_t k, int bits, hipStream_t s);
extern std::atomic<int> g_split_panel;
int solve_refined(int kind, const double* sh, const double* sl, const double* bh, const double* bl, long n, double tol,
                  int max_iter, double* xh, double* xl, hp_refine_stats* st, hipStream_t s);
int split_gemm_variant(size_t m, size_t n, size_t k, int R);
int set_split_gemm_variant(size_t m, size_t n, size_t k, int R, int variant);
int launch_mf4_gemm(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n,
                    size_t k, hipStream_t s);
// MultiFloat<nl> GEMM on device limb planes (nl = 1..8; 4 takes the f256 kernel)
int launch_mfn_gemm(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                    hipStream_t s);
// MultiFloat<4> elementwise op on device limb planes with the generic op codes (0 add, 1 sub,
// 2 mul, 3 div, 4 recip, 5 sqrt, 6 renormalize)
int launch_mfn_elem4(int op, const double* a, const double* b, double* out, size_t n, hipStream_t s);

}  // namespace hp
