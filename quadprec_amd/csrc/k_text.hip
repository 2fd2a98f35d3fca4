// k_text.hip -- the text wire formats of the dd and MF4 value types on the GPU:
//   * Display: f128 "{:.32e}" of hi + lo (crates/quad/src/dd.rs:337-341); MultiFloat<4>
//     "{:.64e}" of limb0 + limb1 (multifloat/fmt.rs:7-17);
//   * human-readable serde: f128 "{:+.18e}+{:+.18e}" (dd.rs:446-450) and its parse
//     (dd.rs:496-511); MultiFloat "{:+.14e}" limbs joined by '+' (multifloat/traits.rs:14-24)
//     and its parse (traits.rs:104-148);
//   * FromStr: the digit-by-digit dd / MF4 accumulation of dd.rs:371-435 and
//     multifloat/fmt.rs:32-114;
//   * Rust's own f64 `{:.Pe}` formatting and `str::parse::<f64>`, which the formats above
//     are built from.
//
// Text arrays are Arrow-style: n strings concatenated, with n + 1 byte offsets.
//
// Exact decimal conversion.  Every finite double is M * 2^E; scaling it by 10^s is
// multiplying by 2^s 5^s.  Both directions use one fixed-size 1024-bit integer per thread
// (32 x u32, indexed only with compile-time indices so it lives in VGPRs): multiply by
// 5^13 chunks, shift, divide by 5^13 chunks (floor division composes, and any nonzero
// remainder is the sticky bit), then round half to even.
//   format: T = floor(2 |x| 10^s) with s = P - floor(log10 |x|) gives the P+1 digits and the
//           rounding bit; the estimate of floor(log10 |x|) is low by at most one, which the
//           size of T shows.
//   parse:  up to 40 significant digits w and decimal exponent q; T = w 5^q (q >= 0) or
//           floor(w 2^A / 5^-q) with A chosen for >= 65 quotient bits, then rounded to 53
//           bits (or fewer, subnormal) with the sticky bit.  Longer digit strings are cut at
//           40 digits: when w and w + 1 round to the same double that double is the answer;
//           otherwise (a string within 1e-40 relative of a halfway point) the element gets
//           status 2 instead of a guess.
// Sizes: the largest intermediate is M 5^388 < 2^954 (format, P = 64, subnormal x) and
// w 2^A < 2^920 (parse), inside the 1024 bits.
//
// Fast paths.  Each call first builds 10^s as truncated 128-bit mantissas F 2^g for
// -365 <= s <= 345 (exactly, with the integer above).  Formatting with P <= 18 and parsing
// up to 19 significant digits then take one 64 x 128 multiply: the exact value lies between
// the products with F and with F + 1, and when both round alike that is the answer; only
// values within about 2^-120 of a rounding boundary (exact ties among them) take the exact
// path, so every result is the exact path's.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int NL = 32;
constexpr int MAX_SIG = 40;
constexpr uint32_t P5_13 = 1220703125u;  // 5^13
constexpr uint32_t P10_9 = 1000000000u;
__constant__ uint32_t POW5[13] = {1u, 5u, 25u, 125u, 625u, 3125u, 15625u, 78125u, 390625u, 1953125u,
                                  9765625u, 48828125u, 244140625u};

struct Big {
  uint32_t w[NL];
};

__device__ __forceinline__ void big_set(Big& b, uint64_t v) {
#pragma unroll
  for (int i = 0; i < NL; ++i) b.w[i] = 0;
  b.w[0] = (uint32_t)v;
  b.w[1] = (uint32_t)(v >> 32);
}
__device__ __forceinline__ void big_mul_add(Big& b, uint32_t m, uint32_t add) {
  uint64_t carry = add;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint64_t t = (uint64_t)b.w[i] * m + carry;
    b.w[i] = (uint32_t)t;
    carry = t >> 32;
  }
}
template <uint32_t D>
__device__ __forceinline__ uint32_t big_div(Big& b) {
  uint64_t r = 0;
#pragma unroll
  for (int i = NL - 1; i >= 0; --i) {
    const uint64_t cur = (r << 32) | b.w[i];
    const uint64_t q = cur / D;
    b.w[i] = (uint32_t)q;
    r = cur - q * D;
  }
  return (uint32_t)r;
}
__device__ __forceinline__ void big_shl(Big& b, int s) {
  for (; s >= 32; s -= 32) {
#pragma unroll
    for (int i = NL - 1; i > 0; --i) b.w[i] = b.w[i - 1];
    b.w[0] = 0;
  }
  if (s > 0) {
#pragma unroll
    for (int i = NL - 1; i > 0; --i) b.w[i] = (b.w[i] << s) | (b.w[i - 1] >> (32 - s));
    b.w[0] <<= s;
  }
}
__device__ __forceinline__ void big_shr(Big& b, int s, bool& sticky) {
  if (s >= 32 * NL) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      sticky |= b.w[i] != 0;
      b.w[i] = 0;
    }
    return;
  }
  for (; s >= 32; s -= 32) {
    sticky |= b.w[0] != 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) b.w[i] = b.w[i + 1];
    b.w[NL - 1] = 0;
  }
  if (s > 0) {
    sticky |= (b.w[0] & ((1u << s) - 1u)) != 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) b.w[i] = (b.w[i] >> s) | (b.w[i + 1] << (32 - s));
    b.w[NL - 1] >>= s;
  }
}
__device__ __forceinline__ void big_mul_pow5(Big& b, int k) {
  for (; k >= 13; k -= 13) big_mul_add(b, P5_13, 0);
  if (k > 0) big_mul_add(b, POW5[k], 0);
}
__device__ __forceinline__ void big_div_pow5(Big& b, int k, bool& sticky) {
  for (; k >= 13; k -= 13) sticky |= big_div<P5_13>(b) != 0;
  for (; k > 0; --k) sticky |= big_div<5u>(b) != 0;
}
__device__ __forceinline__ int big_cmp(const Big& a, const Big& b) {
  int r = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) r = a.w[i] > b.w[i] ? 1 : (a.w[i] < b.w[i] ? -1 : r);
  return r;
}
__device__ __forceinline__ void big_inc(Big& b) {
  uint64_t carry = 1;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint64_t t = (uint64_t)b.w[i] + carry;
    b.w[i] = (uint32_t)t;
    carry = t >> 32;
  }
}
__device__ __forceinline__ bool big_is_zero(const Big& b) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) o |= b.w[i];
  return o == 0;
}
__device__ __forceinline__ void big_pow10(Big& b, int k) {
  big_set(b, 1);
  for (; k >= 9; k -= 9) big_mul_add(b, P10_9, 0);
  for (; k > 0; --k) big_mul_add(b, 10u, 0);
}
// bit length (0 for zero)
__device__ __forceinline__ int big_bitlen(const Big& b) {
  int top = -1;
  uint32_t tw = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i)
    if (b.w[i]) {
      top = i;
      tw = b.w[i];
    }
  return top < 0 ? 0 : top * 32 + (32 - __clz(tw));
}
// bits [pos, pos + 64) and whether any bit below pos is set (pos >= 0)
__device__ __forceinline__ uint64_t big_bits64(const Big& b, int pos, bool& below) {
  const int k = pos >> 5, s = pos & 31;
  uint32_t x0 = 0, x1 = 0, x2 = 0, lowor = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    x0 = i == k ? b.w[i] : x0;
    x1 = i == k + 1 ? b.w[i] : x1;
    x2 = i == k + 2 ? b.w[i] : x2;
    lowor |= i < k ? b.w[i] : 0u;
  }
  below = lowor != 0 || (s && (x0 & ((1u << s) - 1u)) != 0);
  const uint64_t lo = (uint64_t)x0 | ((uint64_t)x1 << 32);
  return s ? (lo >> s) | ((uint64_t)x2 << (64 - s)) : lo;
}

// ------------------------------------------------------------------ powers of ten, 128 bits
// 10^s = F 2^g (1 + d), F in [2^127, 2^128) truncated (0 <= d < 2^-127), for TMIN <= s <= TMAX:
// the fast paths below multiply by F and bracket the exact product between F and F + 1.
constexpr int TMIN = -365, TMAX = 345, TCOUNT = TMAX - TMIN + 1;
struct P10 {
  uint64_t hi, lo;
  int g, pad;
};
// one thread per s, exact with the 1024-bit integer: 5^s (s >= 0) or floor(2^K / 5^-s)
__global__ void k_pow10_table(P10* __restrict__ tab) {
  const int idx = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= TCOUNT) return;
  const int s = TMIN + idx;
  Big b;
  bool sticky = false;
  int g;
  if (s >= 0) {
    big_set(b, 1);
    big_mul_pow5(b, s);
    const int L = big_bitlen(b);
    if (L > 128) big_shr(b, L - 128, sticky);
    else big_shl(b, 128 - L);
    g = s + L - 128;
  } else {
    Big f;
    big_set(f, 1);
    big_mul_pow5(f, -s);
    const int K = 128 + big_bitlen(f);
    big_set(b, 1);
    big_shl(b, K);
    big_div_pow5(b, -s, sticky);
    const int L = big_bitlen(b);  // 128 or 129
    if (L > 128) big_shr(b, L - 128, sticky);
    g = -K + s + (L - 128);
  }
  P10 r;
  r.hi = ((uint64_t)b.w[3] << 32) | b.w[2];
  r.lo = ((uint64_t)b.w[1] << 32) | b.w[0];
  r.g = g;
  r.pad = 0;
  tab[idx] = r;
}
// p2:p1:p0 = a * (fh:fl)
__device__ __forceinline__ void mul_64x128(uint64_t a, uint64_t fh, uint64_t fl, uint64_t& p2, uint64_t& p1,
                                           uint64_t& p0) {
  p0 = a * fl;
  const uint64_t h0 = __umul64hi(a, fl), l1 = a * fh, h1 = __umul64hi(a, fh);
  p1 = l1 + h0;
  p2 = h1 + (p1 < l1 ? 1u : 0u);
}
__device__ __forceinline__ void add_64(uint64_t& p2, uint64_t& p1, uint64_t& p0, uint64_t a) {
  p0 += a;
  const uint64_t c0 = p0 < a ? 1u : 0u;
  p1 += c0;
  p2 += (c0 && p1 == 0) ? 1u : 0u;
}
// bits [a, a + 64) of p2:p1:p0 (0 <= a < 192; bits above 191 read as zero)
__device__ __forceinline__ uint64_t bits3(uint64_t p2, uint64_t p1, uint64_t p0, int a) {
  const int k = a >> 6, r = a & 63;
  const uint64_t x0 = k == 0 ? p0 : (k == 1 ? p1 : p2);
  const uint64_t x1 = k == 0 ? p1 : (k == 1 ? p2 : 0u);
  return r ? (x0 >> r) | (x1 << (64 - r)) : x0;
}
// any of bits [0, b) set (0 <= b <= 192)
__device__ __forceinline__ bool low3(uint64_t p2, uint64_t p1, uint64_t p0, int b) {
  auto part = [](uint64_t x, int n) { return n >= 64 ? x != 0 : (n <= 0 ? false : (x & ((1ull << n) - 1)) != 0); };
  return part(p0, b) || part(p1, b - 64) || part(p2, b - 128);
}
__device__ __forceinline__ int bitlen3(uint64_t p2, uint64_t p1, uint64_t p0) {
  return p2 ? 192 - __clzll(p2) : (p1 ? 128 - __clzll(p1) : (p0 ? 64 - __clzll(p0) : 0));
}

// ------------------------------------------------------------------ formatting
// Rust's `{:.Pe}` / `{:+.Pe}` of x into out; returns the length.  Finite nonzero values
// get P+1 significant digits rounded half to even; the exponent is printed without '+'
// or padding; NaN is "NaN" whatever its sign, infinities "inf", zero "0.000e0".
__constant__ uint64_t POW10_U64[20] = {1ull, 10ull, 100ull, 1000ull, 10000ull, 100000ull, 1000000ull, 10000000ull,
                                       100000000ull, 1000000000ull, 10000000000ull, 100000000000ull,
                                       1000000000000ull, 10000000000000ull, 100000000000000ull,
                                       1000000000000000ull, 10000000000000000ull, 100000000000000000ull,
                                       1000000000000000000ull, 10000000000000000000ull};
// P <= 18: the P+1 digits D of M 2^E 10^s, s = P - e10, from one 64 x 128 multiply.  The
// exact product lies in [M F, M F + M) 2^(E + g); when the integer part and the side of 1/2
// are the same at both ends, the rounding (half to even; a tie cannot lie strictly inside)
// is decided.  e10 may be one low: then the integer part has P+2 digits and s drops by one.
__device__ bool fmt_fast(uint64_t M, int E, int P, const P10* __restrict__ tab, int& e10, uint64_t& D) {
  int ee = e10;
  for (int pass = 0; pass < 2; ++pass) {
    const int s = P - ee;
    if (s < TMIN || s > TMAX) return false;
    const P10 F = tab[s - TMIN];
    const int sh = -(E + F.g);  // integer part = product >> sh
    if (sh < 2 || sh > 190) return false;
    uint64_t p2, p1, p0;
    mul_64x128(M, F.hi, F.lo, p2, p1, p0);
    const bool big = sh + 64 < 192 && bits3(p2, p1, p0, sh + 64) != 0;  // integer part >= 2^64
    const uint64_t N = bits3(p2, p1, p0, sh);
    if (big || N >= POW10_U64[P + 1]) {  // e10 was one low
      if (pass) return false;
      ++ee;
      continue;
    }
    const bool hb = (bits3(p2, p1, p0, sh - 1) & 1u) != 0, rest = low3(p2, p1, p0, sh - 1);
    uint64_t q2 = p2, q1 = p1, q0 = p0;
    add_64(q2, q1, q0, M);
    const uint64_t N2 = bits3(q2, q1, q0, sh);
    const bool hb2 = (bits3(q2, q1, q0, sh - 1) & 1u) != 0, rest2 = low3(q2, q1, q0, sh - 1);
    if (N2 == N && (!hb2 || !rest2)) {  // the top end is at most N + 1/2: the value lies below
      D = N;
    } else if (hb && rest && (N2 == N || (N2 == N + 1 && !hb2 && !rest2))) {
      D = N + 1;
    } else {
      return false;
    }
    if (D == POW10_U64[P + 1]) {  // 9.99..9 rounded up to 10.00..0
      D = POW10_U64[P];
      ++ee;
    }
    e10 = ee;
    return true;
  }
  return false;
}

__device__ int fmt_exp(double x, int P, bool plus, char* out, const P10* __restrict__ tab) {
  int n = 0;
  const uint64_t bits = (uint64_t)__double_as_longlong(x);
  const bool neg = bits >> 63;
  const int bexp = (int)((bits >> 52) & 0x7FF);
  uint64_t mant = bits & ((1ull << 52) - 1);
  if (bexp == 0x7FF && mant) {
    out[0] = 'N'; out[1] = 'a'; out[2] = 'N';
    return 3;
  }
  if (neg) out[n++] = '-';
  else if (plus) out[n++] = '+';
  if (bexp == 0x7FF) {
    out[n] = 'i'; out[n + 1] = 'n'; out[n + 2] = 'f';
    return n + 3;
  }
  int e10 = 0;
  const int d0 = n;  // first digit position
  if (bexp == 0 && mant == 0) {
    out[n++] = '0';
    if (P > 0) {
      out[n++] = '.';
      for (int i = 0; i < P; ++i) out[n++] = '0';
    }
  } else {
    int E;
    if (bexp) {
      mant |= 1ull << 52;
      E = bexp - 1075;
    } else {
      E = -1074;
    }
    const int e2 = 63 - __clzll(mant) + E;
    e10 = (int)floor(e2 * 0.30102999566398119521);  // floor(log10|x|) or one less
    uint64_t D;
    if (P <= 18 && fmt_fast(mant, E, P, tab, e10, D)) {
      const int total = P + 1;
      n = d0 + (P > 0 ? total + 1 : 1);
      if (P > 0) out[d0 + 1] = '.';
      for (int k = total - 1; k >= 0; --k) {
        out[k == 0 ? d0 : d0 + 1 + k] = (char)('0' + D % 10u);
        D /= 10u;
      }
    } else {
    Big t, q, p10;
    big_pow10(p10, P + 1);
    bool sticky;
    for (int pass = 0; pass < 2; ++pass) {
      const int s = P - e10;
      big_set(t, mant);
      sticky = false;
      const int A = E + s + 1;  // t = floor(2 |x| 10^s)
      if (s > 0) big_mul_pow5(t, s);
      if (A > 0) big_shl(t, A);
      if (s < 0) big_div_pow5(t, -s, sticky);
      if (A < 0) big_shr(t, -A, sticky);
      q = t;
      bool dummy = false;
      big_shr(q, 1, dummy);
      if (big_cmp(q, p10) >= 0) {
        ++e10;
        continue;
      }
      break;
    }
    const bool rbit = t.w[0] & 1u;
    if (rbit && (sticky || (q.w[0] & 1u))) big_inc(q);
    const bool carry = big_cmp(q, p10) == 0;  // 9.99..9 rounded up to 10.00..0
    if (carry) {
      ++e10;
      out[n++] = '1';
      if (P > 0) {
        out[n++] = '.';
        for (int i = 0; i < P; ++i) out[n++] = '0';
      }
    } else {
      // digits k = 0..P at out[d0] and out[d0 + 1 + k] (k >= 1), least significant first
      const int total = P + 1;
      n = d0 + (P > 0 ? total + 1 : 1);
      if (P > 0) out[d0 + 1] = '.';
      int k = total - 1;
      while (k >= 0) {
        uint32_t g = big_div<P10_9>(q);
        for (int j = 0; j < 9 && k >= 0; ++j, --k) {
          const char c = (char)('0' + g % 10u);
          g /= 10u;
          out[k == 0 ? d0 : d0 + 1 + k] = c;
        }
      }
    }
    }
  }
  out[n++] = 'e';
  const unsigned ae = (unsigned)(e10 < 0 ? -e10 : e10);  // at most 324
  if (e10 < 0) out[n++] = '-';
  if (ae >= 100) out[n++] = (char)('0' + ae / 100u);
  if (ae >= 10) out[n++] = (char)('0' + ae / 10u % 10u);
  out[n++] = (char)('0' + ae % 10u);
  return n;
}

enum FmtKind { F_F64 = 0, F_DD_DISPLAY = 1, F_MF4_DISPLAY = 2, F_DD_SERDE = 3, F_MF4_SERDE = 4 };

template <int KIND>
__device__ int format_elem(const double* const* pl, long e, int prec, bool plus, char* out, const P10* tab) {
  if (KIND == F_F64) return fmt_exp(pl[0][e], prec, plus, out, tab);
  if (KIND == F_DD_DISPLAY) return fmt_exp(pl[0][e] + pl[1][e], 32, false, out, tab);
  if (KIND == F_MF4_DISPLAY) return fmt_exp(pl[0][e] + pl[1][e], 64, false, out, tab);
  if (KIND == F_DD_SERDE) {
    int n = fmt_exp(pl[0][e], 18, true, out, tab);
    out[n++] = '+';
    return n + fmt_exp(pl[1][e], 18, true, out + n, tab);
  }
  int n = 0;  // F_MF4_SERDE
  for (int i = 0; i < 4; ++i) {
    if (i) out[n++] = '+';
    n += fmt_exp(pl[i][e], 14, true, out + n, tab);
  }
  return n;
}

struct Planes {
  const double* p[4];
};

// one element per thread, its string staged in LDS, then the block's 64 fixed-size slots
// stored with coalesced 16-byte writes
template <int KIND, int SLOT>
__global__ __launch_bounds__(64) void k_format(Planes pl, long n, int prec, int plus, char* __restrict__ slots,
                                               uint32_t* __restrict__ lens, const P10* __restrict__ tab) {
  __shared__ __align__(16) char buf[64 * SLOT];
  const long e = (long)blockIdx.x * 64 + threadIdx.x;
  if (e < n) lens[e] = (uint32_t)format_elem<KIND>(pl.p, e, prec, plus != 0, buf + threadIdx.x * SLOT, tab);
  __syncthreads();
  const long base = (long)blockIdx.x * 64 * SLOT;
  const long nb = (n - (long)blockIdx.x * 64 < 64 ? n - (long)blockIdx.x * 64 : 64) * SLOT;
  for (long o = threadIdx.x * 16L; o < nb; o += 64 * 16)
    *reinterpret_cast<uint4*>(slots + base + o) = *reinterpret_cast<const uint4*>(buf + o);
}

// exclusive scan of the lengths into offsets (n + 1 entries): per-block sums of 1024
// lengths, a one-block scan of those, then each block's own scan with its carry-in
__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ len, long n,
                                                     uint64_t* __restrict__ bsum) {
  __shared__ uint64_t red[256];
  const long b0 = (long)blockIdx.x * 1024;
  uint64_t s = 0;
  for (int j = 0; j < 4; ++j) {
    const long i = b0 + threadIdx.x * 4 + j;
    if (i < n) s += len[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}
// exclusive scan of the nb block sums in place (bsum[nb] = total): one block of 1024
// threads, each a contiguous run of the sums, the run totals scanned through LDS
__global__ __launch_bounds__(1024) void k_scan_sums(uint64_t* __restrict__ bsum, long nb) {
  __shared__ uint64_t part[1024];
  const int t = threadIdx.x;
  const long per = (nb + 1023) / 1024;
  const long b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
  uint64_t own = 0;
  for (long i = b0; i < b1; ++i) own += bsum[i];
  part[t] = own;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
    const uint64_t add = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  uint64_t run = part[t] - own;
  for (long i = b0; i < b1; ++i) {
    const uint64_t v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (t == 1023) bsum[nb] = part[1023];
}
__global__ __launch_bounds__(256) void k_scan_final(const uint32_t* __restrict__ len, long n,
                                                    const uint64_t* __restrict__ bsum, long nb,
                                                    uint64_t* __restrict__ off) {
  __shared__ uint64_t sc[256];
  const long b0 = (long)blockIdx.x * 1024;
  uint64_t v[4], s = 0;
  for (int j = 0; j < 4; ++j) {
    const long i = b0 + threadIdx.x * 4 + j;
    v[j] = i < n ? len[i] : 0;
    s += v[j];
  }
  sc[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {  // inclusive Hillis-Steele scan
    const uint64_t add = (int)threadIdx.x >= d ? sc[threadIdx.x - d] : 0;
    __syncthreads();
    sc[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t run = bsum[blockIdx.x] + sc[threadIdx.x] - s;
  for (int j = 0; j < 4; ++j) {
    const long i = b0 + threadIdx.x * 4 + j;
    if (i < n) off[i] = run;
    run += v[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) off[n] = bsum[nb];
}

// the block's 64 slots, packed at their offsets: consecutive lanes write consecutive bytes
template <int SLOT>
__global__ __launch_bounds__(256) void k_compact(const char* __restrict__ slots, const uint64_t* __restrict__ off,
                                                 long n, char* __restrict__ text) {
  __shared__ __align__(16) char buf[64 * SLOT];
  __shared__ uint64_t so[65];
  const long e0 = (long)blockIdx.x * 64;
  const int cnt = (int)(n - e0 < 64 ? n - e0 : 64);
  for (long o = threadIdx.x * 16L; o < (long)cnt * SLOT; o += 256 * 16)
    *reinterpret_cast<uint4*>(buf + o) = *reinterpret_cast<const uint4*>(slots + e0 * SLOT + o);
  if ((int)threadIdx.x <= cnt) so[threadIdx.x] = off[e0 + threadIdx.x];
  __syncthreads();
  const uint64_t o0 = so[0], o1 = so[cnt];
  for (uint64_t j = threadIdx.x; j < o1 - o0; j += 256) {
    int lo = 0, hi = cnt - 1;  // last element whose offset <= o0 + j
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (so[mid] - o0 <= j) lo = mid;
      else hi = mid - 1;
    }
    text[o0 + j] = buf[lo * SLOT + (int)(j - (so[lo] - o0))];
  }
}

// ------------------------------------------------------------------ parsing
__device__ __forceinline__ bool is_digit(char c) { return c >= '0' && c <= '9'; }
__device__ __forceinline__ char lower(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c + 32) : c; }
__device__ bool eq_nocase(const char* s, long len, const char* word, int wl) {
  if (len != wl) return false;
  for (int i = 0; i < wl; ++i)
    if (lower(s[i]) != word[i]) return false;
  return true;
}

// the double nearest m64 2^(E - 63) (+ sticky below m64's last bit; bit 63 of m64 set),
// half to even, with subnormals and overflow
__device__ double round_m64(uint64_t m64, bool sticky, int E, bool neg) {
  uint64_t out;
  if (E > 1023) {
    out = 0x7FF0000000000000ull;
  } else if (E >= -1022) {
    uint64_t m = m64 >> 11;
    const uint64_t rest = m64 & 0x7FF, half = 0x400;
    if (rest > half || (rest == half && (sticky || (m & 1)))) ++m;
    int Eb = E;
    if (m == (1ull << 53)) {
      m >>= 1;
      ++Eb;
    }
    out = Eb > 1023 ? 0x7FF0000000000000ull : ((uint64_t)(Eb + 1023) << 52) | (m & ((1ull << 52) - 1));
  } else {
    const int sh = -1011 - E;  // value / 2^-1074 = m64 / 2^sh, sh >= 12
    uint64_t m;
    if (sh > 64) {
      m = 0;
    } else if (sh == 64) {
      m = (m64 > (1ull << 63) || (m64 == (1ull << 63) && sticky)) ? 1 : 0;
    } else {
      m = m64 >> sh;
      const uint64_t rest = m64 & ((1ull << sh) - 1), half = 1ull << (sh - 1);
      if (rest > half || (rest == half && (sticky || (m & 1)))) ++m;
    }
    out = m;  // m == 2^52 encodes the smallest normal
  }
  if (neg) out |= 1ull << 63;
  return __longlong_as_double((long long)out);
}

// round T 2^e2off (+ sticky, below T's last bit) to a double, half to even
__device__ double big_to_double(const Big& t, int e2off, bool sticky, bool neg) {
  const int L = big_bitlen(t);
  uint64_t m64;
  if (L <= 64) {
    const uint64_t v = (uint64_t)t.w[0] | ((uint64_t)t.w[1] << 32);
    m64 = v << (64 - L);
  } else {
    bool below;
    m64 = big_bits64(t, L - 64, below);
    sticky |= below;
  }
  return round_m64(m64, sticky, L - 1 + e2off, neg);
}
// p2:p1:p0 2^g (p nonzero) rounded to a double
__device__ double p3_to_double(uint64_t p2, uint64_t p1, uint64_t p0, int g, bool neg) {
  const int L = bitlen3(p2, p1, p0);
  uint64_t m64;
  bool sticky = false;
  if (L <= 64) {
    m64 = p0 << (64 - L);
  } else {
    m64 = bits3(p2, p1, p0, L - 64);
    sticky = low3(p2, p1, p0, L - 64);
  }
  return round_m64(m64, sticky, L - 1 + g, neg);
}
// w 10^q for w < 2^64, q in the table: the roundings of w F 2^g and (w F + w) 2^g, which
// bracket the exact value, agree on every value but those within 2^-120 or so of a halfway
// point; then that is the answer (false: take the exact path)
__device__ bool parse_fast(uint64_t w, int q, bool neg, const P10* __restrict__ tab, double& out) {
  if (q < TMIN || q > TMAX) return false;
  const P10 F = tab[q - TMIN];
  uint64_t p2, p1, p0;
  mul_64x128(w, F.hi, F.lo, p2, p1, p0);
  const double lo = p3_to_double(p2, p1, p0, F.g, neg);
  add_64(p2, p1, p0, w);
  const double hi = p3_to_double(p2, p1, p0, F.g, neg);
  if (__double_as_longlong(lo) != __double_as_longlong(hi)) return false;
  out = lo;
  return true;
}

// w 10^q rounded to a double (w nonzero, |q| within the cut-offs of parse_f64)
__device__ double decimal_to_double(const Big& w, int q, bool neg) {
  Big t = w;
  bool sticky = false;
  if (q >= 0) {
    big_mul_pow5(t, q);
    return big_to_double(t, q, false, neg);
  }
  const int bl5 = (int)(-q * 2.321928094887362) + 1;  // >= bitlen(5^-q)
  int A = 66 + bl5 - big_bitlen(w);
  if (A < 0) A = 0;
  big_shl(t, A);
  big_div_pow5(t, -q, sticky);
  return big_to_double(t, q - A, sticky, neg);
}

// Rust's str::parse::<f64>: [+-] (inf | infinity | nan, any case) or
// [+-] digits [. digits] [(e|E) [+-] digits] with at least one mantissa digit, nothing else.
// status: 0 ok, 1 syntax error, 2 more than 40 significant digits needed to round.
__device__ double parse_f64(const char* s, long len, int& status, const P10* __restrict__ tab) {
  status = 0;
  long i = 0;
  bool neg = false;
  if (len > 0 && (s[0] == '+' || s[0] == '-')) {
    neg = s[0] == '-';
    i = 1;
  }
  if (i >= len) {
    status = 1;
    return 0.0;
  }
  const char* r = s + i;
  const long rl = len - i;
  if (eq_nocase(r, rl, "inf", 3) || eq_nocase(r, rl, "infinity", 8))
    return neg ? -__longlong_as_double(0x7FF0000000000000ll) : __longlong_as_double(0x7FF0000000000000ll);
  if (eq_nocase(r, rl, "nan", 3))
    return __longlong_as_double((long long)((neg ? 0xFFF8000000000000ull : 0x7FF8000000000000ull)));
  Big w;
  big_set(w, 0);
  uint64_t w64 = 0;  // the same digits while there are at most 19
  int nsig = 0;
  long dexp = 0;
  bool trunc = false, any = false;
  for (; i < len && is_digit(s[i]); ++i) {
    const uint32_t d = (uint32_t)(s[i] - '0');
    any = true;
    if (nsig < MAX_SIG) {
      if (nsig > 0 || d) {
        big_mul_add(w, 10u, d);
        w64 = w64 * 10u + d;
        ++nsig;
      }
    } else {
      ++dexp;
      trunc |= d != 0;
    }
  }
  if (i < len && s[i] == '.') {
    for (++i; i < len && is_digit(s[i]); ++i) {
      const uint32_t d = (uint32_t)(s[i] - '0');
      any = true;
      if (nsig < MAX_SIG) {
        if (nsig > 0 || d) {
          big_mul_add(w, 10u, d);
          w64 = w64 * 10u + d;
          ++nsig;
        }
        --dexp;
      } else {
        trunc |= d != 0;
      }
    }
  }
  if (!any) {
    status = 1;
    return 0.0;
  }
  if (i < len && (s[i] == 'e' || s[i] == 'E')) {
    ++i;
    bool eneg = false;
    if (i < len && (s[i] == '+' || s[i] == '-')) {
      eneg = s[i] == '-';
      ++i;
    }
    if (i >= len || !is_digit(s[i])) {
      status = 1;
      return 0.0;
    }
    long ev = 0;
    for (; i < len && is_digit(s[i]); ++i) ev = ev < 100000000 ? ev * 10 + (s[i] - '0') : ev;
    dexp += eneg ? -ev : ev;
  }
  if (i != len) {
    status = 1;
    return 0.0;
  }
  const double zero = neg ? -0.0 : 0.0;
  if (nsig == 0) return zero;
  const long qeff = dexp + nsig - 1;  // decimal exponent of the leading digit
  if (qeff > 310) return neg ? -__longlong_as_double(0x7FF0000000000000ll) : __longlong_as_double(0x7FF0000000000000ll);
  if (qeff < -326) return zero;
  if (!trunc && nsig <= 19) {
    double r;
    if (parse_fast(w64, (int)dexp, neg, tab, r)) return r;
  }
  const double r1 = decimal_to_double(w, (int)dexp, neg);
  if (trunc) {
    Big w1 = w;
    big_inc(w1);
    const double r2 = decimal_to_double(w1, (int)dexp, neg);
    if (__double_as_longlong(r1) != __double_as_longlong(r2)) status = 2;
  }
  return r1;
}

// Unicode White_Space as UTF-8 (what str::trim removes): length of the one at s[i], or 0
__device__ int ws_fwd(const char* s, long len, long i) {
  const unsigned char c = (unsigned char)s[i];
  if (c == 0x20 || (c >= 0x09 && c <= 0x0D)) return 1;
  if (i + 1 < len && c == 0xC2 && ((unsigned char)s[i + 1] == 0x85 || (unsigned char)s[i + 1] == 0xA0)) return 2;
  if (i + 2 < len) {
    const unsigned char c1 = (unsigned char)s[i + 1], c2 = (unsigned char)s[i + 2];
    if (c == 0xE1 && c1 == 0x9A && c2 == 0x80) return 3;
    if (c == 0xE2 && c1 == 0x80 && (c2 <= 0x8A || c2 == 0xA8 || c2 == 0xA9 || c2 == 0xAF) && c2 >= 0x80) return 3;
    if (c == 0xE2 && c1 == 0x81 && c2 == 0x9F) return 3;
    if (c == 0xE3 && c1 == 0x80 && c2 == 0x80) return 3;
  }
  return 0;
}
// length of the White_Space character ending at s[end - 1], or 0
__device__ int ws_back(const char* s, long beg, long end) {
  for (int k = 1; k <= 3; ++k) {
    if (end - k < beg) break;
    const unsigned char c = (unsigned char)s[end - k];
    if (k < 3 && (c & 0xC0) == 0x80) continue;  // continuation byte: look further back
    return ws_fwd(s, end, end - k) == k ? k : 0;
  }
  return 0;
}

// Rust's i32::from_str: [+-] digits, in range
__device__ bool parse_i32(const char* s, long len, int& out) {
  long i = 0;
  bool neg = false;
  if (len > 0 && (s[0] == '+' || s[0] == '-')) {
    neg = s[0] == '-';
    i = 1;
  }
  if (i >= len) return false;
  long long v = 0;
  for (; i < len; ++i) {
    if (!is_digit(s[i])) return false;
    v = v * 10 + (s[i] - '0');
    if (v > 2147483648ll) return false;
  }
  if (!neg && v > 2147483647ll) return false;
  out = (int)(neg ? -v : v);
  return true;
}

// ---- value types for FromStr
struct DdOps {
  using T = dd;
  __device__ static T zero() { return mkdd(0.0); }
  __device__ static T from(double x) { return mkdd(x); }
  __device__ static T mul(T a, T b) { return dd_mul(a, b); }
  __device__ static T add(T a, T b) { return dd_add(a, b); }
  __device__ static T neg(T a) { return dd_neg(a); }
  struct Div {
    __device__ Div() {}
    __device__ T operator()(T a) const { return dd_div(a, mkdd(10.0)); }
  };
  __device__ static bool same(T a, T b) {
    return (__double_as_longlong(a.hi) == __double_as_longlong(b.hi) || (a.hi != a.hi && b.hi != b.hi)) &&
           (__double_as_longlong(a.lo) == __double_as_longlong(b.lo) || (a.lo != a.lo && b.lo != b.lo));
  }
  __device__ static void store(T v, double* const* o, long e) {
    o[0][e] = v.hi;
    o[1][e] = v.lo;
  }
};

struct Mf4Ops {
  using T = mf4;
  __device__ static T zero() { return mkmf(0.0); }
  __device__ static T from(double x) { return mkmf(x); }
  __device__ static T mul(T a, T b) { return mf_mul(a, b); }
  __device__ static T add(T a, T b) { return mf_add(a, b); }
  __device__ static T neg(T a) { return mf_neg(a); }
  struct Div {  // a / 10: the divisor is finite and nonzero, so div.rs:49-77 reduces to
    mf4 r;      // a == 0 -> ZERO, a non-finite -> from_f64(a0 / 10), else a * recip(10)
    __device__ Div() : r(mf_recip(mkmf(10.0))) {}
    __device__ T operator()(T a) const {
      if (a.l0 == 0.0) return mkmf(0.0);
      if (!isfinite(a.l0)) return mkmf(a.l0 / 10.0);
      return mf_mul(a, r);
    }
  };
  __device__ static bool limb_same(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b) || (a != a && b != b);
  }
  __device__ static bool same(T a, T b) {
    return limb_same(a.l0, b.l0) && limb_same(a.l1, b.l1) && limb_same(a.l2, b.l2) && limb_same(a.l3, b.l3);
  }
  __device__ static void store(T v, double* const* o, long e) {
    o[0][e] = v.l0;
    o[1][e] = v.l1;
    o[2][e] = v.l2;
    o[3][e] = v.l3;
  }
};

// FromStr of dd.rs:371-435 / multifloat/fmt.rs:32-114, op for op.  The exponent loops stop
// early once a step leaves the value unchanged (the step is a fixed function of the value,
// so every later step would too): a huge exponent drives the value to inf/NaN or zero within
// about 700 steps.  An exponent of i32::MIN runs no division steps, as the reference's
// `0..(-exp)` does when -exp wraps.
template <class Ops>
__device__ bool from_str(const char* s, long len, typename Ops::T& out) {
  using T = typename Ops::T;
  long b = 0, e = len;
  for (int k; b < e && (k = ws_fwd(s, e, b)) > 0;) b += k;
  for (int k; e > b && (k = ws_back(s, b, e)) > 0;) e -= k;
  if (b >= e) return false;
  bool negative = false;
  if (s[b] == '-' || s[b] == '+') {
    negative = s[b] == '-';
    ++b;
  }
  if (b >= e) return false;
  long m_end = e;
  for (long i = b; i < e; ++i)
    if (s[i] == 'e' || s[i] == 'E') {
      m_end = i;
      break;
    }
  if (m_end == b) return false;
  const T ten = Ops::from(10.0);
  const typename Ops::Div div10;
  T acc = Ops::zero(), pow10_neg = Ops::from(1.0);
  bool saw_dot = false, any_digit = false;
  for (long i = b; i < m_end; ++i) {
    const char c = s[i];
    if (c == '.') {
      if (saw_dot) return false;
      saw_dot = true;
      continue;
    }
    if (!is_digit(c)) return false;
    any_digit = true;
    const T digit = Ops::from((double)(c - '0'));
    if (!saw_dot) {
      acc = Ops::add(Ops::mul(acc, ten), digit);
    } else {
      pow10_neg = div10(pow10_neg);
      acc = Ops::add(acc, Ops::mul(digit, pow10_neg));
    }
  }
  if (!any_digit) return false;
  if (m_end < e) {
    int ex;
    if (!parse_i32(s + m_end + 1, e - m_end - 1, ex)) return false;
    if (ex > 0) {
      for (int k = 0; k < ex; ++k) {
        const T nx = Ops::mul(acc, ten);
        if (Ops::same(nx, acc)) break;
        acc = nx;
      }
    } else if (ex < 0 && ex != (-2147483647 - 1)) {
      for (int k = 0; k < -ex; ++k) {
        const T nx = div10(acc);
        if (Ops::same(nx, acc)) break;
        acc = nx;
      }
    }
  }
  out = negative ? Ops::neg(acc) : acc;
  return true;
}

enum ParseKind { P_F64 = 0, P_DD_FROMSTR = 1, P_MF4_FROMSTR = 2, P_DD_SERDE = 3, P_MF4_SERDE = 4 };

struct OutPlanes {
  double* p[4];
};

template <int KIND>
__global__ __launch_bounds__(64) void k_parse(const char* __restrict__ text, const uint64_t* __restrict__ off, long n,
                                              OutPlanes o, uint8_t* __restrict__ status,
                                              const P10* __restrict__ tab) {
  const long e = (long)blockIdx.x * 64 + threadIdx.x;
  if (e >= n) return;
  const char* s = text + off[e];
  const long len = (long)(off[e + 1] - off[e]);
  int st = 0;
  if (KIND == P_F64) {
    o.p[0][e] = parse_f64(s, len, st, tab);
  } else if (KIND == P_DD_FROMSTR) {
    dd v = mkdd(0.0);
    if (!from_str<DdOps>(s, len, v)) st = 1;
    DdOps::store(v, o.p, e);
  } else if (KIND == P_MF4_FROMSTR) {
    mf4 v = mkmf(0.0);
    if (!from_str<Mf4Ops>(s, len, v)) st = 1;
    Mf4Ops::store(v, o.p, e);
  } else if (KIND == P_DD_SERDE) {
    // dd.rs:496-511: the first '+' at 1 <= i <= len-2 followed by a sign and not preceded
    // by 'e'/'E' splits hi from lo
    long split = -1;
    for (long i = 1; i + 1 < len; ++i)
      if (s[i] == '+' && (s[i + 1] == '+' || s[i + 1] == '-') && s[i - 1] != 'e' && s[i - 1] != 'E') {
        split = i;
        break;
      }
    double hi = 0.0, lo = 0.0;
    if (split < 0) {
      st = 1;
    } else {
      int s1, s2;
      hi = parse_f64(s, split, s1, tab);
      lo = parse_f64(s + split + 1, len - split - 1, s2, tab);
      st = s1 == 1 || s2 == 1 ? 1 : (s1 | s2);
    }
    o.p[0][e] = st == 1 ? 0.0 : hi;
    o.p[1][e] = st == 1 ? 0.0 : lo;
  } else {  // P_MF4_SERDE, traits.rs:104-148
    double limb[4] = {0.0, 0.0, 0.0, 0.0};
    int count = 0;
    long start = 0, i = 1;
    while (i < len && st != 1) {
      const bool boundary = s[i] == '+' && i + 1 < len && (s[i + 1] == '+' || s[i + 1] == '-') &&
                            s[i - 1] != 'e' && s[i - 1] != 'E';
      if (boundary) {
        if (count >= 4) {
          st = 1;
          break;
        }
        int sl;
        const double v = parse_f64(s + start, i - start, sl, tab);
        if (sl == 1) {
          st = 1;
          break;
        }
        st |= sl;
#pragma unroll
        for (int k = 0; k < 4; ++k) limb[k] = k == count ? v : limb[k];
        ++count;
        start = i + 1;
        i = start + 1;
      } else {
        ++i;
      }
    }
    if (st != 1) {
      if (count >= 4) {
        st = 1;
      } else {
        int sl;
        const double v = parse_f64(s + start, len - start, sl, tab);
        if (sl == 1) st = 1;
        else st |= sl;
#pragma unroll
        for (int k = 0; k < 4; ++k) limb[k] = k == count ? v : limb[k];
        ++count;
        if (count != 4) st = 1;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) o.p[k][e] = st == 1 ? 0.0 : limb[k];
  }
  status[e] = (uint8_t)st;
}

__global__ void k_count_bad(const uint8_t* __restrict__ status, long n, unsigned long long* __restrict__ bad) {
  unsigned long long c = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    c += status[i] != 0;
  if (c) atomicAdd(bad, c);
}
__global__ void k_zero_u64(unsigned long long* p) { *p = 0; }

unsigned nblk64(long n) { return (unsigned)((n + 63) / 64); }

int slot_of(int kind) {
  switch (kind) {
    case F_F64: return 80;
    case F_DD_DISPLAY: return 48;
    case F_MF4_DISPLAY: return 80;
    case F_DD_SERDE: return 64;
    default: return 96;
  }
}
int limbs_of_fmt(int kind) { return kind == F_F64 ? 1 : (kind == F_MF4_DISPLAY || kind == F_MF4_SERDE) ? 4 : 2; }
int limbs_of_parse(int kind) { return kind == P_F64 ? 1 : (kind == P_MF4_FROMSTR || kind == P_MF4_SERDE) ? 4 : 2; }

template <int KIND, int SLOT>
void launch_format(const Planes& pl, long n, int prec, int plus, char* slots, uint32_t* lens, const P10* tab,
                   hipStream_t s) {
  hipLaunchKernelGGL((k_format<KIND, SLOT>), dim3(nblk64(n)), dim3(64), 0, s, pl, n, prec, plus, slots, lens, tab);
}

}  // namespace

// Device-side format: planes of n values -> text (capacity cap) + n + 1 offsets.
// need = offsets[n]; HP_ESHAPE when cap < need (offsets are written either way).
int format_dev(int kind, const double* const* planes, long n, int prec, int plus, char* text, size_t cap,
               uint64_t* off, size_t* need, hipStream_t s) {
  if (kind == F_F64 && (prec < 0 || prec > 64)) {
    set_error("precision %d outside 0..64", prec);
    return HP_EINVAL;
  }
  if (n == 0) {
    HP_HIP(hipMemsetAsync(off, 0, sizeof(uint64_t), s));
    HP_HIP(hipStreamSynchronize(s));
    *need = 0;
    return HP_OK;
  }
  const int slot = slot_of(kind);
  const long nb = (n + 1023) / 1024;
  DevBuf slots, lens, bsum, tabb;
  int rc;
  if ((rc = slots.alloc((size_t)nblk64(n) * 64 * slot)) || (rc = lens.alloc(n * sizeof(uint32_t))) ||
      (rc = bsum.alloc((nb + 1) * sizeof(uint64_t))) || (rc = tabb.alloc(TCOUNT * sizeof(P10))))
    return rc;
  const P10* ptab = static_cast<const P10*>(tabb.p);
  hipLaunchKernelGGL(k_pow10_table, dim3((TCOUNT + 63) / 64), dim3(64), 0, s, static_cast<P10*>(tabb.p));
  HP_LAUNCH_CHECK();
  Planes pl{};
  for (int k = 0; k < limbs_of_fmt(kind); ++k) pl.p[k] = planes[k];
  char* sl = static_cast<char*>(slots.p);
  uint32_t* ln = static_cast<uint32_t*>(lens.p);
  switch (kind) {
    case F_F64: launch_format<F_F64, 80>(pl, n, prec, plus, sl, ln, ptab, s); break;
    case F_DD_DISPLAY: launch_format<F_DD_DISPLAY, 48>(pl, n, prec, plus, sl, ln, ptab, s); break;
    case F_MF4_DISPLAY: launch_format<F_MF4_DISPLAY, 80>(pl, n, prec, plus, sl, ln, ptab, s); break;
    case F_DD_SERDE: launch_format<F_DD_SERDE, 64>(pl, n, prec, plus, sl, ln, ptab, s); break;
    case F_MF4_SERDE: launch_format<F_MF4_SERDE, 96>(pl, n, prec, plus, sl, ln, ptab, s); break;
    default: set_error("unknown format kind %d", kind); return HP_EINVAL;
  }
  HP_LAUNCH_CHECK();
  uint64_t* bs = static_cast<uint64_t*>(bsum.p);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(256), 0, s, ln, n, bs);
  HP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, bs, nb);
  HP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(256), 0, s, ln, n, bs, nb, off);
  HP_LAUNCH_CHECK();
  uint64_t total = 0;
  HP_HIP(hipMemcpyAsync(&total, off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  *need = total;
  if (!text || cap < total) {
    set_error("text buffer holds %zu bytes, the strings need %llu", cap, (unsigned long long)total);
    return HP_ESHAPE;
  }
  switch (slot) {
    case 48: hipLaunchKernelGGL(k_compact<48>, dim3(nblk64(n)), dim3(256), 0, s, sl, off, n, text); break;
    case 64: hipLaunchKernelGGL(k_compact<64>, dim3(nblk64(n)), dim3(256), 0, s, sl, off, n, text); break;
    case 80: hipLaunchKernelGGL(k_compact<80>, dim3(nblk64(n)), dim3(256), 0, s, sl, off, n, text); break;
    default: hipLaunchKernelGGL(k_compact<96>, dim3(nblk64(n)), dim3(256), 0, s, sl, off, n, text); break;
  }
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

// Device-side parse: text + n + 1 offsets -> planes and per-element status (0 ok, 1 syntax,
// 2 needs more than 40 significant digits); *bad = number of elements with status != 0.
int parse_dev(int kind, const char* text, const uint64_t* off, long n, double* const* planes, uint8_t* status,
              size_t* bad, hipStream_t s) {
  *bad = 0;
  if (n == 0) return HP_OK;
  OutPlanes o{};
  for (int k = 0; k < limbs_of_parse(kind); ++k) o.p[k] = planes[k];
  DevBuf tabb;
  int rc0;
  if ((rc0 = tabb.alloc(TCOUNT * sizeof(P10)))) return rc0;
  const P10* ptab = static_cast<const P10*>(tabb.p);
  hipLaunchKernelGGL(k_pow10_table, dim3((TCOUNT + 63) / 64), dim3(64), 0, s, static_cast<P10*>(tabb.p));
  HP_LAUNCH_CHECK();
  switch (kind) {
    case P_F64: hipLaunchKernelGGL(k_parse<P_F64>, dim3(nblk64(n)), dim3(64), 0, s, text, off, n, o, status, ptab); break;
    case P_DD_FROMSTR:
      hipLaunchKernelGGL(k_parse<P_DD_FROMSTR>, dim3(nblk64(n)), dim3(64), 0, s, text, off, n, o, status, ptab);
      break;
    case P_MF4_FROMSTR:
      hipLaunchKernelGGL(k_parse<P_MF4_FROMSTR>, dim3(nblk64(n)), dim3(64), 0, s, text, off, n, o, status, ptab);
      break;
    case P_DD_SERDE:
      hipLaunchKernelGGL(k_parse<P_DD_SERDE>, dim3(nblk64(n)), dim3(64), 0, s, text, off, n, o, status, ptab);
      break;
    case P_MF4_SERDE:
      hipLaunchKernelGGL(k_parse<P_MF4_SERDE>, dim3(nblk64(n)), dim3(64), 0, s, text, off, n, o, status, ptab);
      break;
    default: set_error("unknown parse kind %d", kind); return HP_EINVAL;
  }
  HP_LAUNCH_CHECK();
  DevBuf cnt;
  int rc;
  if ((rc = cnt.alloc(sizeof(unsigned long long)))) return rc;
  auto* c = static_cast<unsigned long long*>(cnt.p);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, c);
  hipLaunchKernelGGL(k_count_bad, dim3(std::min<unsigned>(nblk64(n), 1024u)), dim3(256), 0, s, status, n, c);
  HP_LAUNCH_CHECK();
  unsigned long long nb = 0;
  HP_HIP(hipMemcpyAsync(&nb, c, sizeof(nb), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  *bad = (size_t)nb;
  return HP_OK;
}

}  // namespace hp

using namespace hp;

namespace {

int host_format(int kind, const double* v, size_t n, int prec, int plus, char* text, size_t cap, size_t* offsets) {
  if (!offsets || (n && !v)) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  const int limbs = limbs_of_fmt(kind);
  hipStream_t s = hipStreamPerThread;
  DevBuf dv, doff, dtext;
  int rc;
  if ((rc = dv.alloc(limbs * n * sizeof(double))) || (rc = doff.alloc((n + 1) * sizeof(uint64_t)))) return rc;
  if (n && (rc = upload_planes(v, n, limbs, dv.d(), s))) return rc;
  const double* planes[4] = {dv.d(), dv.d() + n, dv.d() + 2 * n, dv.d() + 3 * n};
  if ((rc = dtext.alloc(std::max<size_t>(cap, 1)))) return rc;
  size_t need = 0;
  rc = format_dev(kind, planes, (long)n, prec, plus, text ? static_cast<char*>(dtext.p) : nullptr, text ? cap : 0,
                  static_cast<uint64_t*>(doff.p), &need, s);
  if (rc != HP_OK && rc != HP_ESHAPE) return rc;
  static_assert(sizeof(size_t) == sizeof(uint64_t), "offsets are 64-bit");
  HP_HIP(hipMemcpy(offsets, doff.p, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (rc == HP_ESHAPE) return rc;
  if (need) HP_HIP(hipMemcpy(text, dtext.p, need, hipMemcpyDeviceToHost));
  return HP_OK;
}

int host_parse(int kind, const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  if (!offsets || !out || (n && !text)) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  if (n == 0) return HP_OK;
  for (size_t i = 0; i < n; ++i)
    if (offsets[i + 1] < offsets[i]) {
      set_error("offsets decrease at %zu", i);
      return HP_EINVAL;
    }
  const int limbs = limbs_of_parse(kind);
  const size_t bytes = offsets[n];
  hipStream_t s = hipStreamPerThread;
  DevBuf dtext, doff, dout, dst;
  int rc;
  if ((rc = dtext.alloc(bytes)) || (rc = doff.alloc((n + 1) * sizeof(uint64_t))) ||
      (rc = dout.alloc(limbs * n * sizeof(double))) || (rc = dst.alloc(n)))
    return rc;
  if (bytes) HP_HIP(hipMemcpyAsync(dtext.p, text, bytes, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(doff.p, offsets, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  double* planes[4] = {dout.d(), dout.d() + n, dout.d() + 2 * n, dout.d() + 3 * n};
  size_t bad = 0;
  uint8_t* dstat = static_cast<uint8_t*>(dst.p);
  if ((rc = parse_dev(kind, static_cast<const char*>(dtext.p), static_cast<const uint64_t*>(doff.p), (long)n, planes,
                      dstat, &bad, s)))
    return rc;
  if ((rc = download_planes(dout.d(), n, limbs, out, s))) return rc;
  std::vector<uint8_t> st(n);
  HP_HIP(hipMemcpy(st.data(), dstat, n, hipMemcpyDeviceToHost));
  if (status) std::copy(st.begin(), st.end(), status);
  for (size_t i = 0; i < n; ++i)
    if (st[i]) {
      set_error(st[i] == 1 ? "invalid literal at element %zu" : "element %zu needs more than 40 significant digits to round",
                i);
      return HP_EINVAL;
    }
  return HP_OK;
}

}  // namespace

extern "C" {

int hp_format_f64(const double* v, size_t n, int precision, int force_sign, char* text, size_t cap, size_t* offsets) {
  return host_format(F_F64, v, n, precision, force_sign, text, cap, offsets);
}
int hp_dd_to_string(const double* v, size_t n, char* text, size_t cap, size_t* offsets) {
  return host_format(F_DD_DISPLAY, v, n, 32, 0, text, cap, offsets);
}
int hp_mf4_to_string(const double* v, size_t n, char* text, size_t cap, size_t* offsets) {
  return host_format(F_MF4_DISPLAY, v, n, 64, 0, text, cap, offsets);
}
int hp_dd_serialize_str(const double* v, size_t n, char* text, size_t cap, size_t* offsets) {
  return host_format(F_DD_SERDE, v, n, 18, 1, text, cap, offsets);
}
int hp_mf4_serialize_str(const double* v, size_t n, char* text, size_t cap, size_t* offsets) {
  return host_format(F_MF4_SERDE, v, n, 14, 1, text, cap, offsets);
}
int hp_parse_f64(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  return host_parse(P_F64, text, offsets, n, out, status);
}
int hp_dd_from_str(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  return host_parse(P_DD_FROMSTR, text, offsets, n, out, status);
}
int hp_mf4_from_str(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  return host_parse(P_MF4_FROMSTR, text, offsets, n, out, status);
}
int hp_dd_deserialize_str(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  return host_parse(P_DD_SERDE, text, offsets, n, out, status);
}
int hp_mf4_deserialize_str(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status) {
  return host_parse(P_MF4_SERDE, text, offsets, n, out, status);
}
int hp_dev_format(int kind, const double* const* planes, size_t n, int precision, int force_sign, char* text,
                  size_t cap, size_t* offsets, size_t* need, void* stream) {
  if (!planes || !offsets || !need) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  return format_dev(kind, planes, (long)n, precision, force_sign, text, cap, reinterpret_cast<uint64_t*>(offsets), need,
                    (hipStream_t)stream);
}
int hp_dev_parse(int kind, const char* text, const size_t* offsets, size_t n, double* const* planes, uint8_t* status,
                 size_t* n_bad, void* stream) {
  if (!planes || !offsets || !status || !n_bad) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  return parse_dev(kind, text, reinterpret_cast<const uint64_t*>(offsets), (long)n, planes, status, n_bad,
                   (hipStream_t)stream);
}

}  // extern "C"
