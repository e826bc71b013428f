// Radix-2 FFTs over the Fr domain for G1, G2, GT and the cyclotomic subgroup (gfx950).
//
// Semantics of weighted_btx/src/fft.rs:156-390 (Radix2Domain): the natural-order transform
// X_k = sum_j x_j w^(jk) with w = ROOT_OF_UNITY^(2^(32 - log n)) (w^-1 for the inverse), computed
// as a bit-reversal permutation followed by log n decimation-in-time stages; butterfly
// (u, v) -> (u + v t, u - v t) with the group's scalar action for "v t"; the scaled inverse
// multiplies every output by n^-1.  For GT and cyclotomic values the group law is Fp12
// multiplication, the inverse is conjugation and the scalar action is the decomposed plan of
// cyclo.hpp (fft.rs:437-526, weighted_pfe/src/fft.rs:324-365).
//
// Every stage is one launch with one thread per butterfly; twiddles w^j (j < n/2) and their
// cyclotomic plans are built once per (size, direction) on the host and kept in device memory.
// Group butterflies multiply by the twiddle through the curve endomorphisms: k = k1 + k2 lambda
// on G1 (phi, lambda = |z|^2 - 1) and k = k1 + k2 |z|^2 on G2 (psi^2), a joint double-and-add of
// at most 128 doublings instead of 255.
// Butterflies are numbered twiddle-major: the lanes of a wavefront take butterflies that share
// one twiddle (same exponent, so the same branch pattern of the double-and-add or NAF chain)
// whenever a stage has at least 64 butterflies per twiddle, instead of 64 different exponents
// whose union of add steps every lane would wait through.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "coop12.hpp"
#include "cyclo.hpp"
#include "host_field.hpp"

extern "C" {
int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream);
int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream);
}

namespace bte {

// ---------------------------------------------------------------------------------------------
// Host: domain twiddles and cyclotomic scalar plans
// ---------------------------------------------------------------------------------------------
static const uint64_t X_ABS = 0xd201000000010000ull;

// weighted_btx/src/fft.rs:543-567: canonical scalar -> four base-|x| digits.
static void decompose_base_x(const bteh::Fr& s, uint64_t out[4]) {
  uint64_t limbs[4] = {s.l[0], s.l[1], s.l[2], s.l[3]};
  for (int c = 0; c < 4; ++c) {
    unsigned __int128 rem = 0;
    for (int i = 3; i >= 0; --i) {
      unsigned __int128 d = (rem << 64) | limbs[i];
      limbs[i] = (uint64_t)(d / X_ABS);
      rem = d % X_ABS;
    }
    out[c] = (uint64_t)rem;
  }
}

// weighted_btx/src/fft.rs:569-594: width-4 NAF of a u64 (65 digits for the carry into bit 64).
static int wnaf4(uint64_t value, int8_t digits[CYC_DIGITS]) {
  unsigned __int128 v = value;
  int pos = 0;
  for (int i = 0; i < CYC_DIGITS; ++i) digits[i] = 0;
  while (v != 0) {
    if (v & 1) {
      int d = (int)(v & 15);
      if (d >= 8) d -= 16;
      digits[pos] = (int8_t)d;
      if (d > 0) v -= (unsigned)d; else v += (unsigned)(-d);
    }
    v >>= 1;
    ++pos;
  }
  return pos;
}

static cyc_plan make_plan(const bteh::Fr& s) {
  cyc_plan p;
  memset(&p, 0, sizeof(p));
  uint64_t comps[4];
  decompose_base_x(s, comps);
  int count = 0;
  for (int c = 0; c < 4; ++c) {
    int n = wnaf4(comps[c], p.naf[c]);
    if (n > count) count = n;
  }
  p.count = (uint8_t)count;
  return p;
}

// fft.rs:164-191: generator of the size-2^logn subgroup (or its inverse).
static bteh::Fr domain_generator(int logn, bool inverse) {
  bteh::Fr g;
  memcpy(g.l, bteh::FR_ROOT_OF_UNITY, 32);
  for (int i = logn; i < bteh::FR_S; ++i) g = bteh::fr_mul(g, g);
  return inverse ? bteh::fr_inv(g) : g;
}

// k = q d + r for a canonical k (four u64 limbs) and 2^64 < d < 2^128; q and r must be below
// 2^128 (true for k < r_Fr and the divisors used here, |z|^2 - 1 and |z|^2).
static void divmod_u128(const uint64_t k[4], unsigned __int128 d, unsigned __int128& q, unsigned __int128& rem) {
  q = 0;
  rem = 0;
  for (int i = 255; i >= 0; --i) {
    const bool carry = (rem >> 127) != 0;
    rem = (rem << 1) | ((k[i >> 6] >> (i & 63)) & 1);
    if (carry || rem >= d) {
      rem -= d;
      q |= (unsigned __int128)1 << i;  // i < 128 whenever this fires (q < 2^128)
    }
  }
}

// Twiddle decompositions for the endomorphism-accelerated group butterflies: words 0..7 hold
// (k mod lambda, k div lambda) with lambda = |z|^2 - 1 (G1, phi), words 8..15 hold
// (k mod |z|^2, k div |z|^2) (G2, psi^2); each half as four little-endian u32.
static void twiddle_decomposition(const bteh::Fr& k, uint32_t out[16]) {
  const unsigned __int128 z2 = (unsigned __int128)X_ABS * X_ABS;
  const unsigned __int128 divs[2] = {z2 - 1, z2};
  for (int h = 0; h < 2; ++h) {
    unsigned __int128 q, r;
    divmod_u128(k.l, divs[h], q, r);
    for (int i = 0; i < 4; ++i) {
      out[8 * h + i] = (uint32_t)(r >> (32 * i));
      out[8 * h + 4 + i] = (uint32_t)(q >> (32 * i));
    }
  }
}

// G1 half only: (k mod lambda, k div lambda) as eight little-endian u32, for g1_glv_mul.
void glv_decompose_g1(const bteh::Fr& k, uint32_t out[8]) {
  uint32_t both[16];
  twiddle_decomposition(k, both);
  memcpy(out, both, 8 * sizeof(uint32_t));
}

// Width-3 NAF of a value below 2^128 as four-bit digits, least significant first (digit i in bits
// 4 (i mod 8) of word i / 8); returns the highest nonzero position (-1 for zero).
// A group FFT stage uses the NAF twiddle form when every twiddle of the stage serves at least a
// wavefront of butterflies (twiddle-major numbering then gives each wavefront one digit pattern);
// otherwise lanes with different twiddles would take both of the NAF step's additions, and the
// binary form (at most one addition per step) is faster.
static bool naf_stage(size_t butterflies, int log_half) { return butterflies >> log_half >= 64; }

static int wnaf3(unsigned __int128 k, uint32_t* words) {
  int top = -1;
  for (int i = 0; i < TW_NAF_BLOCKS * 4; ++i) words[i] = 0;
  for (int i = 0; i < TW_NAF_LEN; ++i) {
    int di = 0;
    if (k & 1) {
      di = (int)(k & 7);
      if (di >= 4) di -= 8;
      if (di > 0) k -= (unsigned)di;
      else k += (unsigned)(-di);
    }
    words[i >> 3] |= ((uint32_t)di & 15u) << (4 * (i & 7));
    if (di) top = i;
    k >>= 1;
  }
  return top;
}

static tw_naf twiddle_naf(const bteh::Fr& k) {
  uint32_t dec[16];
  twiddle_decomposition(k, dec);
  tw_naf t;
  memset(&t, 0, sizeof(t));
  for (int g = 0; g < 2; ++g) {
    int top = -1;
    for (int c = 0; c < 2; ++c) {
      unsigned __int128 v = 0;
      for (int i = 3; i >= 0; --i) v = (v << 32) | dec[8 * g + 4 * c + i];
      top = std::max(top, wnaf3(v, t.d[g][c]));
    }
    t.top[g] = (int16_t)top;
  }
  return t;
}

struct DomainTables {
  uint32_t* dec = nullptr;       // n/2 twiddle decompositions (16 x u32 each)
  tw_naf* naf = nullptr;         // the same as width-3 NAF digits
  uint32_t* twiddles = nullptr;  // n/2 canonical scalars (8 x u32 each): w^j
  cyc_plan* plans = nullptr;     // n/2 plans of w^j
  uint32_t* size_inv = nullptr;  // n^-1 (8 x u32)
  cyc_plan* size_inv_plan = nullptr;
};

static std::mutex g_dom_mu;
static std::map<std::tuple<int, int, int>, DomainTables> g_domains;  // (device, logn, inverse)

// Device copy of a host array into a new allocation; on failure nothing stays allocated.
template <class T> static hipError_t upload_new(T** dst, const T* src, size_t count) {
  *dst = nullptr;
  if (test_alloc_should_fail()) return hipErrorOutOfMemory;
  hipError_t e = hipMalloc((void**)dst, count * sizeof(T));
  if (e != hipSuccess) {
    *dst = nullptr;
    return e;
  }
  e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(*dst);
    *dst = nullptr;
  }
  return e;
}

// Tables are built into locals and published in the cache only once every allocation and copy
// has succeeded, so a failure (for example an allocation refused under memory pressure) leaves
// no half-filled entry behind for later calls to read.
static int domain_tables(int logn, bool inverse, bool want_plans, DomainTables** out, hipStream_t s) {
  int dev = 0;
  BTE_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_dom_mu);
  DomainTables& t = g_domains[std::make_tuple(dev, logn, inverse ? 1 : 0)];
  const size_t half = logn ? ((size_t)1 << (logn - 1)) : 0;
  if (!t.twiddles) {
    std::vector<bteh::Fr> tw(half ? half : 1);
    bteh::Fr w = domain_generator(logn, inverse);
    tw[0] = bteh::fr_from_u64(1);
    for (size_t j = 1; j < half; ++j) tw[j] = bteh::fr_mul(tw[j - 1], w);
    bteh::Fr ninv = bteh::fr_inv(bteh::fr_from_u64((uint64_t)1 << logn));
    std::vector<uint32_t> dec(tw.size() * 16);
    for (size_t j = 0; j < tw.size(); ++j) twiddle_decomposition(tw[j], &dec[16 * j]);
    std::vector<tw_naf> naf(tw.size());
    for (size_t j = 0; j < tw.size(); ++j) naf[j] = twiddle_naf(tw[j]);
    DomainTables nt;
    hipError_t e = upload_new(&nt.twiddles, (const uint32_t*)tw.data(), tw.size() * 8);
    if (e == hipSuccess) e = upload_new(&nt.dec, (const uint32_t*)dec.data(), dec.size());
    if (e == hipSuccess) e = upload_new(&nt.naf, (const tw_naf*)naf.data(), naf.size());
    if (e == hipSuccess) e = upload_new(&nt.size_inv, (const uint32_t*)ninv.l, 8);
    if (e != hipSuccess) {
      for (void* p : {(void*)nt.twiddles, (void*)nt.dec, (void*)nt.naf, (void*)nt.size_inv})
        if (p) (void)hipFree(p);
      return set_error(BTE_ENOMEM, "fft domain tables (2^%d): %s", logn, hipGetErrorString(e));
    }
    t.dec = nt.dec;
    t.naf = nt.naf;
    t.size_inv = nt.size_inv;
    t.twiddles = nt.twiddles;  // published last: it marks the entry as built
  }
  if (want_plans && !t.plans) {
    std::vector<cyc_plan> pl(half ? half : 1);
    bteh::Fr w = domain_generator(logn, inverse);
    bteh::Fr cur = bteh::fr_from_u64(1);
    for (size_t j = 0; j < half; ++j) {
      pl[j] = make_plan(cur);
      cur = bteh::fr_mul(cur, w);
    }
    cyc_plan np = make_plan(bteh::fr_inv(bteh::fr_from_u64((uint64_t)1 << logn)));
    cyc_plan *plans = nullptr, *inv_plan = nullptr;
    hipError_t e = upload_new(&plans, (const cyc_plan*)pl.data(), pl.size());
    if (e == hipSuccess) e = upload_new(&inv_plan, (const cyc_plan*)&np, 1);
    if (e != hipSuccess) {
      if (plans) (void)hipFree(plans);
      return set_error(BTE_ENOMEM, "fft domain plans (2^%d): %s", logn, hipGetErrorString(e));
    }
    t.size_inv_plan = inv_plan;
    t.plans = plans;  // published last
  }
  (void)s;
  *out = &t;
  return BTE_OK;
}

static int log2_exact(size_t n) {
  if (n == 0 || (n & (n - 1))) return -1;
  int l = 0;
  while (((size_t)1 << l) < n) ++l;
  return l > bteh::FR_S ? -1 : l;
}

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------
BTE_DEV size_t bitrev(size_t i, int bits) {
  return bits ? (size_t)(__builtin_bitreverse64((uint64_t)i) >> (64 - bits)) : 0;
}

// fft.rs:621-632, over `count` transforms of size n stored back to back.
template <class T> __global__ void k_bit_reverse(T* v, size_t n, int logn, size_t count) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= n * count) return;
  T* b = v + (g >> logn) * n;
  const size_t i = g & (n - 1);
  const size_t r = bitrev(i, logn);
  if (i < r) {
    T a = b[i];
    b[i] = b[r];
    b[r] = a;
  }
}

// Butterfly g of one stage (butterfly size 2 * 2^log_half) over `count` transforms of size n,
// twiddle-major: the nb * count butterflies (nb blocks per transform) of twiddle k are
// consecutive.  Returns the index of the butterfly's first element; k receives the twiddle.
BTE_DEV size_t butterfly_slot(size_t g, size_t n, int log_half, size_t count, size_t& k) {
  const size_t half = (size_t)1 << log_half;
  const size_t nb = n >> (log_half + 1);
  const size_t S = nb * count;
  k = g / S;
  const size_t r = g - k * S;
  const size_t b = r / nb, j = r - b * nb;
  return b * n + j * 2 * half + k;
}

// One DIT stage for butterfly size 2 * 2^log_half; twiddle index k << (logn - 1 - log_half).
template <class F, class TW>
__global__ void __launch_bounds__(64) k_group_stage(jac<F>* __restrict__ v, size_t n, int logn, int log_half,
                                                    const TW* __restrict__ tw, size_t count) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= (n / 2) * count) return;
  const size_t half = (size_t)1 << log_half;
  size_t k;
  const size_t i = butterfly_slot(g, n, log_half, count, k);
  const jac<F> u = v[i];
  jac<F> w = v[i + half];
  if (k) w = twiddle_mul(w, twiddle_at(tw, k << (logn - 1 - log_half)));
  v[i] = jac_add(u, w);
  v[i + half] = jac_sub(u, w);
}

template <class F>
__global__ void __launch_bounds__(64) k_group_scale(jac<F>* __restrict__ v, size_t n, const uint32_t* __restrict__ s) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = group_scalar_mul(v[i], s);
}

// Cyclotomic stages come in two forms.  One thread per butterfly (k_cyc_stage, ...) is the
// faster per element once a stage has enough butterflies to fill the GPU; below that, each
// element runs on three lanes (coop12.hpp, *3 kernels: 21 butterflies per 64-lane workgroup,
// `tab` holding the per-lane exponentiation table, C3_TAB_BYTES per lane of the
// grid), which shortens the per-stage latency of small transforms.  A stage with fewer than
// BTE_TUNE_CYC_THREE_LANE_MAX butterflies (default 32768) takes the three-lane form.
static constexpr int C3_PER_WG = 21;

BTE_DEV c3tab c3_tab_for(uint32_t* tab) {
  return {tab, (size_t)gridDim.x * 64, blockIdx.x * (size_t)64 + (threadIdx.x & 63)};
}

__global__ void __launch_bounds__(64) k_cyc_stage3(fp12* __restrict__ v, size_t n, int logn, int log_half,
                                                   const cyc_plan* __restrict__ plans, size_t count,
                                                   uint32_t* __restrict__ tab) {
  __shared__ uint32_t xbuf[48 * 64];
  const c3ctx c = c3_make(xbuf);
  const size_t g = blockIdx.x * (size_t)C3_PER_WG + c.lane / 3;
  if (c.lane >= 3 * C3_PER_WG || g >= (n / 2) * count) return;
  const size_t half = (size_t)1 << log_half;
  size_t k;
  const size_t i = butterfly_slot(g, n, log_half, count, k);
  const fp4 u = c3_load(c, v + i);
  fp4 w = c3_load(c, v + i + half);
  if (k) w = c3_cyclo_pow_plan(c, c3_tab_for(tab), w, plans + (k << (logn - 1 - log_half)));
  const fp4 lo = c3_mul(c, u, w);
  const fp4 hi = c3_mul(c, u, c3_conj(c, w));
  c3_store(c, v + i, lo);
  c3_store(c, v + i + half, hi);
}

__global__ void __launch_bounds__(64) k_cyc_prefix_last3(fp12* __restrict__ v, size_t n, size_t count,
                                                         const cyc_plan* __restrict__ plans, size_t ntrans,
                                                         uint32_t* __restrict__ tab) {
  __shared__ uint32_t xbuf[48 * 64];
  const c3ctx c = c3_make(xbuf);
  const size_t g = blockIdx.x * (size_t)C3_PER_WG + c.lane / 3;
  if (c.lane >= 3 * C3_PER_WG || g >= count * ntrans) return;
  const size_t k = g / ntrans;  // twiddle-major
  v += (g - k * ntrans) * n;
  const size_t half = n / 2;
  fp4 w = c3_load(c, v + half + k);
  if (k) w = c3_cyclo_pow_plan(c, c3_tab_for(tab), w, plans + k);
  const fp4 out = c3_mul(c, c3_load(c, v + k), w);
  c3_store(c, v + k, out);
}

static unsigned c3_blocks(size_t nbutterflies) { return (unsigned)((nbutterflies + C3_PER_WG - 1) / C3_PER_WG); }

__global__ void __launch_bounds__(64) k_cyc_stage(fp12* __restrict__ v, size_t n, int logn, int log_half,
                                                  const cyc_plan* __restrict__ plans, size_t count) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= (n / 2) * count) return;
  const size_t half = (size_t)1 << log_half;
  size_t k;
  const size_t i = butterfly_slot(g, n, log_half, count, k);
  const fp12 u = v[i];
  fp12 w = v[i + half];
  if (k) w = cyclo_pow_plan(w, plans + (k << (logn - 1 - log_half)));
  v[i] = fp12_mul(u, w);
  v[i + half] = fp12_mul(u, fp12_conj(w));
}

// Last stage of the prefix inverse transform (fft.rs:322-349): only outputs k < count.
__global__ void __launch_bounds__(64) k_cyc_prefix_last(fp12* __restrict__ v, size_t n, size_t count,
                                                        const cyc_plan* __restrict__ plans, size_t ntrans) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= count * ntrans) return;
  const size_t k = g / ntrans;  // twiddle-major
  v += (g - k * ntrans) * n;
  const size_t half = n / 2;
  fp12 w = v[half + k];
  if (k) w = cyclo_pow_plan(w, plans + k);
  v[k] = fp12_mul(v[k], w);
}

__global__ void __launch_bounds__(64) k_cyc_scale(fp12* __restrict__ v, size_t n, const cyc_plan* __restrict__ plan) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = cyclo_pow_plan(v[i], plan);
}

__global__ void __launch_bounds__(64) k_cyc_pow_small(fp12* __restrict__ v, size_t n, const int64_t* __restrict__ c) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = cyclo_pow_small(v[i], c[i]);
}

__global__ void __launch_bounds__(64) k_cyc_pow_plans(fp12* __restrict__ v, size_t n, const cyc_plan* __restrict__ plans) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = cyclo_pow_plan(v[i], plans + i);
}

// ---------------------------------------------------------------------------------------------
// Drivers
// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// Small group transforms as one batched MSM.  A stage of the radix-2 transform is one thread per
// butterfly running a whole scalar multiplication (a serial chain of ~128 doublings), so a size-n
// transform costs log n such chains end to end whatever the GPU's width: 42 ms for G1 at 2^8.
// Written as n^2 scalar products X_k = sum_j (w^(jk)) x_j, one MSM segment per output, the same
// transform is one pass of the batched Pippenger kernels, whose serial part is one window
// combination per output: more arithmetic (n / log n times the butterflies' multiplications) but
// in parallel, so it wins while n^2 count stays small.  The outputs are the same points (the
// transform is a linear map); their Jacobian coordinates may differ from the staged form.
// ---------------------------------------------------------------------------------------------
struct DftScalars {
  uint8_t* w = nullptr;  // n canonical scalars (w^j, w^-j, or w^-j n^-1), 32 bytes each
};
static std::mutex g_dft_mu;
static std::map<std::tuple<int, int, int>, DftScalars> g_dft;  // (device, logn, 0 fwd / 1 inv / 2 inv scaled)
static std::atomic<int> g_dft_max_log{10};                     // sizes up to 2^10 (bte_fft_set_dft_max_log)
static constexpr size_t DFT_MAX_TERMS = (size_t)1 << 20;

static int dft_scalars(int logn, int kind, const uint8_t** out) {
  int dev = 0;
  BTE_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_dft_mu);
  DftScalars& t = g_dft[std::make_tuple(dev, logn, kind)];
  if (!t.w) {
    const size_t n = (size_t)1 << logn;
    const bteh::Fr w = domain_generator(logn, kind != 0);
    bteh::Fr cur = kind == 2 ? bteh::fr_inv(bteh::fr_from_u64(n)) : bteh::fr_from_u64(1);
    std::vector<bteh::Fr> v(n);
    for (size_t j = 0; j < n; ++j) {
      v[j] = cur;
      cur = bteh::fr_mul(cur, w);
    }
    uint8_t* d = nullptr;
    const hipError_t e = upload_new(&d, (const uint8_t*)v.data(), 32 * n);
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "dft scalars (2^%d): %s", logn, hipGetErrorString(e));
    t.w = d;
  }
  *out = t.w;
  return BTE_OK;
}

// Term t = (b n + k) n + j of `count` transforms: point b n + j with scalar index j k mod n, in
// segment b n + k; seg_off[s] = s n.
__global__ void k_dft_terms(int logn, size_t count, uint32_t* __restrict__ pidx, uint32_t* __restrict__ sidx,
                            uint64_t* __restrict__ seg_off) {
  const size_t n = (size_t)1 << logn;
  const size_t total = n * n * count;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t > total) return;
  if (t == total) {
    seg_off[n * count] = total;
    return;
  }
  const size_t seg = t >> logn, j = t & (n - 1), k = seg & (n - 1), b = seg >> logn;
  pidx[t] = (uint32_t)(b * n + j);
  sidx[t] = (uint32_t)((j * k) & (n - 1));
  if (j == 0) seg_off[seg] = t;
}

template <class F> static bool dft_by_msm(int logn, size_t count) {
  const size_t n = (size_t)1 << logn;
  return logn >= 2 && logn <= g_dft_max_log && n * n * count <= DFT_MAX_TERMS;
}

template <class F>
static int group_dft_msm(jac<F>* v, int logn, size_t count, bool inverse, bool scaled, hipStream_t s) {
  const size_t n = (size_t)1 << logn, N = n * count, terms = n * N;
  CallScope sc(s);
  aff<F>* A;
  uint32_t *pidx, *sidx;
  uint64_t* off;
  const uint8_t* w;
  BTE_TRY_RC(dft_scalars(logn, inverse ? (scaled ? 2 : 1) : 0, &w));
  BTE_TRY_RC(sc.dev(&A, N));
  BTE_TRY_RC(sc.dev(&pidx, terms));
  BTE_TRY_RC(sc.dev(&sidx, terms));
  BTE_TRY_RC(sc.dev(&off, N + 1));
  if (sizeof(F) == sizeof(fp))
    BTE_TRY_RC(bte_g1_batch_normalize((const bte_p1*)v, N, (bte_p1_affine*)A, s));
  else
    BTE_TRY_RC(bte_g2_batch_normalize((const bte_p2*)v, N, (bte_p2_affine*)A, s));
  hipLaunchKernelGGL(k_dft_terms, dim3(blocks_for(terms + 1, 256)), dim3(256), 0, s, logn, count, pidx, sidx, off);
  BTE_CHECK_LAUNCH();
  if (sizeof(F) == sizeof(fp))
    return g1_msm_batched_gather(A, N, pidx, w, sidx, terms, off, N, 0, v, s);
  return g2_msm_batched_gather(A, N, pidx, w, sidx, terms, off, N, 0, v, s);
}

template <class F>
static int group_fft(jac<F>* v, size_t n, size_t count, bool inverse, bool scaled, hipStream_t s) {
  const int logn = log2_exact(n);
  if (logn < 0) return set_error(BTE_EINVAL, "fft: InvalidBatchSize(%zu)", n);
  if (count == 0) return BTE_OK;
  if (!v) return set_error(BTE_EINVAL, "fft: null values");
  if (n == 1) return BTE_OK;  // a one-point transform is the identity map (n^-1 = 1)
  if (dft_by_msm<F>(logn, count)) return group_dft_msm<F>(v, logn, count, inverse, scaled, s);
  DomainTables* t;
  int rc = domain_tables(logn, inverse, false, &t, s);
  if (rc) return rc;
  PhaseTimer pt(s);
  pt.begin();
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bit_reverse<jac<F>>), dim3(blocks_for(n * count, 256)), dim3(256), 0, s, v, n,
                     logn, count);
  for (int lh = 0; lh < logn; ++lh) {
    if (naf_stage(n / 2 * count, lh))
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_stage<F, tw_naf>), dim3(blocks_for(n / 2 * count, 64)), dim3(64), 0,
                         s, v, n, logn, lh, (const tw_naf*)t->naf, count);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_stage<F, uint32_t>), dim3(blocks_for(n / 2 * count, 64)), dim3(64),
                         0, s, v, n, logn, lh, (const uint32_t*)t->dec, count);
  }
  pt.mark("stages");
  if (inverse && scaled)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_scale<F>), dim3(blocks_for(n * count, 64)), dim3(64), 0, s, v,
                       n * count, (const uint32_t*)t->size_inv);
  pt.mark("scale");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

// Bytes of per-lane odd-power storage the cyclotomic stages of `count` transforms of size n
// need (zero when they run one thread per butterfly).
size_t cyc_tab_bytes(size_t n, size_t count, size_t three_lane_max) {
  const size_t nbf = n / 2 * count;
  return nbf && nbf < three_lane_max ? (size_t)c3_blocks(nbf) * 64 * C3_TAB_BYTES : 0;
}

static size_t three_lane_max() { return (size_t)tune_get(BTE_TUNE_CYC_THREE_LANE_MAX); }

// mode 0: full transform; mode 1: prefix inverse (last stage only for `out_count` outputs).
// ext_tab: caller-owned storage of cyc_tab_bytes(n, count, c3max) bytes, or null to allocate here;
// stages with fewer than c3max butterflies take the three-lane form.
static int cyclotomic_fft(fp12* v, size_t n, size_t count, bool inverse, bool scaled, size_t out_count, int mode,
                          hipStream_t s, uint32_t* ext_tab = nullptr, size_t c3max = three_lane_max()) {
  const int logn = log2_exact(n);
  if (logn < 0) return set_error(BTE_EINVAL, "fft: InvalidBatchSize(%zu)", n);
  if (count == 0) return BTE_OK;
  if (!v) return set_error(BTE_EINVAL, "fft: null values");
  if (mode == 1 && out_count > n / 2) return set_error(BTE_EINVAL, "prefix fft: output_count > n/2");
  if (n == 1) return BTE_OK;
  DomainTables* t;
  int rc = domain_tables(logn, inverse, true, &t, s);
  if (rc) return rc;
  const size_t nbf = n / 2 * count;
  const bool three = nbf < c3max;
  CallScope sc(s);
  uint32_t* tab = ext_tab;  // exponentiation tables, C3_TAB_BYTES per lane
  if (three && !tab) BTE_TRY_RC(sc.dev((uint8_t**)&tab, (size_t)c3_blocks(nbf) * 64 * C3_TAB_BYTES));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bit_reverse<fp12>), dim3(blocks_for(n * count, 256)), dim3(256), 0, s, v, n,
                     logn, count);
  const int full_stages = mode == 1 ? logn - 1 : logn;
  for (int lh = 0; lh < full_stages; ++lh) {
    if (three)
      hipLaunchKernelGGL(k_cyc_stage3, dim3(c3_blocks(nbf)), dim3(64), 0, s, v, n, logn, lh,
                         (const cyc_plan*)t->plans, count, tab);
    else
      hipLaunchKernelGGL(k_cyc_stage, dim3(blocks_for(nbf, 64)), dim3(64), 0, s, v, n, logn, lh,
                         (const cyc_plan*)t->plans, count);
  }
  if (mode == 1 && out_count) {
    if (three)
      hipLaunchKernelGGL(k_cyc_prefix_last3, dim3(c3_blocks(out_count * count)), dim3(64), 0, s, v, n, out_count,
                         (const cyc_plan*)t->plans, count, tab);
    else
      hipLaunchKernelGGL(k_cyc_prefix_last, dim3(blocks_for(out_count * count, 64)), dim3(64), 0, s, v, n,
                         out_count, (const cyc_plan*)t->plans, count);
  }
  if (mode == 0 && inverse && scaled)
    hipLaunchKernelGGL(k_cyc_scale, dim3(blocks_for(n * count, 64)), dim3(64), 0, s, v, n * count,
                       (const cyc_plan*)t->size_inv_plan);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// ---------------------------------------------------------------------------------------------
int cyc_prefix_batched_enqueue(fp12* v, size_t n, size_t count, size_t out_count, uint32_t* tab,
                               size_t c3max, hipStream_t s) {
  if (cyc_tab_bytes(n, count, c3max) && !tab) return set_error(BTE_EINVAL, "cyclotomic prefix: missing odd-power storage");
  return cyclotomic_fft(v, n, count, true, false, out_count, 1, s, tab, c3max);
}

// Builds the twiddle tables (and cyclotomic plans) of a domain ahead of launch-only use.
int fft_warm(size_t n, bool inverse, bool plans, hipStream_t s) {
  const int logn = log2_exact(n);
  if (logn < 0) return set_error(BTE_EINVAL, "fft: InvalidBatchSize(%zu)", n);
  DomainTables* t;
  return domain_tables(logn, inverse, plans, &t, s);
}

// Segmented transforms: `count` transforms of uneven power-of-two sizes stored back to back.
// Stage lh (butterfly size 2^(lh+1)) uses twiddles w_{2h}^k, which do not depend on the
// transform size, so one table for the largest size serves all: index k << (logmax-1-lh).
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void k_bit_reverse_seg(T* v, const uint64_t* __restrict__ off, const uint8_t* __restrict__ logs,
                                  uint32_t count, uint64_t total) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const uint32_t f = seg_of(off, count, g);
  const int lg = logs[f];
  T* b = v + off[f];
  const size_t i = g - off[f];
  const size_t r = bitrev(i, lg);
  if (i < r) {
    T a = b[i];
    b[i] = b[r];
    b[r] = a;
  }
}

// Segmented stages take a host-built list of butterflies (index of the first element) in
// twiddle-major order: for each twiddle k, every block of every transform that has the stage.
// With outs != nullptr (prefix inverse transforms), a transform whose last stage this is
// contributes only its butterflies k < outs[f] (the truncated stage).
static void stage_order(const std::vector<uint64_t>& off, const std::vector<uint8_t>& logs, const uint32_t* outs,
                        int lh, std::vector<uint32_t>& order) {
  const size_t half = (size_t)1 << lh;
  for (size_t k = 0; k < half; ++k)
    for (size_t f = 0; f < logs.size(); ++f) {
      const int lg = logs[f];
      if (lg - 1 < lh) continue;
      if (outs && lg - 1 == lh) {
        if (k < outs[f]) order.push_back((uint32_t)(off[f] + k));
        continue;
      }
      const size_t nb = (size_t)1 << (lg - 1 - lh);
      for (size_t j = 0; j < nb; ++j) order.push_back((uint32_t)(off[f] + j * 2 * half + k));
    }
}

template <class F, class TW>
__global__ void __launch_bounds__(64) k_group_stage_ord(jac<F>* __restrict__ v, const uint32_t* __restrict__ order,
                                                        uint64_t nbf, const uint64_t* __restrict__ off,
                                                        uint32_t count, int logmax, int log_half,
                                                        const TW* __restrict__ tw) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= nbf) return;
  const size_t i = order[g];
  const uint32_t f = seg_of(off, count, i);
  const size_t half = (size_t)1 << log_half;
  const size_t k = (i - off[f]) & (half - 1);
  const jac<F> u = v[i];
  jac<F> w = v[i + half];
  if (k) w = twiddle_mul(w, twiddle_at(tw, k << (logmax - 1 - log_half)));
  v[i] = jac_add(u, w);
  v[i + half] = jac_sub(u, w);
}

template <class F>
__global__ void __launch_bounds__(64) k_group_scale_seg(jac<F>* __restrict__ v, const uint64_t* __restrict__ off,
                                                        uint32_t count, uint64_t total,
                                                        const uint32_t* __restrict__ inv /* count x 8 */) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const uint32_t f = seg_of(off, count, g);
  v[g] = group_scalar_mul(v[g], inv + 8 * f);
}

template <class F>
static int group_fft_segmented(jac<F>* v, const int* log_sizes, size_t count, bool inverse, bool scaled,
                               hipStream_t s) {
  if (count == 0) return BTE_OK;
  if (!v || !log_sizes) return set_error(BTE_EINVAL, "fft_segmented: null pointer");
  std::vector<uint64_t> off(count + 1);
  std::vector<uint8_t> logs(count);
  std::vector<bteh::Fr> inv(count);
  int logmax = 0;
  for (size_t f = 0; f < count; ++f) {
    if (log_sizes[f] < 0 || log_sizes[f] > 28) return set_error(BTE_EINVAL, "fft_segmented: InvalidBatchSize(2^%d)", log_sizes[f]);
    logs[f] = (uint8_t)log_sizes[f];
    logmax = std::max(logmax, log_sizes[f]);
    off[f + 1] = off[f] + ((uint64_t)1 << log_sizes[f]);
    inv[f] = bteh::fr_inv(bteh::fr_from_u64((uint64_t)1 << log_sizes[f]));
  }
  if (logmax == 0) return BTE_OK;
  const uint64_t total = off[count];
  if (total > 0xffffffffull) return set_error(BTE_EINVAL, "fft_segmented: more than 2^32 elements");
  std::vector<uint32_t> order;
  std::vector<size_t> start(logmax + 1, 0);
  for (int lh = 0; lh < logmax; ++lh) {
    start[lh] = order.size();
    stage_order(off, logs, nullptr, lh, order);
  }
  start[logmax] = order.size();
  DomainTables* t;
  int rc = domain_tables(logmax, inverse, false, &t, s);
  if (rc) return rc;
  CallScope sc(s);
  uint64_t* d_off;
  uint8_t* d_logs;
  uint32_t *d_inv, *d_order;
  BTE_TRY_RC(sc.upload(&d_off, off.data(), count + 1));
  BTE_TRY_RC(sc.upload(&d_logs, logs.data(), count));
  BTE_TRY_RC(sc.upload(&d_inv, (const uint32_t*)inv.data(), 8 * count));
  BTE_TRY_RC(sc.upload(&d_order, order.data(), order.size()));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bit_reverse_seg<jac<F>>), dim3(blocks_for(total, 256)), dim3(256), 0, s, v,
                     (const uint64_t*)d_off, (const uint8_t*)d_logs, (uint32_t)count, total);
  for (int lh = 0; lh < logmax; ++lh) {
    const uint64_t nbf = start[lh + 1] - start[lh];
    if (nbf && naf_stage(nbf, lh))
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_stage_ord<F, tw_naf>), dim3(blocks_for(nbf, 64)), dim3(64), 0, s, v,
                         (const uint32_t*)d_order + start[lh], nbf, (const uint64_t*)d_off, (uint32_t)count, logmax,
                         lh, (const tw_naf*)t->naf);
    else if (nbf)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_stage_ord<F, uint32_t>), dim3(blocks_for(nbf, 64)), dim3(64), 0, s,
                         v, (const uint32_t*)d_order + start[lh], nbf, (const uint64_t*)d_off, (uint32_t)count, logmax,
                         lh, (const uint32_t*)t->dec);
  }
  if (inverse && scaled)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_scale_seg<F>), dim3(blocks_for(total, 64)), dim3(64), 0, s, v,
                       (const uint64_t*)d_off, (uint32_t)count, total, (const uint32_t*)d_inv);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Segmented unscaled inverse cyclotomic prefix transforms (fft.rs:284-350) over transforms of
// sizes 2^logs[f] stored back to back, keeping outs[f] outputs of transform f.  At stage level
// lh a transform with lh < logs[f]-1 runs a full stage and one with lh == logs[f]-1 runs its
// truncated last stage; plans come from the inverse domain of the largest size.
__global__ void __launch_bounds__(64) k_cyc_stage_ord3(fp12* __restrict__ v, const uint32_t* __restrict__ order,
                                                       uint64_t nbf, const uint64_t* __restrict__ off,
                                                       const uint8_t* __restrict__ logs, uint32_t count, int logmax,
                                                       int log_half, const cyc_plan* __restrict__ plans,
                                                       uint32_t* __restrict__ tab) {
  __shared__ uint32_t xbuf[48 * 64];
  const c3ctx c = c3_make(xbuf);
  const uint64_t g = blockIdx.x * (uint64_t)C3_PER_WG + c.lane / 3;
  if (c.lane >= 3 * C3_PER_WG || g >= nbf) return;
  const size_t i = order[g];
  const uint32_t f = seg_of(off, count, i);
  const size_t half = (size_t)1 << log_half;
  const size_t k = (i - off[f]) & (half - 1);
  fp4 w = c3_load(c, v + i + half);
  if (k) w = c3_cyclo_pow_plan(c, c3_tab_for(tab), w, plans + (k << (logmax - 1 - log_half)));
  const fp4 u = c3_load(c, v + i);
  if (logs[f] - 1 == log_half) {  // truncated last stage: left output only
    const fp4 lo = c3_mul(c, u, w);
    c3_store(c, v + i, lo);
    return;
  }
  const fp4 lo = c3_mul(c, u, w);
  const fp4 hi = c3_mul(c, u, c3_conj(c, w));
  c3_store(c, v + i, lo);
  c3_store(c, v + i + half, hi);
}

__global__ void __launch_bounds__(64) k_cyc_stage_ord(fp12* __restrict__ v, const uint32_t* __restrict__ order,
                                                      uint64_t nbf, const uint64_t* __restrict__ off,
                                                      const uint8_t* __restrict__ logs, uint32_t count, int logmax,
                                                      int log_half, const cyc_plan* __restrict__ plans) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= nbf) return;
  const size_t i = order[g];
  const uint32_t f = seg_of(off, count, i);
  const size_t half = (size_t)1 << log_half;
  const size_t k = (i - off[f]) & (half - 1);
  const cyc_plan* pl = plans + (k << (logmax - 1 - log_half));
  fp12 w = v[i + half];
  if (k) w = cyclo_pow_plan(w, pl);
  if (logs[f] - 1 == log_half) {  // truncated last stage: left output only
    v[i] = fp12_mul(v[i], w);
    return;
  }
  const fp12 u = v[i];
  v[i] = fp12_mul(u, w);
  v[i + half] = fp12_mul(u, fp12_conj(w));
}

int cyc_prefix_ifft_segmented(fp12* v, const int* log_sizes, const uint32_t* out_counts, size_t count,
                              hipStream_t s) {
  if (count == 0) return BTE_OK;
  std::vector<uint64_t> off(count + 1);
  std::vector<uint8_t> logs(count);
  int logmax = 0;
  for (size_t f = 0; f < count; ++f) {
    if (log_sizes[f] < 1 || log_sizes[f] > 28) return set_error(BTE_EINVAL, "prefix_segmented: size 2^%d", log_sizes[f]);
    if (out_counts[f] > (1u << (log_sizes[f] - 1))) return set_error(BTE_EINVAL, "prefix_segmented: output count");
    logs[f] = (uint8_t)log_sizes[f];
    logmax = std::max(logmax, log_sizes[f]);
    off[f + 1] = off[f] + ((uint64_t)1 << log_sizes[f]);
  }
  const uint64_t total = off[count];
  if (total > 0xffffffffull) return set_error(BTE_EINVAL, "prefix_segmented: more than 2^32 elements");
  std::vector<uint32_t> order;
  std::vector<size_t> start(logmax + 1, 0);
  for (int lh = 0; lh < logmax; ++lh) {
    start[lh] = order.size();
    stage_order(off, logs, out_counts, lh, order);
  }
  start[logmax] = order.size();
  DomainTables* t;
  int rc = domain_tables(logmax, true, true, &t, s);
  if (rc) return rc;
  CallScope sc(s);
  uint64_t* d_off;
  uint8_t* d_logs;
  uint32_t* d_order;
  BTE_TRY_RC(sc.upload(&d_off, off.data(), count + 1));
  BTE_TRY_RC(sc.upload(&d_logs, logs.data(), count));
  BTE_TRY_RC(sc.upload(&d_order, order.data(), order.size()));
  const size_t c3max = three_lane_max();
  size_t widest3 = 0;  // widest stage that takes the three-lane form
  for (int lh = 0; lh < logmax; ++lh) {
    const size_t nbf = start[lh + 1] - start[lh];
    if (nbf < c3max) widest3 = std::max(widest3, nbf);
  }
  uint32_t* tab = nullptr;
  if (widest3) BTE_TRY_RC(sc.dev((uint8_t**)&tab, (size_t)c3_blocks(widest3) * 64 * C3_TAB_BYTES));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bit_reverse_seg<fp12>), dim3(blocks_for(total, 256)), dim3(256), 0, s, v,
                     (const uint64_t*)d_off, (const uint8_t*)d_logs, (uint32_t)count, total);
  for (int lh = 0; lh < logmax; ++lh) {
    const uint64_t nbf = start[lh + 1] - start[lh];
    if (nbf && nbf < c3max)
      hipLaunchKernelGGL(k_cyc_stage_ord3, dim3(c3_blocks(nbf)), dim3(64), 0, s, v,
                         (const uint32_t*)d_order + start[lh], nbf, (const uint64_t*)d_off, (const uint8_t*)d_logs,
                         (uint32_t)count, logmax, lh, (const cyc_plan*)t->plans, tab);
    else if (nbf)
      hipLaunchKernelGGL(k_cyc_stage_ord, dim3(blocks_for(nbf, 64)), dim3(64), 0, s, v,
                         (const uint32_t*)d_order + start[lh], nbf, (const uint64_t*)d_off, (const uint8_t*)d_logs,
                         (uint32_t)count, logmax, lh, (const cyc_plan*)t->plans);
  }
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Plans for caller-given canonical scalars, built in pinned staging and uploaded into the
// caller scope's scratch.
static int upload_plans(CallScope& sc, const uint8_t* host_scalars, size_t n, cyc_plan** out) {
  cyc_plan* pl;
  BTE_TRY_RC(sc.host(&pl, n));
  for (size_t i = 0; i < n; ++i) {
    bteh::Fr f;
    memcpy(f.l, host_scalars + 32 * i, 32);
    if (bteh::geq<4>(f.l, bteh::FR_R)) return set_error(BTE_EINVAL, "scalar %zu is not canonical", i);
    pl[i] = make_plan(f);
  }
  BTE_TRY_RC(sc.dev(out, n));
  BTE_HIP(hipMemcpyAsync(*out, pl, n * sizeof(cyc_plan), hipMemcpyHostToDevice, sc.stream()));
  return BTE_OK;
}

}  // namespace bte

using namespace bte;

// Work of one unscaled prefix inverse cyclotomic transform of size n keeping output_count outputs
// (fft.rs:284-350), from the same twiddle plans the kernels run: per butterfly with a nonzero
// twiddle, cyclo_pow_plan's table (one cyclotomic squaring, three products, twelve Frobenius
// maps) and chain (count - 1 squarings, nonzero digits - 1 products), then the butterfly's
// products (two per butterfly of the full stages, one per output of the last).  Host only.
extern "C" int bte_cyclotomic_prefix_work(size_t n, size_t output_count, uint64_t* counts) {
  const int logn = log2_exact(n);
  if (logn < 1 || !counts || output_count > n / 2) return set_error(BTE_EINVAL, "cyclotomic_prefix_work: bad shape");
  const bteh::Fr w = domain_generator(logn, true);
  std::vector<cyc_plan> plans(n / 2);
  bteh::Fr cur = bteh::fr_from_u64(1);
  for (size_t j = 0; j < n / 2; ++j) {
    plans[j] = make_plan(cur);
    cur = bteh::fr_mul(cur, w);
  }
  uint64_t sqr = 0, mul = 0, frob = 0;
  auto pow_work = [&](const cyc_plan& p) {
    if (p.count == 0) return;
    int nnz = 0;
    for (int c = 0; c < 4; ++c)
      for (int i = 0; i < CYC_DIGITS; ++i) nnz += p.naf[c][i] != 0;
    sqr += 1 + (uint64_t)(p.count - 1);
    mul += 3 + (uint64_t)(nnz - 1);
    frob += 12;
  };
  for (int lh = 0; lh < logn - 1; ++lh) {  // full stages
    const size_t half = (size_t)1 << lh;
    for (size_t k = 0; k < half; ++k) {
      const size_t reps = n / (2 * half);  // butterflies sharing twiddle k
      for (size_t r = 0; r < reps; ++r) {
        if (k) pow_work(plans[k << (logn - 1 - lh)]);
        mul += 2;
      }
    }
  }
  for (size_t k = 0; k < output_count; ++k) {  // last stage, outputs k < output_count
    if (k) pow_work(plans[k]);
    mul += 1;
  }
  counts[0] = sqr;
  counts[1] = mul;
  counts[2] = frob;
  return BTE_OK;
}

extern "C" int bte_fft_set_dft_max_log(int log_n) {
  if (log_n < 0 || log_n > 10) return set_error(BTE_EINVAL, "fft_set_dft_max_log: %d outside 0..10", log_n);
  g_dft_max_log.store(log_n);
  return BTE_OK;
}

extern "C" int bte_fft_g1(bte_p1* values, size_t n, int inverse, int scaled, void* stream) {
  return group_fft<fp>((jac<fp>*)values, n, 1, inverse != 0, scaled != 0, (hipStream_t)stream);
}

extern "C" int bte_fft_g2(bte_p2* values, size_t n, int inverse, int scaled, void* stream) {
  return group_fft<fp2>((jac<fp2>*)values, n, 1, inverse != 0, scaled != 0, (hipStream_t)stream);
}

extern "C" int bte_fft_gt(bte_fp12* values, size_t n, int inverse, int scaled, void* stream) {
  return cyclotomic_fft((fp12*)values, n, 1, inverse != 0, scaled != 0, 0, 0, (hipStream_t)stream);
}

extern "C" int bte_fft_cyclotomic(bte_fp12* values, size_t n, int inverse, void* stream) {
  return cyclotomic_fft((fp12*)values, n, 1, inverse != 0, false, 0, 0, (hipStream_t)stream);
}

extern "C" int bte_ifft_cyclotomic_prefix_unscaled(bte_fp12* values, size_t n, size_t output_count,
                                                   void* stream) {
  return cyclotomic_fft((fp12*)values, n, 1, true, false, output_count, 1, (hipStream_t)stream);
}

// Batched forms: `count` independent transforms of size n stored back to back (one launch per
// stage for all of them).
extern "C" int bte_fft_g1_batched(bte_p1* values, size_t n, size_t count, int inverse, int scaled, void* stream) {
  return group_fft<fp>((jac<fp>*)values, n, count, inverse != 0, scaled != 0, (hipStream_t)stream);
}

extern "C" int bte_fft_g2_batched(bte_p2* values, size_t n, size_t count, int inverse, int scaled, void* stream) {
  return group_fft<fp2>((jac<fp2>*)values, n, count, inverse != 0, scaled != 0, (hipStream_t)stream);
}

extern "C" int bte_fft_cyclotomic_batched(bte_fp12* values, size_t n, size_t count, int inverse, void* stream) {
  return cyclotomic_fft((fp12*)values, n, count, inverse != 0, false, 0, 0, (hipStream_t)stream);
}

extern "C" int bte_ifft_cyclotomic_prefix_unscaled_batched(bte_fp12* values, size_t n, size_t count,
                                                           size_t output_count, void* stream) {
  return cyclotomic_fft((fp12*)values, n, count, true, false, output_count, 1, (hipStream_t)stream);
}

extern "C" int bte_cyclotomic_pow_scalars(bte_fp12* values, const uint8_t* host_scalars, size_t n,
                                          void* stream) {
  if (n == 0) return BTE_OK;
  if (!values || !host_scalars) return set_error(BTE_EINVAL, "cyclotomic_pow_scalars: null pointer");
  hipStream_t s = (hipStream_t)stream;
  CallScope sc(s);
  cyc_plan* plans;
  BTE_TRY_RC(upload_plans(sc, host_scalars, n, &plans));
  hipLaunchKernelGGL(k_cyc_pow_plans, dim3(blocks_for(n, 64)), dim3(64), 0, s, (fp12*)values, n,
                     (const cyc_plan*)plans);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// K batches of B values with one exponent per slot (element b B + i uses exponent i): lanes run
// slot-major (consecutive lanes = one slot of consecutive batches) so a wavefront shares its
// exponent whenever K >= 64.
__global__ void __launch_bounds__(64) k_cyc_pow_plans_tiled(fp12* __restrict__ v, size_t B, size_t K,
                                                            const cyc_plan* __restrict__ plans) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= B * K) return;
  const size_t i = g / K, b = g - i * K;
  v[b * B + i] = cyclo_pow_plan(v[b * B + i], plans + i);
}

__global__ void __launch_bounds__(64) k_cyc_pow_small_tiled(fp12* __restrict__ v, size_t B, size_t K,
                                                            const int64_t* __restrict__ c) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= B * K) return;
  const size_t i = g / K, b = g - i * K;
  v[b * B + i] = cyclo_pow_small(v[b * B + i], c[i]);
}

int bte::cyc_pow_scalars_tiled(fp12* v, const uint8_t* host_scalars, size_t B, size_t K, hipStream_t s) {
  if (B == 0 || K == 0) return BTE_OK;
  CallScope sc(s);
  cyc_plan* plans;
  BTE_TRY_RC(upload_plans(sc, host_scalars, B, &plans));
  hipLaunchKernelGGL(k_cyc_pow_plans_tiled, dim3(blocks_for(B * K, 64)), dim3(64), 0, s, v, B, K,
                     (const cyc_plan*)plans);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

int bte::cyc_pow_small_tiled(fp12* v, const int64_t* coefficients, size_t B, size_t K, hipStream_t s) {
  if (B == 0 || K == 0) return BTE_OK;
  hipLaunchKernelGGL(k_cyc_pow_small_tiled, dim3(blocks_for(B * K, 64)), dim3(64), 0, s, v, B, K, coefficients);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_cyclotomic_pow_small(bte_fp12* values, const int64_t* coefficients, size_t n,
                                        void* stream) {
  if (n == 0) return BTE_OK;
  if (!values || !coefficients) return set_error(BTE_EINVAL, "cyclotomic_pow_small: null pointer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cyc_pow_small, dim3(blocks_for(n, 64)), dim3(64), 0, s, (fp12*)values, n, coefficients);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// `count` transforms of sizes 2^log_sizes[f] (host array) stored back to back, all stages of
// all of them in one launch per stage level.
extern "C" int bte_ifft_cyclotomic_prefix_unscaled_segmented(bte_fp12* values, const int* log_sizes,
                                                             const uint32_t* output_counts, size_t count,
                                                             void* stream) {
  if (count == 0) return BTE_OK;
  if (!values || !log_sizes || !output_counts) return set_error(BTE_EINVAL, "prefix_segmented: null pointer");
  return cyc_prefix_ifft_segmented((fp12*)values, log_sizes, output_counts, count, (hipStream_t)stream);
}

extern "C" int bte_fft_g1_segmented(bte_p1* values, const int* log_sizes, size_t count, int inverse, int scaled,
                                    void* stream) {
  return group_fft_segmented<fp>((jac<fp>*)values, log_sizes, count, inverse != 0, scaled != 0, (hipStream_t)stream);
}

extern "C" int bte_fft_g2_segmented(bte_p2* values, const int* log_sizes, size_t count, int inverse, int scaled,
                                    void* stream) {
  return group_fft_segmented<fp2>((jac<fp2>*)values, log_sizes, count, inverse != 0, scaled != 0,
                                  (hipStream_t)stream);
}
