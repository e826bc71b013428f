// Optimal-ate Miller loops and the split final exponentiation of BLS12-381 on gfx950.
//
// Replaces, from weighted_btx/src/final_exponentiation.rs:
//   PreparedG2Lines::from_affine        :186-195  (blst_precompute_lines)   -> bte_g2_prepare_lines
//   PreparedG2Lines::miller_loop        :207-219  (blst_miller_loop_lines)  -> bte_miller_loop_prepared
//   unprepared_multi_miller_loop        :229-273  (blst_miller_loop_n)      -> bte_multi_miller_loop
//   easy / batch_easy / hard / full     :276-354                            -> bte_final_exp_*
//
// Miller loop: f = prod over the bits of |x| (x = -0xd201000000010000) of doubling lines, with
// addition lines at the 5 set bits below the top; 63 + 5 = 68 lines per G2 point, then f is
// conjugated because x < 0.  T runs in homogeneous projective coordinates on the twist
// y^2 = x^3 + 4(u+1); a line through T evaluated at P = (xp, yp) in G1, scaled by a factor that
// the final exponentiation removes, is
//     l = l0 + (cx * xp) v + (cy * yp) v w     (three Fp2 coefficients, an Fp6-sized record)
//   doubling:  l0 = Y^2 - 3b'Z^2, cx = -3X^2, cy = 2YZ
//   addition:  theta = Y - yq Z, lambda = X - xq Z; l0 = theta xq - lambda yq, cx = -theta,
//              cy = lambda
// Raw Miller outputs therefore differ from blst's by such factors; values agree after the final
// exponentiation, which is where the reference's outputs are canonical.
//
// A multi-Miller loop over a segment of pairs equals the product of the per-pair loops (the
// shared squaring distributes over the product), so pairs run one per thread and each segment
// is multiplied together by one workgroup.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "coop12.hpp"
#include "cyclo.hpp"

namespace bte {

static constexpr uint64_t MILLER_X_ABS = 0xd201000000010000ull;
static constexpr int MILLER_LINES = 68;

struct line_t {
  fp2 l0, cx, cy;  // laid out as one blst_fp6
};
static_assert(sizeof(line_t) == 288, "line layout");

// T <- 2T and its tangent line.
BTE_DEV line_t dbl_step(g2j& T) {
  const fp2 X = T.x, Y = T.y, Z = T.z;
  fp2 A = fp2_half(fp2_mul(X, Y));
  fp2 B = fp2_sqr(Y);
  fp2 C = fp2_sqr(Z);
  fp2 C3 = fp2_mul3(fp2_mul_xi(C));           // 3 (u+1) Z^2
  fp2 E = fp2_dbl(fp2_dbl(C3));                // 3 b' Z^2 with b' = 4 (u+1)
  fp2 F = fp2_mul3(E);
  fp2 G = fp2_half(fp2_add(B, F));
  fp2 H = fp2_sub(fp2_sqr(fp2_add(Y, Z)), fp2_add(B, C));
  line_t l;
  l.l0 = fp2_sub(B, E);
  l.cx = fp2_neg(fp2_mul3(fp2_sqr(X)));
  l.cy = H;
  T.x = fp2_mul(A, fp2_sub(B, F));
  T.y = fp2_sub(fp2_sqr(G), fp2_mul3(fp2_sqr(E)));
  T.z = fp2_mul(B, H);
  return l;
}

// T <- T + Q (Q affine) and the chord line.
BTE_DEV line_t add_step(g2j& T, const g2a& Q) {
  fp2 theta = fp2_sub(T.y, fp2_mul(Q.y, T.z));
  fp2 lambda = fp2_sub(T.x, fp2_mul(Q.x, T.z));
  fp2 C = fp2_sqr(theta);
  fp2 D = fp2_sqr(lambda);
  fp2 E = fp2_mul(lambda, D);
  fp2 F = fp2_mul(T.z, C);
  fp2 G = fp2_mul(T.x, D);
  fp2 H = fp2_sub(fp2_add(E, F), fp2_dbl(G));
  line_t l;
  l.l0 = fp2_sub(fp2_mul(theta, Q.x), fp2_mul(lambda, Q.y));
  l.cx = fp2_neg(theta);
  l.cy = lambda;
  T.x = fp2_mul(lambda, H);
  T.y = fp2_sub(fp2_mul(theta, fp2_sub(G, H)), fp2_mul(T.y, E));
  T.z = fp2_mul(T.z, E);
  return l;
}

BTE_DEV fp12 apply_line(const fp12& f, const line_t& l, const g1a& P) {
  return fp12_mul_line(f, l.l0, fp2_mul_fp(l.cx, P.x), fp2_mul_fp(l.cy, P.y));
}

// One Miller loop f_{|x|,Q}(P), conjugated; identity inputs give one.
BTE_DEV fp12 miller_loop_unprepared(const g1a& P, const g2a& Q) {
  if (aff_is_identity(P) || aff_is_identity(Q)) return fp12_one();
  g2j T = {Q.x, Q.y, fp2_one()};
  fp12 f = fp12_one();
#pragma unroll 1
  for (int b = 62; b >= 0; --b) {
    if (b != 62) f = fp12_sqr(f);
    f = apply_line(f, dbl_step(T), P);
    if ((MILLER_X_ABS >> b) & 1) f = apply_line(f, add_step(T, Q), P);
  }
  return fp12_conj(f);
}

__global__ void __launch_bounds__(64) k_prepare_lines(const g2a* __restrict__ q, size_t n, line_t* __restrict__ lines) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  line_t* out = lines + i * MILLER_LINES;
  const g2a Q = q[i];
  if (aff_is_identity(Q)) {
    // an all-zero table marks the identity (a real first line has cy = 2YZ != 0)
    for (int k = 0; k < MILLER_LINES; ++k) out[k] = {fp2_zero(), fp2_zero(), fp2_zero()};
    return;
  }
  g2j T = {Q.x, Q.y, fp2_one()};
  int k = 0;
#pragma unroll 1
  for (int b = 62; b >= 0; --b) {
    out[k++] = dbl_step(T);
    if ((MILLER_X_ABS >> b) & 1) out[k++] = add_step(T, Q);
  }
}

// Miller loop of P against a prepared table (all-zero table = identity Q), conjugated.
BTE_DEV fp12 miller_with_lines(const line_t* __restrict__ L, const g1a& P) {
  if (aff_is_identity(P) || fp2_is_zero(L[0].cy)) return fp12_one();
  fp12 f = fp12_one();
  int k = 0;
#pragma unroll 1
  for (int b = 62; b >= 0; --b) {
    if (b != 62) f = fp12_sqr(f);
    f = apply_line(f, L[k++], P);
    if ((MILLER_X_ABS >> b) & 1) f = apply_line(f, L[k++], P);
  }
  return fp12_conj(f);
}

__global__ void __launch_bounds__(64) k_miller_prepared(const line_t* __restrict__ lines, size_t ntables,
                                                        const g1a* __restrict__ p, size_t n, fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = miller_with_lines(lines + (i % ntables) * MILLER_LINES, p[i]);
}

__global__ void __launch_bounds__(64) k_miller_prepared_seg(const line_t* const* __restrict__ tables,
                                                            const uint64_t* __restrict__ seg_off, uint32_t nseg,
                                                            const g1a* __restrict__ p, size_t n,
                                                            fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = seg_of(seg_off, nseg, i);
  out[i] = miller_with_lines(tables[s] + (i - seg_off[s]) * MILLER_LINES, p[i]);
}

int miller_prepared_segmented(const void* const* tables, const uint64_t* seg_off, size_t nseg, const void* p,
                              size_t n, void* out, hipStream_t s) {
  if (n == 0) return BTE_OK;
  hipLaunchKernelGGL(k_miller_prepared_seg, dim3(blocks_for(n, 64)), dim3(64), 0, s, (const line_t* const*)tables,
                     seg_off, (uint32_t)nseg, (const g1a*)p, n, (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

__global__ void __launch_bounds__(64) k_miller_pairs(const g1a* __restrict__ p, const g2a* __restrict__ q, size_t n,
                                                     fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = miller_loop_unprepared(p[i], q[i]);
}

// Multi-Miller loop over prepared tables with a shared accumulator (the squarings of
// blst_miller_loop_n shared by the pairs of a chunk): thread j runs pairs
// [chunk_lo[j], chunk_hi[j]) of one segment, pair g against table tables[g] and point P[g];
// pairs with an identity are skipped (final_exponentiation.rs:238-256).  out[j] is conjugated.
__global__ void __launch_bounds__(64) k_multi_miller_prepared(const line_t* const* __restrict__ tables,
                                                              const g1a* __restrict__ P,
                                                              const uint64_t* __restrict__ chunk_lo,
                                                              const uint64_t* __restrict__ chunk_hi, size_t nchunks,
                                                              fp12* __restrict__ out) {
  const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (j >= nchunks) return;
  const uint64_t a = chunk_lo[j], b = chunk_hi[j];
  fp12 f = fp12_one();
  int k = 0;
#pragma unroll 1
  for (int bit = 62; bit >= 0; --bit) {
    if (bit != 62) f = fp12_sqr(f);
    const int nsteps = ((MILLER_X_ABS >> bit) & 1) ? 2 : 1;
#pragma unroll 1
    for (int st = 0; st < nsteps; ++st, ++k) {
#pragma unroll 1
      for (uint64_t g = a; g < b; ++g) {
        const line_t* L = tables[g];
        const g1a Pg = P[g];
        if (aff_is_identity(Pg) || fp2_is_zero(L[0].cy)) continue;
        f = apply_line(f, L[k], Pg);
      }
    }
  }
  out[j] = fp12_conj(f);
}

// ---- hard part on three lanes per element (coop12.hpp) -------------------------------------
// The chain of hard_part (cyclo.hpp) with each Fp12 spread over three lanes: a third of the Fp
// products per lane on the same values, which shortens the chain when too few elements fill the
// GPU one per lane.  Inputs are cyclotomic (the easy part's outputs), as hard_part assumes.
// Products and squares are out-of-line calls, as the one-lane Fp12 operations are, so the
// kernels stay small; the exchange buffer is re-typed as LDS inside them so that the calls keep
// LDS instructions (a plain pointer argument would turn them into flat accesses).
typedef __attribute__((address_space(3))) uint32_t lds_word;

BTE_DEV c3ctx c3_in_lds(const c3ctx& c0) {
  c3ctx c = c0;
  c.buf = (volatile uint32_t*)(volatile lds_word*)c0.buf;
  return c;
}
BTE_NOINL fp4 c3_mul_out(const c3ctx& c, const fp4& A, const fp4& B) { return c3_mul(c3_in_lds(c), A, B); }
BTE_NOINL fp4 c3_sqr_out(const c3ctx& c, const fp4& A) { return c3_cyclotomic_sqr(c3_in_lds(c), A); }

// a^(|x|/2), conjugated: the runs of squarings of raise_to_z_div_by_2 (cyclo.hpp), one product
// before each run.
BTE_NOINL fp4 c3_raise_to_z_div_by_2(const c3ctx& c, const fp4& a) {
  fp4 r = c3_sqr_out(c, a);
#pragma unroll 1
  for (int run = 0; run < 5; ++run) {
    const int len = run == 0 ? 2 : run == 1 ? 3 : run == 2 ? 9 : run == 3 ? 32 : 15;
    r = c3_mul_out(c, r, a);
#pragma unroll 1
    for (int i = 0; i < len; ++i) r = c3_sqr_out(c, r);
  }
  return c3_conj(c, r);
}

BTE_DEV fp4 c3_raise_to_z(const c3ctx& c, const fp4& a) { return c3_sqr_out(c, c3_raise_to_z_div_by_2(c, a)); }

BTE_DEV fp4 c3_hard_part(const c3ctx& c, const fp4& f) {
  const fp4 y0 = c3_sqr_out(c, f);
  fp4 y1 = c3_raise_to_z(c, y0);
  fp4 y2 = c3_raise_to_z_div_by_2(c, y1);
  fp4 y3 = c3_conj(c, f);
  y1 = c3_mul_out(c, y1, y3);
  y1 = c3_conj(c, y1);
  y1 = c3_mul_out(c, y1, y2);
  y2 = c3_raise_to_z(c, y1);
  y3 = c3_raise_to_z(c, y2);
  y1 = c3_conj(c, y1);
  y3 = c3_mul_out(c, y3, y1);
  y1 = c3_conj(c, y1);
  y1 = c3_frobenius(c, y1, 3);
  y2 = c3_frobenius(c, y2, 2);
  y1 = c3_mul_out(c, y1, y2);
  y2 = c3_raise_to_z(c, y3);
  y2 = c3_mul_out(c, y2, y0);
  y2 = c3_mul_out(c, y2, f);
  y1 = c3_mul_out(c, y1, y2);
  y2 = c3_frobenius(c, y3, 1);
  return c3_mul_out(c, y1, y2);
}

// 21 elements per 64-lane workgroup (lanes 3g, 3g+1, 3g+2 hold element g; lane 63 idles).
static constexpr int C3_ELEMS_PER_WG = 21;

// in == out is allowed (each lane group reads its element before it writes it).
__global__ void __launch_bounds__(64) k_final_exp_hard3(const fp12* in, size_t n, fp12* out) {
  __shared__ uint32_t xbuf[48 * 64];
  const c3ctx c = c3_make(xbuf);
  const size_t i = blockIdx.x * (size_t)C3_ELEMS_PER_WG + c.lane / 3;
  if (c.lane >= 3 * C3_ELEMS_PER_WG || i >= n) return;
  const fp4 f = c3_load(c, in + i);
  c3_store(c, out + i, c3_hard_part(c, f));
}

// Multi-pairings in one launch: thread j runs chunk j of the prepared multi-Miller loop (as
// k_multi_miller_prepared) and stores its value.  The chunks of a segment form groups of up to
// `group` consecutive chunks; the thread that finishes a group's last chunk multiplies the
// group's values and stores the group product, and the thread that stores a segment's last group
// product multiplies those and applies the easy part of the final exponentiation; the wavefront
// then runs the hard parts of the segments its lanes closed, three lanes per segment.  So no
// launch boundary
// separates the Miller loops from the final exponentiation, no thread waits for another, and the
// longest serial chain is one group plus one segment's group products.  "Last" is decided by a
// counter: each thread stores its value, fences at device scope (release only), increments the
// counter, and the one that sees the count complete fences again (acquire only) before it reads
// the others' values.  The values equal the staged path's (a product in Fp12 does not depend on the
// order of its factors, and the easy and hard parts are the same maps).
// seg_chunks: the chunks of segment s are seg_chunks[s] .. seg_chunks[s+1]; seg_groups: its
// group products are gpart[seg_groups[s] .. seg_groups[s+1]]; the counters and out[] are set by
// k_multi_pairing_init first (out[s] = 1 stays for a segment without pairs).
__global__ void k_multi_pairing_init(uint32_t* __restrict__ seg_done, fp12* __restrict__ out, size_t nseg,
                                     uint32_t* __restrict__ group_done, size_t ngroups) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < nseg) {
    seg_done[i] = 0;
    out[i] = fp12_one();
  }
  if (i < ngroups) group_done[i] = 0;
}

// Product of v[lo .. hi), hi > lo.
BTE_DEV fp12 fp12_product(const fp12* v, uint64_t lo, uint64_t hi) {
  fp12 acc = v[lo];
#pragma unroll 1
  for (uint64_t c = lo + 1; c < hi; ++c) acc = fp12_mul(acc, v[c]);
  return acc;
}

__global__ void __launch_bounds__(64) k_multi_pairing_fused(const line_t* const* __restrict__ tables,
                                                            const g1a* __restrict__ P,
                                                            const uint64_t* __restrict__ chunk_lo,
                                                            const uint64_t* __restrict__ chunk_hi,
                                                            const uint64_t* __restrict__ seg_chunks,
                                                            const uint64_t* __restrict__ seg_groups, uint32_t nseg,
                                                            uint32_t group, size_t nchunks, fp12* part,
                                                            fp12* gpart, uint32_t* seg_done, uint32_t* group_done,
                                                            fp12* out) {
  __shared__ uint32_t xbuf[48 * 64];
  // Chunk j = lane * (workgroups) + workgroup: consecutive chunks (the chunks of a small segment)
  // sit in different wavefronts, so the segments a wavefront closes, and their hard parts, spread
  // over the grid instead of piling up in the wavefronts that hold many small segments.  Every
  // lane of the wavefront stays to the end (lanes past the last chunk run an empty loop): the
  // hard parts are finished by the whole wavefront together.
  const size_t j = threadIdx.x * (size_t)gridDim.x + blockIdx.x;
  const bool live = j < nchunks;
  const uint64_t a = live ? chunk_lo[j] : 0, b = live ? chunk_hi[j] : 0;
  fp12 f = fp12_one();
  int k = 0;
#pragma unroll 1
  for (int bit = 62; bit >= 0; --bit) {
    if (bit != 62) f = fp12_sqr(f);
    const int nsteps = ((MILLER_X_ABS >> bit) & 1) ? 2 : 1;
#pragma unroll 1
    for (int st = 0; st < nsteps; ++st, ++k) {
#pragma unroll 1
      for (uint64_t g = a; g < b; ++g) {
        const line_t* L = tables[g];
        const g1a Pg = P[g];
        if (aff_is_identity(Pg) || fp2_is_zero(L[0].cy)) continue;
        f = apply_line(f, L[k], Pg);
      }
    }
  }
  bool fin = false;  // this lane closed segment s: out[s] holds its easy part
  uint32_t s = 0;
  if (live) {
    part[j] = fp12_conj(f);
    s = seg_of(seg_chunks, nseg, j);
    const uint64_t c0 = seg_chunks[s], c1 = seg_chunks[s + 1];
    const uint64_t gi = (j - c0) / group;
    const uint64_t glo = c0 + gi * group, ghi = glo + group < c1 ? glo + group : c1;
    const uint64_t gid = seg_groups[s] + gi;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (atomicAdd(&group_done[gid], 1u) + 1u == (uint32_t)(ghi - glo)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      gpart[gid] = fp12_product(part, glo, ghi);
      const uint64_t g0 = seg_groups[s], g1 = seg_groups[s + 1];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (atomicAdd(&seg_done[s], 1u) + 1u == (uint32_t)(g1 - g0)) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const fp12 acc = fp12_product(gpart, g0, g1);
        out[s] = easy_part_with_inverse(acc, fp12_inv(acc));
        fin = true;
      }
    }
  }
  // Hard parts of the segments this wavefront closed, C3_ELEMS_PER_WG at a time on three lanes
  // each: lane group q takes the q-th closing lane's segment.
  __threadfence_block();
  const uint64_t mask = __ballot(fin);
  const int cnt = __popcll(mask);
  const c3ctx c = c3_make(xbuf);
#pragma unroll 1
  for (int base = 0; base < cnt; base += C3_ELEMS_PER_WG) {
    const int q = base + c.lane / 3;
    uint64_t m = mask;
#pragma unroll 1
    for (int t = 0; t < q && m; ++t) m &= m - 1;
    const uint32_t seg = __shfl(s, m ? __builtin_ctzll(m) : 0);
    if (c.lane < 3 * C3_ELEMS_PER_WG && q < cnt) {
      const fp4 e = c3_load(c, out + seg);
      c3_store(c, out + seg, c3_hard_part(c, e));
    }
  }
}

// out[s] = prod of f[off[s] .. off[s+1]); one 64-thread workgroup per segment.
__global__ void __launch_bounds__(64) k_segment_product(const fp12* __restrict__ f, const uint64_t* __restrict__ off,
                                                        size_t nseg, fp12* __restrict__ out) {
  __shared__ fp12 sh[64];
  const size_t s = blockIdx.x;
  if (s >= nseg) return;
  const uint64_t a = off[s], b = off[s + 1];
  const int tid = threadIdx.x;
  fp12 acc = fp12_one();
  bool any = false;
  for (uint64_t i = a + tid; i < b; i += 64) {
    acc = any ? fp12_mul(acc, f[i]) : f[i];
    any = true;
  }
  sh[tid] = acc;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (tid < w && a + tid + w < b) sh[tid] = fp12_mul(sh[tid], sh[tid + w]);
    __syncthreads();
  }
  if (tid == 0) out[s] = sh[0];
}

// Easy part over chunks of `chunk` elements sharing one Fp12 inversion (Montgomery's trick, as
// batch_easy_final_exponentiation does for the whole batch; the values are identical).
__global__ void __launch_bounds__(64) k_final_exp_easy(const fp12* __restrict__ in, size_t n, uint32_t chunk,
                                                       fp12* __restrict__ out) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t start = t * chunk;
  if (start >= n) return;
  const size_t stop = start + chunk < n ? start + chunk : n;
  // prefix products staged in out[]
  fp12 acc = fp12_one();
  for (size_t i = start; i < stop; ++i) {
    out[i] = acc;
    acc = fp12_mul(acc, in[i]);
  }
  fp12 inv = fp12_inv(acc);
  for (size_t i = stop; i-- > start;) {
    const fp12 f = in[i];
    const fp12 fi = fp12_mul(inv, out[i]);
    inv = fp12_mul(inv, f);
    out[i] = easy_part_with_inverse(f, fi);
  }
}

// in == out is allowed (each thread reads its element before writing it).
__global__ void __launch_bounds__(64) k_final_exp_hard(const fp12* in, size_t n, fp12* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = hard_part(in[i]);
}

// Scratch of bte_multi_miller_loop_noalloc: one Fp12 per pair in a file-scope device array.  Each
// GPU the library runs on holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives on that device (the array's name is a host-side
// stand-in, not a device address).
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static constexpr size_t NOALLOC_MAX_PAIRS = NOALLOC_SCRATCH_BYTES / sizeof(fp12);
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_address(void** out) {
  int dev = 0;
  BTE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(BTE_EDEVICE, "multi_miller_loop_noalloc: device id %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(BTE_EDEVICE, "multi_miller_loop_noalloc: scratch address: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return BTE_OK;
}

}  // namespace bte

using namespace bte;

extern "C" int bte_g2_prepare_lines(const bte_p2_affine* q, size_t n, bte_fp6* lines, void* stream) {
  if (n == 0) return BTE_OK;
  if (!q || !lines) return set_error(BTE_EINVAL, "prepare_lines: null pointer");
  hipLaunchKernelGGL(k_prepare_lines, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream,
                     (const g2a*)q, n, (line_t*)lines);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_miller_loop_prepared(const bte_fp6* lines, const bte_p1_affine* p, size_t n, bte_fp12* out,
                                        void* stream) {
  if (n == 0) return BTE_OK;
  if (!lines || !p || !out) return set_error(BTE_EINVAL, "miller_loop_prepared: null pointer");
  hipLaunchKernelGGL(k_miller_prepared, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream,
                     (const line_t*)lines, n, (const g1a*)p, n, (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Pair i uses line table i % ntables (several batches against one prepared kernel).
extern "C" int bte_miller_loop_prepared_tiled(const bte_fp6* lines, size_t ntables, const bte_p1_affine* p, size_t n,
                                              bte_fp12* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!lines || !p || !out || ntables == 0) return set_error(BTE_EINVAL, "miller_loop_prepared_tiled: bad arguments");
  hipLaunchKernelGGL(k_miller_prepared, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream,
                     (const line_t*)lines, ntables, (const g1a*)p, n, (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_miller_loop_prepared_segmented(const bte_fp6* const* tables, const uint64_t* seg_offsets,
                                                  size_t nseg, const bte_p1_affine* p, size_t n, bte_fp12* out,
                                                  void* stream) {
  if (n == 0) return BTE_OK;
  if (!tables || !seg_offsets || !p || !out || nseg == 0)
    return set_error(BTE_EINVAL, "miller_loop_prepared_segmented: bad arguments");
  return miller_prepared_segmented((const void* const*)tables, seg_offsets, nseg, p, n, out, (hipStream_t)stream);
}

extern "C" int bte_miller_loop_pairs(const bte_p1_affine* p, const bte_p2_affine* q, size_t n, bte_fp12* out,
                                     void* stream) {
  if (n == 0) return BTE_OK;
  if (!p || !q || !out) return set_error(BTE_EINVAL, "miller_loop_pairs: null pointer");
  hipLaunchKernelGGL(k_miller_pairs, dim3(blocks_for(n, 64)), dim3(64), 0, (hipStream_t)stream,
                     (const g1a*)p, (const g2a*)q, n, (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_multi_miller_loop(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs,
                                     const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream) {
  if (nseg == 0) return BTE_OK;
  if (!seg_offsets || !out || (npairs && (!p || !q)))
    return set_error(BTE_EINVAL, "multi_miller_loop: null pointer");
  hipStream_t s = (hipStream_t)stream;
  CallScope sc(s);
  fp12* f = nullptr;
  BTE_TRY_RC(sc.dev(&f, npairs));
  if (npairs)
    hipLaunchKernelGGL(k_miller_pairs, dim3(blocks_for(npairs, 64)), dim3(64), 0, s, (const g1a*)p,
                       (const g2a*)q, npairs, f);
  hipLaunchKernelGGL(k_segment_product, dim3((unsigned)nseg), dim3(64), 0, s, (const fp12*)f, seg_offsets, nseg,
                     (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// bte_multi_miller_loop with its per-pair Miller values in the file-scope scratch array instead
// of the stream workspace: the same kernels, no allocation of any kind.  Every pair's value is
// written before it is read, so nothing needs clearing; calls that overlap on one GPU (several
// streams or threads) share the array and race.
extern "C" int bte_multi_miller_loop_noalloc(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs,
                                             const uint64_t* seg_offsets, size_t nseg, bte_fp12* out,
                                             void* stream) {
  if (npairs > NOALLOC_MAX_PAIRS)
    return set_error(BTE_EINVAL, "multi_miller_loop_noalloc: %zu pairs exceed the scratch of %zu pairs", npairs,
                     NOALLOC_MAX_PAIRS);
  if (nseg == 0) return BTE_OK;
  if (!seg_offsets || !out || (npairs && (!p || !q)))
    return set_error(BTE_EINVAL, "multi_miller_loop_noalloc: null pointer");
  hipStream_t s = (hipStream_t)stream;
  void* f = nullptr;
  BTE_TRY_RC(scratch_address(&f));
  if (npairs)
    hipLaunchKernelGGL(k_miller_pairs, dim3(blocks_for(npairs, 64)), dim3(64), 0, s, (const g1a*)p,
                       (const g2a*)q, npairs, (fp12*)f);
  hipLaunchKernelGGL(k_segment_product, dim3((unsigned)nseg), dim3(64), 0, s, (const fp12*)f, seg_offsets, nseg,
                     (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Pairs sharing one accumulator: up to BTE_TUNE_MM_CHUNK_MAX (default 8), fewer when that would
// leave fewer than BTE_TUNE_MM_MIN_THREADS threads (default 2^16, one wave per SIMD), since the
// shared squarings save less than idle SIMDs cost.

void bte::mm_chunks(const uint64_t* lo, const uint64_t* hi, size_t nseg, MMChunks& out) {
  uint64_t npairs = 0;
  for (size_t g = 0; g < nseg; ++g) npairs += hi[g] - lo[g];
  const uint64_t cmax = (uint64_t)tune_get(BTE_TUNE_MM_CHUNK_MAX), tmin = (uint64_t)tune_get(BTE_TUNE_MM_MIN_THREADS);
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(cmax, npairs / tmin));
  out.lo.clear();
  out.hi.clear();
  out.seg.assign(1, 0);
  for (size_t g = 0; g < nseg; ++g) {
    for (uint64_t c = lo[g]; c < hi[g]; c += chunk) {
      out.lo.push_back(c);
      out.hi.push_back(std::min(hi[g], c + chunk));
    }
    out.seg.push_back(out.lo.size());
  }
}

int bte::mm_enqueue(const void* const* tables, const void* p, const uint64_t* d_lo, const uint64_t* d_hi,
                    size_t nchunks, const uint64_t* d_seg, size_t nseg, void* part, void* out, hipStream_t s) {
  if (nseg == 0) return BTE_OK;
  if (nchunks)
    hipLaunchKernelGGL(k_multi_miller_prepared, dim3(blocks_for(nchunks, 64)), dim3(64), 0, s,
                       (const line_t* const*)tables, (const g1a*)p, d_lo, d_hi, nchunks, (fp12*)part);
  hipLaunchKernelGGL(k_segment_product, dim3((unsigned)nseg), dim3(64), 0, s, (const fp12*)part, d_seg, nseg,
                     (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

int bte::multi_miller_prepared_ranges(const void* const* pair_tables, const void* p, const uint64_t* lo,
                                      const uint64_t* hi, size_t nseg, void* out, hipStream_t s) {
  if (nseg == 0) return BTE_OK;
  MMChunks ch;
  mm_chunks(lo, hi, nseg, ch);
  const size_t nchunks = ch.lo.size();
  uint64_t *d_lo, *d_hi, *d_seg;
  fp12* part;
  CallScope sc(s);
  BTE_TRY_RC(sc.upload(&d_lo, ch.lo.data(), nchunks));
  BTE_TRY_RC(sc.upload(&d_hi, ch.hi.data(), nchunks));
  BTE_TRY_RC(sc.upload(&d_seg, ch.seg.data(), ch.seg.size()));
  BTE_TRY_RC(sc.dev(&part, nchunks));
  return mm_enqueue(pair_tables, p, d_lo, d_hi, nchunks, d_seg, nseg, part, out, s);
}

int bte::multi_miller_prepared(const void* const* pair_tables, const void* p, size_t npairs, const uint64_t* seg_offsets,
                               size_t nseg, void* out, hipStream_t s) {
  if (nseg == 0) return BTE_OK;
  for (size_t g = 0; g < nseg; ++g)
    if (seg_offsets[g + 1] < seg_offsets[g] || seg_offsets[g + 1] > npairs)
      return set_error(BTE_EINVAL, "multi_miller_prepared: bad segment offsets");
  return multi_miller_prepared_ranges(pair_tables, p, seg_offsets, seg_offsets + 1, nseg, out, s);
}

extern "C" int bte_multi_miller_loop_prepared(const bte_fp6* const* pair_tables, const bte_p1_affine* p, size_t npairs,
                                              const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream) {
  if (nseg == 0) return BTE_OK;
  if (!seg_offsets || !out || (npairs && (!p || !pair_tables)))
    return set_error(BTE_EINVAL, "multi_miller_loop_prepared: null pointer");
  if (seg_offsets[0] != 0 || seg_offsets[nseg] != npairs)
    return set_error(BTE_EINVAL, "multi_miller_loop_prepared: offsets must run from 0 to npairs");
  return multi_miller_prepared((const void* const*)pair_tables, p, npairs, seg_offsets, nseg, out,
                               (hipStream_t)stream);
}

extern "C" int bte_multi_pairing_prepared_fused(const bte_fp6* const* pair_tables, const bte_p1_affine* p,
                                                size_t npairs, const uint64_t* seg_offsets, size_t nseg, bte_fp12* out,
                                                void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nseg == 0) return BTE_OK;
  if (!seg_offsets || !out || (npairs && (!p || !pair_tables)))
    return set_error(BTE_EINVAL, "multi_pairing_prepared_fused: null pointer");
  if (seg_offsets[0] != 0 || seg_offsets[nseg] != npairs)
    return set_error(BTE_EINVAL, "multi_pairing_prepared_fused: offsets must run from 0 to npairs");
  for (size_t g = 0; g < nseg; ++g)
    if (seg_offsets[g + 1] < seg_offsets[g])
      return set_error(BTE_EINVAL, "multi_pairing_prepared_fused: bad segment offsets");
  if (nseg > 0xffffffffu) return set_error(BTE_EINVAL, "multi_pairing_prepared_fused: too many segments");
  MMChunks ch;
  mm_chunks(seg_offsets, seg_offsets + 1, nseg, ch);
  const size_t nchunks = ch.lo.size();
  // Groups of about sqrt(c) chunks for the segment with the most chunks c: the serial products
  // after the Miller loops are then about 2 sqrt(c) multiplications long.
  uint64_t cmax = 1;
  for (size_t g = 0; g < nseg; ++g) cmax = std::max<uint64_t>(cmax, ch.seg[g + 1] - ch.seg[g]);
  uint64_t group = 1;
  while (group * group < cmax) ++group;
  std::vector<uint64_t> seg_groups(nseg + 1, 0);
  for (size_t g = 0; g < nseg; ++g) seg_groups[g + 1] = seg_groups[g] + (ch.seg[g + 1] - ch.seg[g] + group - 1) / group;
  const size_t ngroups = seg_groups[nseg];
  CallScope sc(s);
  uint64_t *d_lo, *d_hi, *d_seg, *d_groups;
  fp12 *part, *gpart;
  uint32_t *seg_done, *group_done;
  BTE_TRY_RC(sc.upload(&d_lo, ch.lo.data(), nchunks));
  BTE_TRY_RC(sc.upload(&d_hi, ch.hi.data(), nchunks));
  BTE_TRY_RC(sc.upload(&d_seg, ch.seg.data(), ch.seg.size()));
  BTE_TRY_RC(sc.upload(&d_groups, seg_groups.data(), seg_groups.size()));
  BTE_TRY_RC(sc.dev(&part, nchunks));
  BTE_TRY_RC(sc.dev(&gpart, ngroups));
  BTE_TRY_RC(sc.dev(&seg_done, nseg));
  BTE_TRY_RC(sc.dev(&group_done, ngroups));
  hipLaunchKernelGGL(k_multi_pairing_init, dim3(blocks_for(std::max(nseg, ngroups), 64)), dim3(64), 0, s, seg_done,
                     (fp12*)out, nseg, group_done, ngroups);
  if (nchunks)
    hipLaunchKernelGGL(k_multi_pairing_fused, dim3(blocks_for(nchunks, 64)), dim3(64), 0, s,
                       (const line_t* const*)pair_tables, (const g1a*)p, (const uint64_t*)d_lo,
                       (const uint64_t*)d_hi, (const uint64_t*)d_seg, (const uint64_t*)d_groups, (uint32_t)nseg,
                       (uint32_t)group, nchunks, part, gpart, seg_done, group_done, (fp12*)out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Easy part of n elements: one element per thread while that is at most one wavefront per SIMD
// (EASY_FULL threads at this kernel's occupancy on 256 CUs), else 2 or 4 elements sharing an
// inversion, which trades three Fp12 products per element for the inversion they share.
static constexpr size_t EASY_FULL = 65536;
static void easy_part_enqueue(const fp12* in, size_t n, fp12* out, hipStream_t s) {
  const uint32_t chunk = n <= EASY_FULL ? 1 : n <= 2 * EASY_FULL ? 2 : 4;
  hipLaunchKernelGGL(k_final_exp_easy, dim3(blocks_for((n + chunk - 1) / chunk, 64)), dim3(64), 0, s, in, n, chunk,
                     out);
}

extern "C" int bte_final_exp_easy(const bte_fp12* in, size_t n, bte_fp12* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!in || !out || (const void*)in == (const void*)out)
    return set_error(BTE_EINVAL, "final_exp_easy: null or aliased buffers");
  easy_part_enqueue((const fp12*)in, n, (fp12*)out, (hipStream_t)stream);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// Hard part of n elements: one lane per element, or three lanes per element when n is below the
// three-lane switch (BTE_TUNE_CYC_THREE_LANE_MAX, the same switch as the cyclotomic stages), where
// one lane per element leaves most SIMDs idle and the shorter per-lane chain wins.
static void hard_part_enqueue(const fp12* in, size_t n, fp12* out, hipStream_t s) {
  if (n < (size_t)tune_get(BTE_TUNE_CYC_THREE_LANE_MAX))
    hipLaunchKernelGGL(k_final_exp_hard3, dim3(blocks_for(n, C3_ELEMS_PER_WG)), dim3(64), 0, s, in, n, out);
  else
    hipLaunchKernelGGL(k_final_exp_hard, dim3(blocks_for(n, 64)), dim3(64), 0, s, in, n, out);
}

extern "C" int bte_final_exp_hard(const bte_fp12* in, size_t n, bte_fp12* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!in || !out) return set_error(BTE_EINVAL, "final_exp_hard: null pointer");
  hard_part_enqueue((const fp12*)in, n, (fp12*)out, (hipStream_t)stream);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// The easy part shares an inversion over chunks of up to 4 elements (one thread per chunk);
// the hard part then runs one thread (or three lanes) per element, so every SIMD gets work at
// any n.
extern "C" int bte_final_exp(const bte_fp12* in, size_t n, bte_fp12* out, void* stream) {
  if (n == 0) return BTE_OK;
  if (!in || !out || (const void*)in == (const void*)out)
    return set_error(BTE_EINVAL, "final_exp: null or aliased buffers");
  easy_part_enqueue((const fp12*)in, n, (fp12*)out, (hipStream_t)stream);
  hard_part_enqueue((const fp12*)out, n, (fp12*)out, (hipStream_t)stream);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}
