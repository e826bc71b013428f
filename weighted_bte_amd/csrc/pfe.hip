// weighted_pfe batch decryption on the GPU: the host control flow of
// weighted_pfe/src/decryption.rs (prepare_decryption :421-537, precompute_batch :540-573,
// open_batch :576-644, the Cauchy transforms cauchy_scaled_group :683-708 and
// cauchy_scaled_cyclotomic :710-734, mul_group_signed_small :736-760) and partial_decrypt
// (:230-251) over this library's kernels.  Protocol bookkeeping stays with the caller.
//
// Key material layout (device): one G2 affine array holding d1 (B blocks of W points,
// d1[i W + w] = [Z(w) g_{owner(w), i}]_2) followed by d2 (B blocks of W,
// d2[i W + w] = [Z(w) u_{owner(w), i}]_2), i.e. the reference's slot-major d1 then d2.
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "cyclo.hpp"
#include "host_field.hpp"

extern "C" {
int bte_fft_g1(bte_p1* values, size_t n, int inverse, int scaled, void* stream);
int bte_fft_cyclotomic(bte_fp12* values, size_t n, int inverse, void* stream);
int bte_cyclotomic_pow_small(bte_fp12* values, const int64_t* coefficients, size_t n, void* stream);
int bte_cyclotomic_pow_scalars(bte_fp12* values, const uint8_t* host_scalars, size_t n, void* stream);
int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream);
int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream);
int bte_g2_prepare_lines(const bte_p2_affine* q, size_t n, bte_fp6* lines, void* stream);
int bte_miller_loop_prepared(const bte_fp6* lines, const bte_p1_affine* p, size_t n, bte_fp12* out, void* stream);
int bte_multi_miller_loop(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs, const uint64_t* seg_offsets,
                          size_t nseg, bte_fp12* out, void* stream);
int bte_final_exp_easy(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
int bte_final_exp_hard(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
int bte_lagrange_at_zero(size_t total_weight, const uint64_t* selected, size_t n, uint8_t* out);
int bte_g1_msm(const bte_p1_affine* points, const uint8_t* scalars, size_t n, bte_p1* out, void* stream);
int bte_fft_g1_batched(bte_p1* values, size_t n, size_t count, int inverse, int scaled, void* stream);
int bte_fft_cyclotomic_batched(bte_fp12* values, size_t n, size_t count, int inverse, void* stream);
int bte_miller_loop_prepared_tiled(const bte_fp6* lines, size_t ntables, const bte_p1_affine* p, size_t n,
                                   bte_fp12* out, void* stream);
}

namespace bte {

#define PFE_TRY(x)        \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

// v * c for a small signed integer by double-and-add (decryption.rs:736-760).
template <class F> BTE_DEV jac<F> mul_signed_small(const jac<F>& v, int64_t c) {
  if (c == 0) return jac_identity<F>();
  uint64_t mag = c < 0 ? (uint64_t)(-c) : (uint64_t)c;
  jac<F> base = v, result = jac_identity<F>();
  while (mag) {
    if (mag & 1) result = jac_add(result, base);
    mag >>= 1;
    if (mag) base = jac_dbl(base);
  }
  return c < 0 ? jac_neg(result) : result;
}

__global__ void __launch_bounds__(64) k_g1_mul_small(g1j* __restrict__ v, const int64_t* __restrict__ c, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = mul_signed_small(v[i], c[i]);
}

__global__ void __launch_bounds__(64) k_g1_mul_small_const(g1j* __restrict__ v, int64_t c, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = mul_signed_small(v[i], c);
}

// out[i] = -in[(i - 1) mod n]  (decryption.rs:694-701)
__global__ void k_g1_shift_neg(const g1j* __restrict__ in, size_t n, g1j* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = jac_neg(in[(i + n - 1) % n]);
}

// out[i] = conj(in[(i - 1) mod n])  (decryption.rs:720-727)
__global__ void k_cyc_shift_conj(const fp12* __restrict__ in, size_t n, fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fp12_conj(in[(i + n - 1) % n]);
}

// Batched forms: element i of batch b sits at b B + i; slot-major lanes share the coefficient.
__global__ void __launch_bounds__(64) k_g1_mul_small_tiled(g1j* __restrict__ v, const int64_t* __restrict__ c, size_t B,
                                                           size_t K) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= B * K) return;
  const size_t i = g / K, b = g - i * K;
  v[b * B + i] = mul_signed_small(v[b * B + i], c[i]);
}

__global__ void k_g1_shift_neg_tiled(const g1j* __restrict__ in, size_t B, size_t n, g1j* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= n) return;
  const size_t b = g / B, i = g % B;
  out[g] = jac_neg(in[b * B + (i + B - 1) % B]);
}

__global__ void k_cyc_shift_conj_tiled(const fp12* __restrict__ in, size_t B, size_t n, fp12* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= n) return;
  const size_t b = g / B, i = g % B;
  out[g] = fp12_conj(in[b * B + (i + B - 1) % B]);
}

// left[(b B + slot) T + p] = shares2B[b T + p] - cauchy[b B + slot]; tab = line table of the
// committee's d1 point (slot, p).
__global__ void k_pfe_left_tiled(const g1j* __restrict__ shares2b, const g1j* __restrict__ cauchy,
                                 const bte_fp6* __restrict__ d1_lines, size_t T, size_t B, size_t count,
                                 g1j* __restrict__ out, const bte_fp6** __restrict__ tab) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= T * B * count) return;
  const size_t p = g % T, slot = (g / T) % B, b = g / (T * B);
  out[g] = jac_sub(shares2b[b * T + p], cauchy[b * B + slot]);
  tab[g] = d1_lines + (slot * T + p) * 68;
}

template <class F> __global__ void k_jac_from_aff(const aff<F>* __restrict__ in, size_t n, jac<F>* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = jac_from_aff(in[i]);
}

// out[s] = sum of in[s * len .. (s+1) * len)
__global__ void __launch_bounds__(64) k_g2_sum_runs(const g2j* __restrict__ in, size_t len, size_t nout,
                                                    g2j* __restrict__ out) {
  __shared__ g2j sh[64];
  const size_t s = blockIdx.x;
  if (s >= nout) return;
  const int tid = threadIdx.x;
  g2j acc = jac_identity<fp2>();
  for (size_t i = tid; i < len; i += 64) acc = jac_add(acc, in[s * len + i]);
  sh[tid] = acc;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (tid < w) sh[tid] = jac_add(sh[tid], sh[tid + w]);
    __syncthreads();
  }
  if (tid == 0) out[s] = sh[0];
}

// mask[i] = d2[i] * (a_i (a_i + 1)) - V * a_i  (decryption.rs:492-501); scalars canonical 8 x u32.
__global__ void __launch_bounds__(64) k_pfe_mask(const g2j* __restrict__ d2, const g2a* __restrict__ V,
                                                 const uint32_t* __restrict__ aa1, const uint32_t* __restrict__ a,
                                                 size_t n, g2j* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  g2j x = group_scalar_mul(d2[i], aa1 + 8 * i);
  g2j y = group_scalar_mul(jac_from_aff(*V), a + 8 * i);
  out[i] = jac_sub(x, y);
}

// left[slot T + p] = shares2B[p] - cauchy[slot]  (decryption.rs:597-605)
__global__ void k_pfe_left(const g1j* __restrict__ shares2b, const g1j* __restrict__ cauchy, size_t T, size_t B,
                           g1j* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= T * B) return;
  out[g] = jac_sub(shares2b[g % T], cauchy[g / T]);
}

// first[i] * cauchy[i] (cyclotomic products before the decomposed power)
__global__ void __launch_bounds__(64) k_fp12_mul_pairs(const fp12* __restrict__ a, const fp12* __restrict__ b, size_t n,
                                                       fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fp12_mul(a[i], b[i]);
}

// combined = mask_pairing * conj(subtracted); message = ct2 * conj(hard(combined)) is finished
// by the caller-side kernels below.
__global__ void __launch_bounds__(64) k_pfe_combine(const fp12* __restrict__ mask_pair, const fp12* __restrict__ sub,
                                                    size_t n, fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fp12_mul(mask_pair[i], fp12_conj(sub[i]));
}

__global__ void __launch_bounds__(64) k_pfe_unmask(const fp12* __restrict__ ct2, const fp12* __restrict__ mask, size_t n,
                                                   fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fp12_mul(ct2[i], fp12_conj(mask[i]));
}

// Scratch and uploads of one driver call through the stream's workspace and pinned staging.
struct PScratch : CallScope {
  using CallScope::CallScope;
  template <class T> int get(T** p, size_t count) { return dev(p, count); }
  int done() { return BTE_OK; }
};

}  // namespace bte

using namespace bte;

// decryption.rs:421-537 prepare_decryption.  material = d1 then d2 (B x W each, device);
// alpha: B canonical scalars (host); global_key: V (G2 affine, device).  Outputs (device):
// d1_by_party (B x T G2 affine, slot-major, scaled by 1/(2B)), total_lines (B tables, the
// per-slot sums), mask_lines (B tables).
extern "C" int bte_pfe_prepare(const bte_p2_affine* material, size_t total_weight, size_t batch_size,
                               const uint64_t* party_offsets, size_t party_count, const uint32_t* accepted,
                               size_t n_accepted, const uint8_t* alpha, const bte_p2_affine* global_key,
                               bte_p2_affine* d1_by_party, bte_fp6* total_lines, bte_fp6* mask_lines,
                               void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t W = total_weight, B = batch_size, T = n_accepted;
  if (!material || !party_offsets || !accepted || !alpha || !global_key || !d1_by_party || !total_lines ||
      !mask_lines)
    return set_error(BTE_EINVAL, "pfe_prepare: null pointer");
  if (B == 0 || T == 0) return set_error(BTE_EINVAL, "pfe_prepare: empty batch or committee");
  if (party_offsets[party_count] != W) return set_error(BTE_EINVAL, "pfe_prepare: offsets do not sum to W");
  std::vector<uint64_t> sel, coff(1, 0);
  for (size_t p = 0; p < T; ++p) {
    const uint32_t j = accepted[p];
    if (j >= party_count) return set_error(BTE_EINVAL, "pfe_prepare: InvalidPartyIndex(%u)", j);
    for (uint64_t v = party_offsets[j]; v < party_offsets[j + 1]; ++v) sel.push_back(v);
    coff.push_back(sel.size());
  }
  const size_t WT = sel.size();
  const bteh::Fr inv2b = bteh::fr_inv(bteh::fr_from_u64(2 * B));
  // d1 MSMs: segment (slot, p) over d1_slot(slot)[range_p] with lambda / 2B
  std::vector<uint32_t> i1(B * WT), si1(B * WT), i2(B * WT), si2(B * WT);
  std::vector<uint64_t> o1(B * T + 1), o2(B + 1);
  for (size_t slot = 0; slot < B; ++slot) {
    for (size_t p = 0; p < T; ++p) o1[slot * T + p] = slot * WT + coff[p];
    o2[slot] = slot * WT;
    for (size_t k = 0; k < WT; ++k) {
      i1[slot * WT + k] = (uint32_t)(slot * W + sel[k]);
      si1[slot * WT + k] = (uint32_t)k;
      i2[slot * WT + k] = (uint32_t)(B * W + slot * W + sel[k]);
      si2[slot * WT + k] = (uint32_t)k;
    }
  }
  o1[B * T] = B * WT;
  o2[B] = B * WT;
  std::vector<bteh::Fr> aa1(B), av(B);
  for (size_t i = 0; i < B; ++i) {
    memcpy(av[i].l, alpha + 32 * i, 32);
    aa1[i] = bteh::fr_mul(av[i], bteh::fr_add(av[i], bteh::fr_from_u64(1)));
  }
  if (2 * B * W > 0xffffffffu) return set_error(BTE_EINVAL, "pfe_prepare: material too large");
  PScratch sc(s);
  uint8_t *d_lam, *d_lams;
  uint32_t *d_i1, *d_si1, *d_i2, *d_si2, *d_aa1, *d_a;
  uint64_t *d_o1, *d_o2;
  g2j *d1p, *d1t, *d2t, *mask;
  g2a *d1t_aff, *mask_aff;
  PFE_TRY(sc.get(&d_lam, 32 * WT));
  PFE_TRY(sc.get(&d_lams, 32 * WT));
  PFE_TRY(lagrange_at_zero_device(W, sel.data(), WT, nullptr, d_lam, s));
  PFE_TRY(lagrange_at_zero_device(W, sel.data(), WT, (const uint8_t*)inv2b.l, d_lams, s));
  PFE_TRY(sc.upload(&d_i1, i1.data(), i1.size()));
  PFE_TRY(sc.upload(&d_si1, si1.data(), si1.size()));
  PFE_TRY(sc.upload(&d_i2, i2.data(), i2.size()));
  PFE_TRY(sc.upload(&d_si2, si2.data(), si2.size()));
  PFE_TRY(sc.upload(&d_o1, o1.data(), o1.size()));
  PFE_TRY(sc.upload(&d_o2, o2.data(), o2.size()));
  PFE_TRY(sc.upload(&d_aa1, (const uint32_t*)aa1.data(), 8 * B));
  PFE_TRY(sc.upload(&d_a, (const uint32_t*)av.data(), 8 * B));
  PFE_TRY(sc.done());
  PFE_TRY(sc.get(&d1p, B * T));
  PFE_TRY(sc.get(&d1t, B));
  PFE_TRY(sc.get(&d2t, B));
  PFE_TRY(sc.get(&mask, B));
  PFE_TRY(sc.get(&d1t_aff, B));
  PFE_TRY(sc.get(&mask_aff, B));
  const size_t npts = 2 * B * W;
  PFE_TRY(g2_msm_batched_gather(material, npts, d_i1, d_lams, d_si1, B * WT, d_o1, B * T, 0, d1p, s));
  PFE_TRY(g2_msm_batched_gather(material, npts, d_i2, d_lam, d_si2, B * WT, d_o2, B, 0, d2t, s));
  hipLaunchKernelGGL(k_g2_sum_runs, dim3((unsigned)B), dim3(64), 0, s, (const g2j*)d1p, T, B, d1t);
  hipLaunchKernelGGL(k_pfe_mask, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const g2j*)d2t, (const g2a*)global_key,
                     (const uint32_t*)d_aa1, (const uint32_t*)d_a, B, mask);
  PFE_TRY(bte_g2_batch_normalize((const bte_p2*)d1p, B * T, d1_by_party, s));
  PFE_TRY(bte_g2_batch_normalize((const bte_p2*)d1t, B, (bte_p2_affine*)d1t_aff, s));
  PFE_TRY(bte_g2_batch_normalize((const bte_p2*)mask, B, (bte_p2_affine*)mask_aff, s));
  PFE_TRY(bte_g2_prepare_lines((const bte_p2_affine*)d1t_aff, B, total_lines, s));
  PFE_TRY(bte_g2_prepare_lines((const bte_p2_affine*)mask_aff, B, mask_lines, s));
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// decryption.rs:540-573 precompute_batch: ct1 = B G1 affine; outputs scaled_first_cauchy (B G1
// Jacobian) and pairing_cauchy (B Fp12, cyclotomic, before the hard part).
extern "C" int bte_pfe_precompute(const bte_p1_affine* ct1, size_t batch_size, const bte_fp6* total_lines,
                                  bte_p1* first_cauchy, bte_fp12* pairing_cauchy, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size;
  if (!ct1 || !total_lines || !first_cauchy || !pairing_cauchy) return set_error(BTE_EINVAL, "pfe_precompute: null");
  if (B == 0 || (B & (B - 1))) return set_error(BTE_EINVAL, "pfe_precompute: InvalidBatchSize(%zu)", B);
  std::vector<int64_t> coef(B);
  for (size_t k = 0; k < B; ++k) coef[k] = (int64_t)B - 1 - 2 * (int64_t)k;  // decryption.rs:198-200
  PScratch sc(s);
  int64_t* d_coef;
  g1j *c, *spec;
  fp12 *ml, *cyc, *cspec;
  PFE_TRY(sc.upload(&d_coef, coef.data(), B));
  PFE_TRY(sc.done());
  PFE_TRY(sc.get(&c, B));
  PFE_TRY(sc.get(&spec, B));
  PFE_TRY(sc.get(&ml, B));
  PFE_TRY(sc.get(&cyc, B));
  PFE_TRY(sc.get(&cspec, B));
  // cauchy_scaled_group: fft, * (B-1-2k), shift with negation, unscaled ifft
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_jac_from_aff<fp>), dim3(blocks_for(B, 256)), dim3(256), 0, s, (const g1a*)ct1, B, spec);
  PFE_TRY(bte_fft_g1((bte_p1*)spec, B, 0, 0, s));
  hipLaunchKernelGGL(k_g1_mul_small, dim3(blocks_for(B, 64)), dim3(64), 0, s, spec, (const int64_t*)d_coef, B);
  hipLaunchKernelGGL(k_g1_shift_neg, dim3(blocks_for(B, 256)), dim3(256), 0, s, (const g1j*)spec, B, (g1j*)first_cauchy);
  PFE_TRY(bte_fft_g1(first_cauchy, B, 1, 0, s));
  // cauchy_scaled_cyclotomic over the easy-exponentiated prepared Miller loops
  PFE_TRY(bte_miller_loop_prepared(total_lines, ct1, B, (bte_fp12*)ml, s));
  PFE_TRY(bte_final_exp_easy((const bte_fp12*)ml, B, (bte_fp12*)cyc, s));
  PFE_TRY(bte_fft_cyclotomic((bte_fp12*)cyc, B, 0, s));
  PFE_TRY(bte_cyclotomic_pow_small((bte_fp12*)cyc, d_coef, B, s));
  hipLaunchKernelGGL(k_cyc_shift_conj, dim3(blocks_for(B, 256)), dim3(256), 0, s, (const fp12*)cyc, B, (fp12*)pairing_cauchy);
  PFE_TRY(bte_fft_cyclotomic(pairing_cauchy, B, 1, s));
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// decryption.rs:576-644 open_batch.  shares: T G1 Jacobian (committee order); alpha: B canonical
// scalars (host).  messages_out: B Fp12.
extern "C" int bte_pfe_open(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* d1_by_party,
                            const bte_fp6* mask_lines, const bte_p1_affine* ct1, const bte_fp12* ct2,
                            const bte_p1* first_cauchy, const bte_fp12* pairing_cauchy, const uint8_t* alpha,
                            size_t batch_size, bte_fp12* messages_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, T = n_accepted;
  if (!shares || !d1_by_party || !mask_lines || !ct1 || !ct2 || !first_cauchy || !pairing_cauchy || !alpha ||
      !messages_out)
    return set_error(BTE_EINVAL, "pfe_open: null pointer");
  if (B == 0 || T == 0) return set_error(BTE_EINVAL, "pfe_open: empty batch or committee");
  std::vector<uint64_t> off(B + 1);
  for (size_t i = 0; i <= B; ++i) off[i] = i * T;
  std::vector<uint8_t> aa1(32 * B);
  for (size_t i = 0; i < B; ++i) {
    bteh::Fr a;
    memcpy(a.l, alpha + 32 * i, 32);
    bteh::Fr p = bteh::fr_mul(a, bteh::fr_add(a, bteh::fr_from_u64(1)));
    memcpy(&aa1[32 * i], p.l, 32);
  }
  PScratch sc(s);
  uint64_t* d_off;
  g1j *sh2, *left;
  g1a* left_aff;
  fp12 *ml, *easy, *prod, *comb, *mask;
  PFE_TRY(sc.upload(&d_off, off.data(), B + 1));
  PFE_TRY(sc.done());
  PFE_TRY(sc.get(&sh2, T));
  PFE_TRY(sc.get(&left, B * T));
  PFE_TRY(sc.get(&left_aff, B * T));
  PFE_TRY(sc.get(&ml, 2 * B));
  PFE_TRY(sc.get(&easy, 2 * B));
  PFE_TRY(sc.get(&prod, B));
  PFE_TRY(sc.get(&comb, B));
  PFE_TRY(sc.get(&mask, B));
  BTE_HIP(hipMemcpyAsync(sh2, shares, T * sizeof(g1j), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_g1_mul_small_const, dim3(blocks_for(T, 64)), dim3(64), 0, s, sh2, (int64_t)(2 * B), T);
  hipLaunchKernelGGL(k_pfe_left, dim3(blocks_for(B * T, 256)), dim3(256), 0, s, (const g1j*)sh2,
                     (const g1j*)first_cauchy, T, B, left);
  PFE_TRY(bte_g1_batch_normalize((const bte_p1*)left, B * T, (bte_p1_affine*)left_aff, s));
  PFE_TRY(bte_multi_miller_loop((const bte_p1_affine*)left_aff, d1_by_party, B * T, d_off, B, (bte_fp12*)ml, s));
  PFE_TRY(bte_miller_loop_prepared(mask_lines, ct1, B, (bte_fp12*)(ml + B), s));
  PFE_TRY(bte_final_exp_easy((const bte_fp12*)ml, 2 * B, (bte_fp12*)easy, s));
  hipLaunchKernelGGL(k_fp12_mul_pairs, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const fp12*)easy,
                     (const fp12*)pairing_cauchy, B, prod);
  PFE_TRY(bte_cyclotomic_pow_scalars((bte_fp12*)prod, aa1.data(), B, s));
  hipLaunchKernelGGL(k_pfe_combine, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const fp12*)(easy + B),
                     (const fp12*)prod, B, comb);
  PFE_TRY(bte_final_exp_hard((const bte_fp12*)comb, B, (bte_fp12*)mask, s));
  hipLaunchKernelGGL(k_pfe_unmask, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const fp12*)ct2, (const fp12*)mask, B,
                     (fp12*)messages_out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// precompute_batch for `count` batches of B ciphertexts of one committee, every stage one launch
// over all batches (same outputs as bte_pfe_precompute per batch).
extern "C" int bte_pfe_precompute_batched(const bte_p1_affine* ct1, size_t batch_size, size_t count,
                                          const bte_fp6* total_lines, bte_p1* first_cauchy, bte_fp12* pairing_cauchy,
                                          void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, K = count, n = B * K;
  if (!ct1 || !total_lines || !first_cauchy || !pairing_cauchy) return set_error(BTE_EINVAL, "pfe_precompute: null");
  if (B == 0 || (B & (B - 1))) return set_error(BTE_EINVAL, "pfe_precompute: InvalidBatchSize(%zu)", B);
  if (K == 0) return BTE_OK;
  std::vector<int64_t> coef(B);
  for (size_t k = 0; k < B; ++k) coef[k] = (int64_t)B - 1 - 2 * (int64_t)k;  // decryption.rs:198-200
  PScratch sc(s);
  int64_t* d_coef;
  g1j* spec;
  fp12 *ml, *cyc;
  PFE_TRY(sc.upload(&d_coef, coef.data(), B));
  PFE_TRY(sc.done());
  PFE_TRY(sc.get(&spec, n));
  PFE_TRY(sc.get(&ml, n));
  PFE_TRY(sc.get(&cyc, n));
  PhaseTimer pt(s);
  pt.begin();
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_jac_from_aff<fp>), dim3(blocks_for(n, 256)), dim3(256), 0, s, (const g1a*)ct1, n, spec);
  PFE_TRY(bte_fft_g1_batched((bte_p1*)spec, B, K, 0, 0, s));
  hipLaunchKernelGGL(k_g1_mul_small_tiled, dim3(blocks_for(n, 64)), dim3(64), 0, s, spec, (const int64_t*)d_coef, B, K);
  hipLaunchKernelGGL(k_g1_shift_neg_tiled, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const g1j*)spec, B, n,
                     (g1j*)first_cauchy);
  PFE_TRY(bte_fft_g1_batched(first_cauchy, B, K, 1, 0, s));
  pt.mark("group_cauchy");
  PFE_TRY(bte_miller_loop_prepared_tiled(total_lines, B, ct1, n, (bte_fp12*)ml, s));
  PFE_TRY(bte_final_exp_easy((const bte_fp12*)ml, n, (bte_fp12*)cyc, s));
  pt.mark("miller_easy");
  PFE_TRY(bte_fft_cyclotomic_batched((bte_fp12*)cyc, B, K, 0, s));
  PFE_TRY(cyc_pow_small_tiled(cyc, d_coef, B, K, s));
  hipLaunchKernelGGL(k_cyc_shift_conj_tiled, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const fp12*)cyc, B, n,
                     (fp12*)pairing_cauchy);
  PFE_TRY(bte_fft_cyclotomic_batched(pairing_cauchy, B, K, 1, s));
  pt.mark("cyclotomic_cauchy");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

// open_batch for `count` batches of one committee against the line tables of d1_by_party
// (bte_g2_prepare_lines over its B x T points, committee state): shares count x T, ct1 / ct2 /
// first_cauchy / pairing_cauchy count x B; alpha B canonical scalars (host).
extern "C" int bte_pfe_open_prepared_batched(const bte_p1* shares, size_t n_accepted, const bte_fp6* d1_lines,
                                             const bte_fp6* mask_lines, const bte_p1_affine* ct1,
                                             const bte_fp12* ct2, const bte_p1* first_cauchy,
                                             const bte_fp12* pairing_cauchy, const uint8_t* alpha, size_t batch_size,
                                             size_t count, bte_fp12* messages_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, T = n_accepted, K = count, n = B * K;
  if (!shares || !d1_lines || !mask_lines || !ct1 || !ct2 || !first_cauchy || !pairing_cauchy || !alpha ||
      !messages_out)
    return set_error(BTE_EINVAL, "pfe_open: null pointer");
  if (B == 0 || T == 0) return set_error(BTE_EINVAL, "pfe_open: empty batch or committee");
  if (K == 0) return BTE_OK;
  std::vector<uint64_t> off(n + 1);
  for (size_t i = 0; i <= n; ++i) off[i] = i * T;
  std::vector<uint8_t> aa1(32 * B);
  for (size_t i = 0; i < B; ++i) {
    bteh::Fr a;
    memcpy(a.l, alpha + 32 * i, 32);
    bteh::Fr p = bteh::fr_mul(a, bteh::fr_add(a, bteh::fr_from_u64(1)));
    memcpy(&aa1[32 * i], p.l, 32);
  }
  PScratch sc(s);
  g1j *sh2, *left;
  g1a* left_aff;
  const bte_fp6** tab;
  fp12 *ml, *easy, *prod, *comb, *mask;
  PFE_TRY(sc.get(&sh2, K * T));
  PFE_TRY(sc.get(&left, n * T));
  PFE_TRY(sc.get(&left_aff, n * T));
  PFE_TRY(sc.get(&tab, n * T));
  PFE_TRY(sc.get(&ml, 2 * n));
  PFE_TRY(sc.get(&easy, 2 * n));
  PFE_TRY(sc.get(&prod, n));
  PFE_TRY(sc.get(&comb, n));
  PFE_TRY(sc.get(&mask, n));
  PhaseTimer pt(s);
  pt.begin();
  BTE_HIP(hipMemcpyAsync(sh2, shares, K * T * sizeof(g1j), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_g1_mul_small_const, dim3(blocks_for(K * T, 64)), dim3(64), 0, s, sh2, (int64_t)(2 * B), K * T);
  hipLaunchKernelGGL(k_pfe_left_tiled, dim3(blocks_for(n * T, 256)), dim3(256), 0, s, (const g1j*)sh2,
                     (const g1j*)first_cauchy, d1_lines, T, B, K, left, tab);
  PFE_TRY(bte_g1_batch_normalize((const bte_p1*)left, n * T, (bte_p1_affine*)left_aff, s));
  PFE_TRY(multi_miller_prepared((const void* const*)tab, left_aff, n * T, off.data(), n, ml, s));
  PFE_TRY(bte_miller_loop_prepared_tiled(mask_lines, B, ct1, n, (bte_fp12*)(ml + n), s));
  pt.mark("miller");
  PFE_TRY(bte_final_exp_easy((const bte_fp12*)ml, 2 * n, (bte_fp12*)easy, s));
  hipLaunchKernelGGL(k_fp12_mul_pairs, dim3(blocks_for(n, 64)), dim3(64), 0, s, (const fp12*)easy,
                     (const fp12*)pairing_cauchy, n, prod);
  PFE_TRY(cyc_pow_scalars_tiled(prod, aa1.data(), B, K, s));
  hipLaunchKernelGGL(k_pfe_combine, dim3(blocks_for(n, 64)), dim3(64), 0, s, (const fp12*)(easy + n),
                     (const fp12*)prod, n, comb);
  PFE_TRY(bte_final_exp_hard((const bte_fp12*)comb, n, (bte_fp12*)mask, s));
  hipLaunchKernelGGL(k_pfe_unmask, dim3(blocks_for(n, 64)), dim3(64), 0, s, (const fp12*)ct2, (const fp12*)mask, n,
                     (fp12*)messages_out);
  pt.mark("final");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

__global__ void k_pfe_proof_bases(const g1a* __restrict__ commitments, const g1a* __restrict__ ct1, size_t B,
                                  g1a* __restrict__ bases) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g > 2 * B) return;
  if (g == 0) {
    g1a G;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      G.x.l[j] = G1_GEN[0][j];
      G.y.l[j] = G1_GEN[1][j];
    }
    bases[0] = G;
  } else {
    bases[g] = g <= B ? commitments[g - 1] : ct1[g - 1 - B];
  }
}

// weighted_pfe/src/proof.rs:74-122 verify_batch after the transcript hashing: the Schnorr
// equations s_i G = A_i + e_i C_i folded with nonzero coefficients c_i into one G1 MSM of 2B+1
// terms, sum c_i s_i G - sum c_i A_i - sum c_i e_i C_i == identity.  responses s_i, challenges
// e_i and coefficients c_i are canonical scalars (HOST); commitments A_i and ct1 C_i device.
extern "C" int bte_pfe_verify_proofs_batched(const bte_p1_affine* commitments, const bte_p1_affine* ct1,
                                             const uint8_t* responses, const uint8_t* challenges,
                                             const uint8_t* coefficients, size_t batch_size, int* result,
                                             void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size;
  if (!result) return set_error(BTE_EINVAL, "pfe_verify_proofs: null result");
  *result = 0;
  if (B == 0) return BTE_OK;  // the reference rejects an empty batch
  if (!commitments || !ct1 || !responses || !challenges || !coefficients)
    return set_error(BTE_EINVAL, "pfe_verify_proofs: null pointer");
  std::vector<bteh::Fr> sc(2 * B + 1);
  bteh::Fr gsum = bteh::fr_from_u64(0), zero = bteh::fr_from_u64(0);
  for (size_t i = 0; i < B; ++i) {
    bteh::Fr c, e, r;
    memcpy(c.l, coefficients + 32 * i, 32);
    memcpy(e.l, challenges + 32 * i, 32);
    memcpy(r.l, responses + 32 * i, 32);
    gsum = bteh::fr_add(gsum, bteh::fr_mul(c, r));
    sc[1 + i] = bteh::fr_sub(zero, c);
    sc[1 + B + i] = bteh::fr_sub(zero, bteh::fr_mul(c, e));
  }
  sc[0] = gsum;
  PScratch ps(s);
  g1a* bases;
  uint8_t* d_sc;
  g1j* out;
  PFE_TRY(ps.get(&bases, 2 * B + 1));
  PFE_TRY(ps.upload(&d_sc, (const uint8_t*)sc.data(), 32 * (2 * B + 1)));
  PFE_TRY(ps.get(&out, 1));
  PFE_TRY(ps.done());
  hipLaunchKernelGGL(k_pfe_proof_bases, dim3(blocks_for(2 * B + 1, 256)), dim3(256), 0, s, (const g1a*)commitments,
                     (const g1a*)ct1, B, bases);
  PFE_TRY(bte_g1_msm((const bte_p1_affine*)bases, d_sc, 2 * B + 1, (bte_p1*)out, s));
  g1j h;
  BTE_HIP(hipMemcpyAsync(&h, out, sizeof(g1j), hipMemcpyDeviceToHost, s));
  BTE_HIP(hipStreamSynchronize(s));
  bool zero_z = true;
  for (int j = 0; j < 12; ++j) zero_z = zero_z && h.z.l[j] == 0;
  *result = zero_z ? 1 : 0;
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// weighted_pfe/src/decryption.rs:230-251 partial_decrypt for N parties at once:
// sigma_j = sum_i g_{j,i} ct1_i with g_{j,i} given (N x B canonical scalars, device).
extern "C" int bte_pfe_partial_decrypt(const bte_p1_affine* ct1, size_t batch_size, const uint8_t* fractions,
                                       size_t n_parties, bte_p1* sigma_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, N = n_parties;
  if (!ct1 || !fractions || !sigma_out) return set_error(BTE_EINVAL, "pfe_partial_decrypt: null pointer");
  if (N == 0) return BTE_OK;
  std::vector<uint32_t> pidx(N * B);
  std::vector<uint64_t> off(N + 1);
  for (size_t j = 0; j < N; ++j) {
    off[j] = j * B;
    for (size_t i = 0; i < B; ++i) pidx[j * B + i] = (uint32_t)i;
  }
  off[N] = N * B;
  PScratch sc(s);
  uint32_t* d_pidx;
  uint64_t* d_off;
  PFE_TRY(sc.upload(&d_pidx, pidx.data(), N * B));
  PFE_TRY(sc.upload(&d_off, off.data(), N + 1));
  PFE_TRY(sc.done());
  PFE_TRY(g1_msm_batched_gather(ct1, B, d_pidx, fractions, nullptr, N * B, d_off, N, 0, sigma_out, s));
  return BTE_OK;
}

// ---------------------------------------------------------------------------------------------
// Plan of the fused weighted_pfe step: bte_pfe_precompute_batched then
// bte_pfe_open_prepared_batched for `count` batches of one committee (same outputs), with the
// Cauchy intermediates in plan-owned buffers, and its HIP graph: the first launch captures one
// run (the entry points allocate nothing eager and own their scratch while captured), later
// launches replay it.  No counterpart in the reference (its steps are rayon loops).
// ---------------------------------------------------------------------------------------------
struct bte_pfe_plan {
  size_t B = 0, K = 0, T = 0;
  const bte_fp6 *total_lines = nullptr, *d1_lines = nullptr, *mask_lines = nullptr;
  std::vector<uint8_t> alpha;
  bte_p1* first_cauchy = nullptr;
  bte_fp12* pairing_cauchy = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* bound[4] = {nullptr, nullptr, nullptr, nullptr};
  ~bte_pfe_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (first_cauchy) (void)hipFree(first_cauchy);
    if (pairing_cauchy) (void)hipFree(pairing_cauchy);
  }
};

extern "C" int bte_pfe_plan_create(size_t batch_size, size_t count, size_t n_accepted, const bte_fp6* total_lines,
                                   const bte_fp6* d1_lines, const bte_fp6* mask_lines, const uint8_t* alpha,
                                   bte_pfe_plan** out, void* stream) {
  (void)stream;
  if (!out || !total_lines || !d1_lines || !mask_lines || !alpha)
    return set_error(BTE_EINVAL, "pfe_plan_create: null pointer");
  if (batch_size == 0 || (batch_size & (batch_size - 1)) || count == 0 || n_accepted == 0)
    return set_error(BTE_EINVAL, "pfe_plan_create: bad shape (B must be a power of two)");
  *out = nullptr;
  bte_pfe_plan* p = new bte_pfe_plan();
  p->B = batch_size;
  p->K = count;
  p->T = n_accepted;
  p->total_lines = total_lines;
  p->d1_lines = d1_lines;
  p->mask_lines = mask_lines;
  p->alpha.assign(alpha, alpha + 32 * batch_size);
  const size_t n = batch_size * count;
  hipError_t e = hipMalloc((void**)&p->first_cauchy, n * sizeof(bte_p1));
  if (e == hipSuccess) e = hipMalloc((void**)&p->pairing_cauchy, n * sizeof(bte_fp12));
  if (e != hipSuccess) {
    delete p;
    return set_error(BTE_ENOMEM, "pfe plan: %s", hipGetErrorString(e));
  }
  *out = p;
  return BTE_OK;
}

static int pfe_plan_enqueue(bte_pfe_plan* p, const bte_p1_affine* ct1, const bte_p1* shares, const bte_fp12* ct2,
                            bte_fp12* messages_out, void* stream) {
  PFE_TRY(bte_pfe_precompute_batched(ct1, p->B, p->K, p->total_lines, p->first_cauchy, p->pairing_cauchy, stream));
  return bte_pfe_open_prepared_batched(shares, p->T, p->d1_lines, p->mask_lines, ct1, ct2, p->first_cauchy,
                                       p->pairing_cauchy, p->alpha.data(), p->B, p->K, messages_out, stream);
}

extern "C" int bte_pfe_plan_run(bte_pfe_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                                const bte_fp12* ct2, bte_fp12* messages_out, void* stream) {
  if (!plan || !ct1 || !shares || !ct2 || !messages_out) return set_error(BTE_EINVAL, "pfe_plan_run: null pointer");
  return pfe_plan_enqueue(plan, ct1, shares, ct2, messages_out, stream);
}

extern "C" int bte_pfe_plan_launch_graph(bte_pfe_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                                         const bte_fp12* ct2, bte_fp12* messages_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!plan || !ct1 || !shares || !ct2 || !messages_out) return set_error(BTE_EINVAL, "pfe_plan_graph: null pointer");
  if (!s) return set_error(BTE_EINVAL, "pfe_plan_graph: the null stream cannot be captured");
  const void* args[4] = {ct1, shares, ct2, messages_out};
  if (!plan->exec) {
    if (PhaseTimer(s).on)  // phase timing synchronises the stream, which a capture cannot contain
      return set_error(BTE_EINVAL, "pfe_plan_graph: disable profiling before the first launch");
    BTE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = pfe_plan_enqueue(plan, ct1, shares, ct2, messages_out, stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_error(BTE_EDEVICE, "pfe_plan_graph: capture failed: %s", hipGetErrorString(e));
    plan->graph = g;
    BTE_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
    for (int i = 0; i < 4; ++i) plan->bound[i] = args[i];
  } else {
    for (int i = 0; i < 4; ++i)
      if (plan->bound[i] != args[i])
        return set_error(BTE_EINVAL, "pfe_plan_graph: the graph is bound to the buffers of its first launch");
  }
  BTE_HIP(hipGraphLaunch(plan->exec, s));
  return BTE_OK;
}

extern "C" int bte_pfe_plan_destroy(bte_pfe_plan* plan) {
  delete plan;
  return BTE_OK;
}
