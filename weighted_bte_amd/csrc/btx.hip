// weighted_btx batch decryption on the GPU: the host control flow of
// weighted_btx/src/decryption.rs (prepare_decryption :427-551, precompute_batch :555-603,
// open_batch :607-695, partial_decrypt :205-233) over this library's kernels; the Lagrange
// coefficients (weighted_btx/src/interpolation.rs:31-140) come from fr_ntt.hip.  Protocol
// bookkeeping (proof checks, digests, setup and committee binding) stays with the caller; these
// entry points take the already-validated arrays.
//
// Key material layout (device): one array of G2 affine points holding the reference's
// negative_material (L blocks of W, block i-1 = [z(w_v) q_owner^-i]_2) followed by its
// positive_material (L-1 blocks of W, block i-1 = [z(w_v) q_owner^i]_2), i.e.
// material[(i-1) W + v] and material[L W + (i-1) W + v].
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "cyclo.hpp"
#include "fr.hpp"
#include "host_field.hpp"

extern "C" {
int bte_fft_g1(bte_p1* values, size_t n, int inverse, int scaled, void* stream);
int bte_fft_g2(bte_p2* values, size_t n, int inverse, int scaled, void* stream);
int bte_ifft_cyclotomic_prefix_unscaled(bte_fp12* values, size_t n, size_t output_count, void* stream);
int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream);
int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream);
int bte_g2_prepare_lines(const bte_p2_affine* q, size_t n, bte_fp6* lines, void* stream);
int bte_miller_loop_prepared(const bte_fp6* lines, const bte_p1_affine* p, size_t n, bte_fp12* out, void* stream);
int bte_miller_loop_prepared_tiled(const bte_fp6* lines, size_t ntables, const bte_p1_affine* p, size_t n,
                                   bte_fp12* out, void* stream);
int bte_fft_g1_batched(bte_p1* values, size_t n, size_t count, int inverse, int scaled, void* stream);
int bte_ifft_cyclotomic_prefix_unscaled_batched(bte_fp12* values, size_t n, size_t count, size_t output_count,
                                                void* stream);
int bte_multi_miller_loop(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs, const uint64_t* seg_offsets,
                          size_t nseg, bte_fp12* out, void* stream);
int bte_final_exp_easy(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
int bte_final_exp_hard(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
int bte_final_exp(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
}

namespace bte {

// ---------------------------------------------------------------------------------------------
// Glue kernels
// ---------------------------------------------------------------------------------------------
template <class F> __global__ void k_fill_identity(jac<F>* v, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = jac_identity<F>();
}

// out[s] = sum of in[s * len .. (s+1) * len), one 64-thread group per output.
template <class F>
__global__ void __launch_bounds__(64) k_sum_runs(const jac<F>* __restrict__ in, size_t len, size_t nout,
                                                 jac<F>* __restrict__ out) {
  __shared__ jac<F> sh[64];
  const size_t s = blockIdx.x;
  if (s >= nout) return;
  const int tid = threadIdx.x;
  jac<F> acc = jac_identity<F>();
  for (size_t i = tid; i < len; i += 64) acc = jac_add(acc, in[s * len + i]);
  sh[tid] = acc;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (tid < w) sh[tid] = jac_add(sh[tid], sh[tid + w]);
    __syncthreads();
  }
  if (tid == 0) out[s] = sh[0];
}

// kernel[d] = neg[d-1] (d = 1..B-1), kernel[m-d] = pos[d-1] (decryption.rs:494-521).
__global__ void k_place_kernel(const g2j* __restrict__ neg, const g2j* __restrict__ pos, size_t B, size_t m,
                               g2j* __restrict__ kernel) {
  const size_t d = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (d == 0 || d >= B) return;
  kernel[d] = neg[d - 1];
  kernel[m - d] = pos[d - 1];
}

__global__ void __launch_bounds__(64) k_scale_all_g2(g2j* __restrict__ v, size_t n, const uint32_t* __restrict__ s) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = group_scalar_mul(v[i], s);
}

// count batches: out[b m + i] = in[b B + i] (i < B, slot valid) else the identity (an invalid
// slot contributes the identity everywhere, decryption.rs:113-125).
__global__ void k_g1_pad(const g1a* __restrict__ in, const uint8_t* __restrict__ valid, size_t B, size_t m,
                         size_t count, g1j* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= m * count) return;
  const size_t b = g / m, i = g % m;
  out[g] = (i < B && valid[b * B + i]) ? jac_from_aff(in[b * B + i]) : jac_identity<fp>();
}

// out[b B + k] = in[b m + k] for k < B.
__global__ void k_gather_prefix(const fp12* __restrict__ in, size_t m, size_t B, size_t count, fp12* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= B * count) return;
  out[g] = in[(g / B) * m + g % B];
}

// Pairs of batch b, slot i, party p: P = shares[b T + p], Q = keys[i T + p].
__global__ void k_open_pairs(const g1a* __restrict__ shares, const g2a* __restrict__ keys, size_t T, size_t B,
                             size_t count, g1a* __restrict__ P, g2a* __restrict__ Q) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= T * B * count) return;
  const size_t p = g % T, i = (g / T) % B, b = g / (T * B);
  P[g] = shares[b * T + p];
  Q[g] = keys[i * T + p];
}

template <class T> __global__ void k_tile(const T* __restrict__ in, size_t n, size_t reps, T* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n * reps) out[i] = in[i % n];
}

// beta for invalid slots is the identity (decryption.rs:585-594).
__global__ void k_mask_invalid(fp12* __restrict__ v, const uint8_t* __restrict__ valid, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n && !valid[i]) v[i] = fp12_one();
}

// m_i = ct2_i - alpha_i + beta_i in additive GT notation (decryption.rs:689-694).
__global__ void __launch_bounds__(64) k_open_combine(const fp12* __restrict__ ct2, const fp12* __restrict__ alpha,
                                                     const fp12* __restrict__ beta, const uint8_t* __restrict__ valid,
                                                     size_t n, fp12* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!valid[i]) {
    out[i] = fp12_one();
    return;
  }
  out[i] = fp12_mul(fp12_mul(ct2[i], fp12_conj(alpha[i])), beta[i]);
}

// Device view of one request of bte_btx_decrypt_requests.
struct ReqDev {
  const g1a* ct1;
  const fp12* ct2;
  const g2a* keys;         // batch_size x T, slot-major
  const bte_fp6* klines;   // their line tables (68 per key) or null
  const g1j* shares;
  fp12* out;
  uint64_t count, T;
};

static constexpr size_t LINES_PER_TABLE = 68;

// Request k owns m-slots [m_off[k], m_off[k+1]) of the padded FFT input: its ciphertexts' first
// components (identity for invalid ones), then identities up to nextpow2(2 batch_size).
__global__ void k_req_pad(const ReqDev* __restrict__ R, const uint64_t* __restrict__ m_off,
                          const uint64_t* __restrict__ slot_off, const uint8_t* __restrict__ valid, uint32_t K,
                          uint64_t total, g1j* __restrict__ c) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const uint32_t k = seg_of(m_off, K, g);
  const uint64_t i = g - m_off[k];
  c[g] = (i < R[k].count && valid[slot_off[k] + i]) ? jac_from_aff(R[k].ct1[i]) : jac_identity<fp>();
}

// pref[slot_off[k] + i] = conv[m_off[k] + i] for i < count_k.
__global__ void k_req_gather_prefix(const fp12* __restrict__ conv, const uint64_t* __restrict__ m_off,
                                    const uint64_t* __restrict__ slot_off, uint32_t K, uint64_t S,
                                    fp12* __restrict__ pref) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= S) return;
  const uint32_t k = seg_of(slot_off, K, j);
  pref[j] = conv[m_off[k] + (j - slot_off[k])];
}

__global__ void k_req_gather_shares(const ReqDev* __restrict__ R, const uint64_t* __restrict__ share_off,
                                    uint32_t K, uint64_t total, g1j* __restrict__ out) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const uint32_t k = seg_of(share_off, K, g);
  out[g] = R[k].shares[g - share_off[k]];
}

// Pair g belongs to slot j (pair_off[j] <= g < pair_off[j+1]; invalid slots own no pairs) of
// request k; party p = g - pair_off[j]: P = share p of k, Q = opening key (slot, p) of k.
__global__ void k_req_pairs(const ReqDev* __restrict__ R, const uint64_t* __restrict__ slot_off,
                            const uint64_t* __restrict__ pair_off, const uint64_t* __restrict__ share_off,
                            uint32_t K, uint32_t S, uint64_t npairs, const g1a* __restrict__ sa,
                            g1a* __restrict__ P, g2a* __restrict__ Q, const bte_fp6** __restrict__ tab) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= npairs) return;
  const uint32_t j = seg_of(pair_off, S, g);
  const uint32_t k = seg_of(slot_off, K, j);
  const uint64_t p = g - pair_off[j], i = j - slot_off[k];
  P[g] = sa[share_off[k] + p];
  if (tab)
    tab[g] = R[k].klines + (i * R[k].T + p) * LINES_PER_TABLE;
  else
    Q[g] = R[k].keys[i * R[k].T + p];
}

// Prepared opening: pair g of batch b, slot i, party p: P = shares[b T + p] and the line table
// of opening key (i, p).
__global__ void k_open_pairs_prepared(const g1a* __restrict__ shares, const bte_fp6* __restrict__ key_lines, size_t T,
                                      size_t B, size_t count, g1a* __restrict__ P, const bte_fp6** __restrict__ tab) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= T * B * count) return;
  const size_t p = g % T, i = (g / T) % B, b = g / (T * B);
  P[g] = shares[b * T + p];
  tab[g] = key_lines + (i * T + p) * LINES_PER_TABLE;
}

__global__ void __launch_bounds__(64) k_req_combine(const ReqDev* __restrict__ R, const uint64_t* __restrict__ slot_off,
                                                    uint32_t K, uint64_t S, const uint8_t* __restrict__ valid,
                                                    const fp12* __restrict__ alpha, const fp12* __restrict__ beta) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= S) return;
  const uint32_t k = seg_of(slot_off, K, j);
  const uint64_t i = j - slot_off[k];
  R[k].out[i] = valid[j] ? fp12_mul(fp12_mul(R[k].ct2[i], fp12_conj(alpha[j])), beta[j]) : fp12_one();
}

// Scratch and uploads of one driver call through the stream's workspace and pinned staging
// (ctx.hpp): nothing is allocated or waited for once the workspace has grown to this shape.
struct Scratch : CallScope {
  using CallScope::CallScope;
  template <class T> int get(T** p, size_t count) { return dev(p, count); }
};

#define BTE_TRY(x)        \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

static size_t next_pow2(size_t v) {
  size_t m = 1;
  while (m < v) m <<= 1;
  return m;
}

}  // namespace bte

using namespace bte;

// decryption.rs:427-551 prepare_decryption.  Host inputs: party_offsets (N+1), accepted party
// indices (T, in the committee order), batch size B <= L.  Device outputs: kernel_lines
// (m = nextpow2(2B) tables of 68 lines, the transformed kernel scaled by 1/m) and opening_keys
// (B x T G2 affine, slot-major).
extern "C" int bte_btx_prepare(const bte_p2_affine* material, size_t total_weight, size_t max_batch,
                               const uint64_t* party_offsets, size_t party_count, const uint32_t* accepted,
                               size_t n_accepted, size_t batch_size, bte_fp6* kernel_lines,
                               bte_p2_affine* opening_keys, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t W = total_weight, L = max_batch, B = batch_size, T = n_accepted;
  if (!material || !party_offsets || !accepted || !kernel_lines || !opening_keys)
    return set_error(BTE_EINVAL, "btx_prepare: null pointer");
  if (B == 0 || B > L) return set_error(BTE_EINVAL, "btx_prepare: InvalidBatchSize(%zu)", B);
  if (T == 0) return set_error(BTE_EINVAL, "btx_prepare: empty committee");
  // selected virtual indices in committee order and per-party offsets into them
  std::vector<uint64_t> sel;
  std::vector<uint64_t> coff(1, 0);
  for (size_t p = 0; p < T; ++p) {
    const uint32_t j = accepted[p];
    if (j >= party_count) return set_error(BTE_EINVAL, "btx_prepare: InvalidPartyIndex(%u)", j);
    for (uint64_t v = party_offsets[j]; v < party_offsets[j + 1]; ++v) sel.push_back(v);
    coff.push_back(sel.size());
  }
  if (party_offsets[party_count] != W) return set_error(BTE_EINVAL, "btx_prepare: offsets do not sum to W");
  const size_t WT = sel.size();
  const size_t m = next_pow2(2 * B);

  // Negative MSMs: segment (slot, party p) over negative_power(slot+1)[range_p] with lambda.
  std::vector<uint32_t> nidx(B * WT), nsidx(B * WT);
  std::vector<uint64_t> noff(B * T + 1);
  for (size_t slot = 0; slot < B; ++slot)
    for (size_t p = 0; p < T; ++p) {
      noff[slot * T + p] = slot * WT + coff[p];
      for (uint64_t k = coff[p]; k < coff[p + 1]; ++k) {
        nidx[slot * WT + k] = (uint32_t)(slot * W + sel[k]);
        nsidx[slot * WT + k] = (uint32_t)k;
      }
    }
  noff[B * T] = B * WT;
  // Positive MSMs: segment d-1 over positive_power(d) at the selected indices (d = 1..B-1).
  const size_t P = B - 1;
  std::vector<uint32_t> pidx(P * WT), psidx(P * WT);
  std::vector<uint64_t> poff(P + 1);
  for (size_t d = 1; d <= P; ++d) {
    poff[d - 1] = (d - 1) * WT;
    for (size_t k = 0; k < WT; ++k) {
      pidx[(d - 1) * WT + k] = (uint32_t)(L * W + (d - 1) * W + sel[k]);
      psidx[(d - 1) * WT + k] = (uint32_t)k;
    }
  }
  poff[P] = P * WT;
  const size_t n_points = L * W + (L - 1) * W;
  if (n_points > 0xffffffffu) return set_error(BTE_EINVAL, "btx_prepare: material too large for 32-bit indices");

  Scratch sc(s);
  uint8_t* d_lam;
  uint32_t *d_nidx, *d_nsidx, *d_pidx = nullptr, *d_psidx = nullptr;
  uint64_t *d_noff, *d_poff = nullptr;
  g2j *opening, *negsum, *possum, *kernel;
  g2a* kernel_aff;
  uint32_t* d_minv;
  BTE_TRY(sc.get(&d_lam, 32 * WT));
  BTE_TRY(lagrange_at_zero_device(W, sel.data(), WT, nullptr, d_lam, s));
  BTE_TRY(sc.upload(&d_nidx, nidx.data(), nidx.size()));
  BTE_TRY(sc.upload(&d_nsidx, nsidx.data(), nsidx.size()));
  BTE_TRY(sc.upload(&d_noff, noff.data(), noff.size()));
  if (P) {
    BTE_TRY(sc.upload(&d_pidx, pidx.data(), pidx.size()));
    BTE_TRY(sc.upload(&d_psidx, psidx.data(), psidx.size()));
    BTE_TRY(sc.upload(&d_poff, poff.data(), poff.size()));
  }
  BTE_TRY(sc.get(&opening, B * T));
  BTE_TRY(sc.get(&negsum, P ? P : 1));
  BTE_TRY(sc.get(&possum, P ? P : 1));
  BTE_TRY(sc.get(&kernel, m));
  BTE_TRY(sc.get(&kernel_aff, m));
  bteh::Fr minv = bteh::fr_inv(bteh::fr_from_u64(m));
  BTE_TRY(sc.upload(&d_minv, (const uint32_t*)minv.l, 8));

  PhaseTimer pt(s);
  pt.begin();
  BTE_TRY(g2_msm_batched_gather(material, n_points, d_nidx, d_lam, d_nsidx, B * WT, d_noff, B * T, 0, opening, s));
  if (P) BTE_TRY(g2_msm_batched_gather(material, n_points, d_pidx, d_lam, d_psidx, P * WT, d_poff, P, 0, possum, s));
  pt.mark("g2_msms");
  // kernel[d] = sum_p opening[(d-1) T + p]
  if (P) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sum_runs<fp2>), dim3((unsigned)P), dim3(64), 0, s, (const g2j*)opening, T, P, negsum);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fill_identity<fp2>), dim3(blocks_for(m, 256)), dim3(256), 0, s, kernel, m);
  if (P) hipLaunchKernelGGL(k_place_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const g2j*)negsum, (const g2j*)possum, B, m, kernel);
  BTE_TRY(bte_fft_g2((bte_p2*)kernel, m, 0, 0, s));
  hipLaunchKernelGGL(k_scale_all_g2, dim3(blocks_for(m, 64)), dim3(64), 0, s, kernel, m, (const uint32_t*)d_minv);
  pt.mark("kernel_fft");
  BTE_TRY(bte_g2_batch_normalize((const bte_p2*)kernel, m, (bte_p2_affine*)kernel_aff, s));
  BTE_TRY(bte_g2_prepare_lines((const bte_p2_affine*)kernel_aff, m, kernel_lines, s));
  BTE_TRY(bte_g2_batch_normalize((const bte_p2*)opening, B * T, opening_keys, s));
  pt.mark("lines");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

// decryption.rs:555-603 precompute_batch for `count` batches of B ciphertexts against one
// prepared kernel (same committee and batch size).  ct1: count x B G1 affine (identity in
// invalid slots), valid: count x B flags (host).  beta_out: count x B Fp12 on the device.
extern "C" int bte_btx_precompute_batched(const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size,
                                          size_t count, const bte_fp6* kernel_lines, bte_fp12* beta_out,
                                          void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, K = count;
  if (!ct1 || !valid || !kernel_lines || !beta_out) return set_error(BTE_EINVAL, "btx_precompute: null pointer");
  if (B == 0) return set_error(BTE_EINVAL, "btx_precompute: InvalidBatchSize(0)");
  if (K == 0) return BTE_OK;
  const size_t m = next_pow2(2 * B);
  Scratch sc(s);
  g1j* c;
  g1a* ca;
  fp12 *ml, *conv, *pref;
  uint8_t* d_valid;
  BTE_TRY(sc.get(&c, m * K));
  BTE_TRY(sc.get(&ca, m * K));
  BTE_TRY(sc.get(&ml, m * K));
  BTE_TRY(sc.get(&conv, m * K));
  BTE_TRY(sc.get(&pref, B * K));
  BTE_TRY(sc.upload(&d_valid, valid, B * K));
  PhaseTimer pt(s);
  pt.begin();
  hipLaunchKernelGGL(k_g1_pad, dim3(blocks_for(m * K, 256)), dim3(256), 0, s, (const g1a*)ct1, (const uint8_t*)d_valid, B, m, K,
                     c);
  BTE_TRY(bte_fft_g1_batched((bte_p1*)c, m, K, 0, 0, s));
  BTE_TRY(bte_g1_batch_normalize((const bte_p1*)c, m * K, (bte_p1_affine*)ca, s));
  pt.mark("g1_fft");
  BTE_TRY(bte_miller_loop_prepared_tiled(kernel_lines, m, (const bte_p1_affine*)ca, m * K, (bte_fp12*)ml, s));
  pt.mark("miller");
  BTE_TRY(bte_final_exp_easy((const bte_fp12*)ml, m * K, (bte_fp12*)conv, s));
  pt.mark("easy");
  BTE_TRY(bte_ifft_cyclotomic_prefix_unscaled_batched((bte_fp12*)conv, m, K, B, s));
  hipLaunchKernelGGL(k_gather_prefix, dim3(blocks_for(B * K, 256)), dim3(256), 0, s, (const fp12*)conv, m, B, K, pref);
  pt.mark("cyclo_ifft");
  BTE_TRY(bte_final_exp_hard((const bte_fp12*)pref, B * K, beta_out, s));
  hipLaunchKernelGGL(k_mask_invalid, dim3(blocks_for(B * K, 256)), dim3(256), 0, s, (fp12*)beta_out,
                     (const uint8_t*)d_valid, B * K);
  pt.mark("hard");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

extern "C" int bte_btx_precompute(const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size,
                                  const bte_fp6* kernel_lines, bte_fp12* beta_out, void* stream) {
  return bte_btx_precompute_batched(ct1, valid, batch_size, 1, kernel_lines, beta_out, stream);
}

// decryption.rs:607-695 open_batch for `count` batches of one committee.  shares: count x T G1
// Jacobian (committee order per batch), opening_keys from bte_btx_prepare, ct2 / beta: count x B
// Fp12, valid: count x B flags (host).  messages_out: count x B Fp12; invalid slots (the
// reference's None) are written as one.
extern "C" int bte_btx_open_batched(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* opening_keys,
                                    const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid,
                                    size_t batch_size, size_t count, bte_fp12* messages_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, T = n_accepted, K = count;
  if (!shares || !opening_keys || !ct2 || !beta || !valid || !messages_out)
    return set_error(BTE_EINVAL, "btx_open: null pointer");
  if (B == 0 || T == 0) return set_error(BTE_EINVAL, "btx_open: empty batch or committee");
  if (K == 0) return BTE_OK;
  const size_t npairs = K * B * T;
  Scratch sc(s);
  g1a *sa, *P;
  g2a* Q;
  fp12 *ml, *alpha;
  uint64_t* d_off;
  uint8_t* d_valid;
  std::vector<uint64_t> off(K * B + 1);
  for (size_t i = 0; i <= K * B; ++i) off[i] = i * T;
  BTE_TRY(sc.get(&sa, K * T));
  BTE_TRY(sc.get(&P, npairs));
  BTE_TRY(sc.get(&Q, npairs));
  BTE_TRY(sc.get(&ml, K * B));
  BTE_TRY(sc.get(&alpha, K * B));
  BTE_TRY(sc.upload(&d_off, off.data(), K * B + 1));
  BTE_TRY(sc.upload(&d_valid, valid, K * B));
  PhaseTimer pt(s);
  pt.begin();
  BTE_TRY(bte_g1_batch_normalize(shares, K * T, (bte_p1_affine*)sa, s));
  hipLaunchKernelGGL(k_open_pairs, dim3(blocks_for(npairs, 256)), dim3(256), 0, s, (const g1a*)sa,
                     (const g2a*)opening_keys, T, B, K, P, Q);
  BTE_TRY(bte_multi_miller_loop((const bte_p1_affine*)P, (const bte_p2_affine*)Q, npairs, d_off, K * B,
                                (bte_fp12*)ml, s));
  pt.mark("multi_miller");
  BTE_TRY(bte_final_exp((const bte_fp12*)ml, K * B, (bte_fp12*)alpha, s));
  pt.mark("final_exp");
  hipLaunchKernelGGL(k_open_combine, dim3(blocks_for(K * B, 64)), dim3(64), 0, s, (const fp12*)ct2,
                     (const fp12*)alpha, (const fp12*)beta, (const uint8_t*)d_valid, K * B, (fp12*)messages_out);
  pt.mark("combine");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

extern "C" int bte_btx_open(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* opening_keys,
                            const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid, size_t batch_size,
                            bte_fp12* messages_out, void* stream) {
  return bte_btx_open_batched(shares, n_accepted, opening_keys, ct2, beta, valid, batch_size, 1, messages_out,
                              stream);
}

extern "C" int bte_btx_decrypt_requests(const bte_btx_request* requests, size_t count, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (count == 0) return BTE_OK;
  if (!requests) return set_error(BTE_EINVAL, "btx_decrypt_requests: null pointer");
  std::vector<ReqDev> R;
  std::vector<const void*> tables;
  std::vector<int> logs;
  std::vector<uint32_t> outs;
  std::vector<uint64_t> m_off{0}, slot_off{0}, share_off{0}, pair_off{0};
  std::vector<uint8_t> valid;
  bool all_lines = true;
  for (size_t k = 0; k < count; ++k) {
    const bte_btx_request& q = requests[k];
    if (q.count == 0) continue;
    if (!q.ct1 || !q.ct2 || !q.kernel_lines || !q.opening_keys || !q.shares || !q.messages_out)
      return set_error(BTE_EINVAL, "btx_decrypt_requests: null pointer in request %zu", k);
    if (q.count > q.batch_size || q.n_accepted == 0)
      return set_error(BTE_EINVAL, "btx_decrypt_requests: request %zu has count > batch_size or no shares", k);
    const size_t m = next_pow2(2 * q.batch_size);
    int lg = 0;
    while (((size_t)1 << lg) < m) ++lg;
    R.push_back({(const g1a*)q.ct1, (const fp12*)q.ct2, (const g2a*)q.opening_keys, q.opening_lines,
                 (const g1j*)q.shares, (fp12*)q.messages_out, q.count, q.n_accepted});
    all_lines = all_lines && q.opening_lines != nullptr;
    tables.push_back(q.kernel_lines);
    logs.push_back(lg);
    outs.push_back((uint32_t)q.count);
    m_off.push_back(m_off.back() + m);
    slot_off.push_back(slot_off.back() + q.count);
    share_off.push_back(share_off.back() + q.n_accepted);
    for (size_t i = 0; i < q.count; ++i) {
      const uint8_t v = q.valid ? (q.valid[i] != 0) : 1;
      valid.push_back(v);
      pair_off.push_back(pair_off.back() + (v ? q.n_accepted : 0));
    }
  }
  const size_t K = R.size();
  if (K == 0) return BTE_OK;
  const uint64_t M = m_off.back(), S = slot_off.back(), SH = share_off.back(), NP = pair_off.back();
  if (S > 0xffffffffull) return set_error(BTE_EINVAL, "btx_decrypt_requests: too many ciphertexts");
  Scratch sc(s);
  ReqDev* d_R;
  const void** d_tables;
  uint64_t *d_moff, *d_soff, *d_shoff, *d_poff;
  uint8_t* d_valid;
  g1j *c, *shj;
  g1a *ca, *sa, *P;
  g2a* Q = nullptr;
  const bte_fp6** tab = nullptr;
  fp12 *ml, *conv, *pref, *beta, *mm, *alpha;
  BTE_TRY(sc.upload(&d_R, R.data(), K));
  BTE_TRY(sc.upload(&d_tables, tables.data(), K));
  BTE_TRY(sc.upload(&d_moff, m_off.data(), K + 1));
  BTE_TRY(sc.upload(&d_soff, slot_off.data(), K + 1));
  BTE_TRY(sc.upload(&d_shoff, share_off.data(), K + 1));
  BTE_TRY(sc.upload(&d_poff, pair_off.data(), S + 1));
  BTE_TRY(sc.upload(&d_valid, valid.data(), S));
  BTE_TRY(sc.get(&c, M));
  BTE_TRY(sc.get(&ca, M));
  BTE_TRY(sc.get(&ml, M));
  BTE_TRY(sc.get(&conv, M));
  BTE_TRY(sc.get(&pref, S));
  BTE_TRY(sc.get(&beta, S));
  BTE_TRY(sc.get(&shj, SH));
  BTE_TRY(sc.get(&sa, SH));
  BTE_TRY(sc.get(&P, NP));
  if (all_lines)
    BTE_TRY(sc.get(&tab, NP));
  else
    BTE_TRY(sc.get(&Q, NP));
  BTE_TRY(sc.get(&mm, S));
  BTE_TRY(sc.get(&alpha, S));
  PhaseTimer pt(s);
  pt.begin();
  hipLaunchKernelGGL(k_req_pad, dim3(blocks_for(M, 256)), dim3(256), 0, s, (const ReqDev*)d_R,
                     (const uint64_t*)d_moff, (const uint64_t*)d_soff, (const uint8_t*)d_valid, (uint32_t)K, M, c);
  BTE_TRY(bte_fft_g1_segmented((bte_p1*)c, logs.data(), K, 0, 0, s));
  BTE_TRY(bte_g1_batch_normalize((const bte_p1*)c, M, (bte_p1_affine*)ca, s));
  pt.mark("g1_fft");
  BTE_TRY(miller_prepared_segmented((const void* const*)d_tables, d_moff, K, ca, M, ml, s));
  pt.mark("miller");
  BTE_TRY(bte_final_exp_easy((const bte_fp12*)ml, M, (bte_fp12*)conv, s));
  pt.mark("easy");
  BTE_TRY(cyc_prefix_ifft_segmented(conv, logs.data(), outs.data(), K, s));
  hipLaunchKernelGGL(k_req_gather_prefix, dim3(blocks_for(S, 256)), dim3(256), 0, s, (const fp12*)conv,
                     (const uint64_t*)d_moff, (const uint64_t*)d_soff, (uint32_t)K, S, pref);
  pt.mark("cyclo_ifft");
  BTE_TRY(bte_final_exp_hard((const bte_fp12*)pref, S, (bte_fp12*)beta, s));
  pt.mark("hard");
  hipLaunchKernelGGL(k_req_gather_shares, dim3(blocks_for(SH, 256)), dim3(256), 0, s, (const ReqDev*)d_R,
                     (const uint64_t*)d_shoff, (uint32_t)K, SH, shj);
  BTE_TRY(bte_g1_batch_normalize((const bte_p1*)shj, SH, (bte_p1_affine*)sa, s));
  if (NP)
    hipLaunchKernelGGL(k_req_pairs, dim3(blocks_for(NP, 256)), dim3(256), 0, s, (const ReqDev*)d_R,
                       (const uint64_t*)d_soff, (const uint64_t*)d_poff, (const uint64_t*)d_shoff, (uint32_t)K,
                       (uint32_t)S, NP, (const g1a*)sa, P, Q, tab);
  if (all_lines)
    BTE_TRY(multi_miller_prepared((const void* const*)tab, P, NP, pair_off.data(), S, mm, s));
  else
    BTE_TRY(bte_multi_miller_loop((const bte_p1_affine*)P, (const bte_p2_affine*)Q, NP, d_poff, S, (bte_fp12*)mm, s));
  pt.mark("multi_miller");
  BTE_TRY(bte_final_exp((const bte_fp12*)mm, S, (bte_fp12*)alpha, s));
  pt.mark("final_exp");
  hipLaunchKernelGGL(k_req_combine, dim3(blocks_for(S, 64)), dim3(64), 0, s, (const ReqDev*)d_R,
                     (const uint64_t*)d_soff, (uint32_t)K, S, (const uint8_t*)d_valid, (const fp12*)alpha,
                     (const fp12*)beta);
  pt.mark("combine");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

// open_batch with the opening keys' line tables (bte_g2_prepare_lines over the B x T keys,
// committee state kept resident): same outputs as bte_btx_open_batched; invalid slots run no
// pairs.
extern "C" int bte_btx_open_prepared_batched(const bte_p1* shares, size_t n_accepted, const bte_fp6* key_lines,
                                             const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid,
                                             size_t batch_size, size_t count, bte_fp12* messages_out,
                                             void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, T = n_accepted, K = count;
  if (!shares || !key_lines || !ct2 || !beta || !valid || !messages_out)
    return set_error(BTE_EINVAL, "btx_open_prepared: null pointer");
  if (B == 0 || T == 0) return set_error(BTE_EINVAL, "btx_open_prepared: empty batch or committee");
  if (K == 0) return BTE_OK;
  const size_t npairs = K * B * T;
  Scratch sc(s);
  g1a *sa, *P;
  const bte_fp6** tab;
  fp12 *ml, *alpha;
  uint8_t* d_valid;
  BTE_TRY(sc.get(&sa, K * T));
  BTE_TRY(sc.get(&P, npairs));
  BTE_TRY(sc.get(&tab, npairs));
  BTE_TRY(sc.get(&ml, K * B));
  BTE_TRY(sc.get(&alpha, K * B));
  BTE_TRY(sc.upload(&d_valid, valid, K * B));
  PhaseTimer pt(s);
  pt.begin();
  BTE_TRY(bte_g1_batch_normalize(shares, K * T, (bte_p1_affine*)sa, s));
  hipLaunchKernelGGL(k_open_pairs_prepared, dim3(blocks_for(npairs, 256)), dim3(256), 0, s, (const g1a*)sa,
                     key_lines, T, B, K, P, tab);
  // slot j owns pairs [j T, j T + T) when valid; an invalid slot's segment is empty
  std::vector<uint64_t> lo(K * B), hi(K * B);
  for (size_t j = 0; j < K * B; ++j) {
    lo[j] = j * T;
    hi[j] = valid[j] ? j * T + T : j * T;
  }
  BTE_TRY(multi_miller_prepared_ranges((const void* const*)tab, P, lo.data(), hi.data(), K * B, ml, s));
  pt.mark("multi_miller");
  BTE_TRY(bte_final_exp((const bte_fp12*)ml, K * B, (bte_fp12*)alpha, s));
  pt.mark("final_exp");
  hipLaunchKernelGGL(k_open_combine, dim3(blocks_for(K * B, 64)), dim3(64), 0, s, (const fp12*)ct2,
                     (const fp12*)alpha, (const fp12*)beta, (const uint8_t*)d_valid, K * B, (fp12*)messages_out);
  pt.mark("combine");
  BTE_CHECK_LAUNCH();
  if (pt.on) {
    BTE_HIP(hipStreamSynchronize(s));
    pt.publish();
  }
  return BTE_OK;
}

// decryption.rs:205-233 partial_decrypt for every party at once: sigma_j = sum_i q_j^i C_i
// (i = 1..B).  q: N canonical scalars (host); ct1: B G1 affine (device); sigma_out: N Jacobian.
extern "C" int bte_btx_partial_decrypt(const bte_p1_affine* ct1, size_t batch_size, const uint8_t* q, size_t n_parties,
                                       bte_p1* sigma_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size, N = n_parties;
  if (!ct1 || !q || !sigma_out) return set_error(BTE_EINVAL, "btx_partial_decrypt: null pointer");
  if (N == 0) return BTE_OK;
  std::vector<bteh::Fr> pw(N * B);
  std::vector<uint32_t> pidx(N * B);
  std::vector<uint64_t> off(N + 1);
  for (size_t j = 0; j < N; ++j) {
    bteh::Fr qj;
    memcpy(qj.l, q + 32 * j, 32);
    bteh::Fr cur = qj;
    off[j] = j * B;
    for (size_t i = 0; i < B; ++i) {
      pw[j * B + i] = cur;
      pidx[j * B + i] = (uint32_t)i;
      cur = bteh::fr_mul(cur, qj);
    }
  }
  off[N] = N * B;
  Scratch sc(s);
  uint8_t* d_pw;
  uint32_t* d_pidx;
  uint64_t* d_off;
  BTE_TRY(sc.upload(&d_pw, (const uint8_t*)pw.data(), 32 * N * B));
  BTE_TRY(sc.upload(&d_pidx, pidx.data(), N * B));
  BTE_TRY(sc.upload(&d_off, off.data(), N + 1));
  BTE_TRY(g1_msm_batched_gather(ct1, B, d_pidx, d_pw, nullptr, N * B, d_off, N, 0, sigma_out, s));
  return BTE_OK;
}

// ---------------------------------------------------------------------------------------------
// Launch-only plan of the fused decrypt step (precompute_batch + open_batch over `count`
// batches of one committee, the opening keys prepared), and its HIP graph.  Creation allocates
// every buffer, uploads the slot validity and the multi-Miller chunk ranges and builds the FFT
// domain tables; a run then only enqueues kernels, so it can be captured and replayed.
// ---------------------------------------------------------------------------------------------
struct bte_btx_plan {
  size_t B = 0, K = 0, T = 0, m = 0, nchunks = 0;
  const bte_fp6* lines = nullptr;
  const bte_fp6* key_lines = nullptr;
  g1j* c = nullptr;
  g1a *ca = nullptr, *sa = nullptr, *P = nullptr;
  fp12 *ml = nullptr, *conv = nullptr, *pref = nullptr, *beta = nullptr, *mm = nullptr, *alpha = nullptr,
       *part = nullptr;
  const bte_fp6** tab = nullptr;
  uint8_t* valid = nullptr;
  uint32_t* cyctab = nullptr;
  size_t c3max = 0;  // three-lane switch the plan was created with
  uint64_t *lo = nullptr, *hi = nullptr, *seg = nullptr;
  std::vector<void*> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* bound[4] = {nullptr, nullptr, nullptr, nullptr};
  template <class T> int alloc(T** p, size_t count) {
    hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "btx plan: %s", hipGetErrorString(e));
    allocs.push_back((void*)*p);
    return BTE_OK;
  }
  ~bte_btx_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : allocs) (void)hipFree(p);
  }
};

extern "C" int bte_btx_plan_create(size_t batch_size, size_t count, size_t n_accepted, const bte_fp6* kernel_lines,
                                   const bte_fp6* key_lines, const uint8_t* valid, bte_btx_plan** out,
                                   void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!out || !kernel_lines || !key_lines || !valid) return set_error(BTE_EINVAL, "btx_plan_create: null pointer");
  if (batch_size == 0 || count == 0 || n_accepted == 0) return set_error(BTE_EINVAL, "btx_plan_create: empty shape");
  *out = nullptr;
  bte_btx_plan* p = new bte_btx_plan();
  p->B = batch_size;
  p->K = count;
  p->T = n_accepted;
  p->m = next_pow2(2 * batch_size);
  p->lines = kernel_lines;
  p->key_lines = key_lines;
  const size_t M = p->m * p->K, S = p->B * p->K, NP = S * p->T;
  std::vector<uint64_t> lo(S), hi(S);
  for (size_t j = 0; j < S; ++j) {
    lo[j] = j * p->T;
    hi[j] = valid[j] ? j * p->T + p->T : j * p->T;
  }
  MMChunks ch;
  mm_chunks(lo.data(), hi.data(), S, ch);
  p->nchunks = ch.lo.size();
  p->c3max = (size_t)tune_get(BTE_TUNE_CYC_THREE_LANE_MAX);
  const size_t ctab = cyc_tab_bytes(p->m, p->K, p->c3max);
  int rc = BTE_OK;
  do {
    if ((rc = p->alloc(&p->c, M)) || (rc = p->alloc(&p->ca, M)) || (rc = p->alloc(&p->ml, M)) ||
        (rc = p->alloc(&p->conv, M)) || (rc = p->alloc(&p->pref, S)) || (rc = p->alloc(&p->beta, S)) ||
        (rc = p->alloc(&p->sa, p->K * p->T)) || (rc = p->alloc(&p->P, NP)) || (rc = p->alloc(&p->tab, NP)) ||
        (rc = p->alloc(&p->mm, S)) || (rc = p->alloc(&p->alpha, S)) || (rc = p->alloc(&p->valid, S)) ||
        (rc = p->alloc(&p->lo, p->nchunks)) || (rc = p->alloc(&p->hi, p->nchunks)) ||
        (rc = p->alloc(&p->seg, ch.seg.size())) || (rc = p->alloc(&p->part, p->nchunks)))
      break;
    if (ctab && (rc = p->alloc((uint8_t**)&p->cyctab, ctab))) break;
    if (hipMemcpy(p->valid, valid, S, hipMemcpyHostToDevice) != hipSuccess ||
        (p->nchunks && (hipMemcpy(p->lo, ch.lo.data(), p->nchunks * 8, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(p->hi, ch.hi.data(), p->nchunks * 8, hipMemcpyHostToDevice) != hipSuccess)) ||
        hipMemcpy(p->seg, ch.seg.data(), ch.seg.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
      rc = set_error(BTE_EDEVICE, "btx plan: upload failed");
      break;
    }
    if ((rc = fft_warm(p->m, false, false, s)) || (rc = fft_warm(p->m, true, true, s))) break;
  } while (0);
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return BTE_OK;
}

// Kernel launches only (no allocation, copy or synchronisation).
static int plan_enqueue(bte_btx_plan* p, const bte_p1_affine* ct1, const bte_p1* shares, const bte_fp12* ct2,
                        bte_fp12* messages_out, hipStream_t s) {
  const size_t B = p->B, K = p->K, T = p->T, m = p->m, M = m * K, S = B * K;
  hipLaunchKernelGGL(k_g1_pad, dim3(blocks_for(M, 256)), dim3(256), 0, s, (const g1a*)ct1,
                     (const uint8_t*)p->valid, B, m, K, p->c);
  BTE_TRY(bte_fft_g1_batched((bte_p1*)p->c, m, K, 0, 0, s));
  BTE_TRY(bte_g1_batch_normalize((const bte_p1*)p->c, M, (bte_p1_affine*)p->ca, s));
  BTE_TRY(bte_miller_loop_prepared_tiled(p->lines, m, (const bte_p1_affine*)p->ca, M, (bte_fp12*)p->ml, s));
  BTE_TRY(bte_final_exp_easy((const bte_fp12*)p->ml, M, (bte_fp12*)p->conv, s));
  BTE_TRY(cyc_prefix_batched_enqueue(p->conv, m, K, B, p->cyctab, p->c3max, s));
  hipLaunchKernelGGL(k_gather_prefix, dim3(blocks_for(S, 256)), dim3(256), 0, s, (const fp12*)p->conv, m, B, K,
                     p->pref);
  BTE_TRY(bte_final_exp_hard((const bte_fp12*)p->pref, S, (bte_fp12*)p->beta, s));
  BTE_TRY(bte_g1_batch_normalize(shares, K * T, (bte_p1_affine*)p->sa, s));
  hipLaunchKernelGGL(k_open_pairs_prepared, dim3(blocks_for(S * T, 256)), dim3(256), 0, s, (const g1a*)p->sa,
                     p->key_lines, T, B, K, p->P, p->tab);
  BTE_TRY(mm_enqueue((const void* const*)p->tab, p->P, p->lo, p->hi, p->nchunks, p->seg, S, p->part, p->mm, s));
  BTE_TRY(bte_final_exp((const bte_fp12*)p->mm, S, (bte_fp12*)p->alpha, s));
  hipLaunchKernelGGL(k_open_combine, dim3(blocks_for(S, 64)), dim3(64), 0, s, (const fp12*)ct2,
                     (const fp12*)p->alpha, (const fp12*)p->beta, (const uint8_t*)p->valid, S, (fp12*)messages_out);
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

extern "C" int bte_btx_plan_run(bte_btx_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                                const bte_fp12* ct2, bte_fp12* messages_out, void* stream) {
  if (!plan || !ct1 || !shares || !ct2 || !messages_out) return set_error(BTE_EINVAL, "btx_plan_run: null pointer");
  return plan_enqueue(plan, ct1, shares, ct2, messages_out, (hipStream_t)stream);
}

extern "C" int bte_btx_plan_launch_graph(bte_btx_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                                         const bte_fp12* ct2, bte_fp12* messages_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!plan || !ct1 || !shares || !ct2 || !messages_out) return set_error(BTE_EINVAL, "btx_plan_graph: null pointer");
  if (!s) return set_error(BTE_EINVAL, "btx_plan_graph: the null stream cannot be captured");
  const void* args[4] = {ct1, shares, ct2, messages_out};
  if (!plan->exec) {
    if (PhaseTimer(s).on)  // phase timing synchronises the stream, which a capture cannot contain
      return set_error(BTE_EINVAL, "btx_plan_graph: disable profiling before the first launch");
    BTE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = plan_enqueue(plan, ct1, shares, ct2, messages_out, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_error(BTE_EDEVICE, "btx_plan_graph: capture failed: %s", hipGetErrorString(e));
    plan->graph = g;
    BTE_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
    for (int i = 0; i < 4; ++i) plan->bound[i] = args[i];
  } else {
    for (int i = 0; i < 4; ++i)
      if (plan->bound[i] != args[i])
        return set_error(BTE_EINVAL, "btx_plan_graph: the graph is bound to the buffers of its first launch");
  }
  BTE_HIP(hipGraphLaunch(plan->exec, s));
  return BTE_OK;
}

extern "C" int bte_btx_plan_destroy(bte_btx_plan* plan) {
  delete plan;
  return BTE_OK;
}
