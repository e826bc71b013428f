// weighted_btx decryption-share verification on the GPU:
//   verify_decryption_share          weighted_btx/src/decryption.rs:236-272
//   verify_decryption_shares_batched weighted_btx/src/decryption.rs:356-422
// A share sigma_j of party j is valid when prod_{valid slots i} e(C_i, vk_{i+1,j}) e(-sigma_j, G2)
// is one, vk_{i,j} = [q_j^i]_2.  The batched check folds the shares with caller-chosen nonzero
// challenges rho_j: sigma = sum rho_j sigma_j (one G1 MSM) and V_i = sum_j rho_j vk_{i+1,j} (one
// indexed G2 MSM per valid slot, all in one batched launch sequence), then one multi-pairing.
// Challenge sampling (OsRng in the reference) stays with the caller.
//
// Also the per-ciphertext Schnorr check of weighted_btx/src/proof.rs:56-70 (SchnorrProof::verify)
// for a whole batch: s_i G == A_i + e_i C_i, with the Fiat-Shamir challenges e_i hashed on the host.
#include <string.h>

#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "curve.hpp"
#include "cyclo.hpp"
#include "host_field.hpp"

extern "C" {
int bte_g1_msm(const bte_p1_affine* points, const uint8_t* scalars, size_t n, bte_p1* out, void* stream);
int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream);
int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream);
int bte_multi_miller_loop(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs, const uint64_t* seg_offsets,
                          size_t nseg, bte_fp12* out, void* stream);
int bte_final_exp(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
int bte_g1_generator_mul(const uint8_t* scalars, size_t n, bte_p1_affine* out, void* stream);
}

namespace bte {

BTE_DEV g2a g2_gen_affine() {
  g2a g;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    g.x.c0.l[j] = G2_GEN[0][j];
    g.x.c1.l[j] = G2_GEN[1][j];
    g.y.c0.l[j] = G2_GEN[2][j];
    g.y.c1.l[j] = G2_GEN[3][j];
  }
  return g;
}

// Pairs of check j (nv + 1 each): (C_slot[k], vk[slot[k] N + party[j]]) for k < nv, then
// (-sigma_j, G2).  sigma: affine, one per check.
__global__ void k_verify_pairs(const g1a* __restrict__ ct1, const uint32_t* __restrict__ slots, size_t nv,
                               const g2a* __restrict__ vks, size_t N, const uint32_t* __restrict__ party,
                               const g1a* __restrict__ sigma, size_t nchecks, g1a* __restrict__ P,
                               g2a* __restrict__ Q) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= nchecks * (nv + 1)) return;
  const size_t j = g / (nv + 1), k = g - j * (nv + 1);
  if (k < nv) {
    P[g] = ct1[slots[k]];
    Q[g] = vks[(size_t)slots[k] * N + party[j]];
  } else {
    P[g] = aff_neg(sigma[j]);
    Q[g] = g2_gen_affine();
  }
}

// Batched form: pairs (C_slot[k], V_k) for k < nv, then (-sigma, G2).
__global__ void k_verify_pairs_folded(const g1a* __restrict__ ct1, const uint32_t* __restrict__ slots, size_t nv,
                                      const g2a* __restrict__ V, const g1a* __restrict__ sigma, g1a* __restrict__ P,
                                      g2a* __restrict__ Q) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k > nv) return;
  if (k < nv) {
    P[k] = ct1[slots[k]];
    Q[k] = V[k];
  } else {
    P[k] = aff_neg(*sigma);
    Q[k] = g2_gen_affine();
  }
}

__global__ void k_fp12_is_one(const fp12* __restrict__ v, size_t n, uint8_t* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fp12_eq(v[i], fp12_one()) ? 1 : 0;
}

// ok_i = (s_i G == A_i + e_i C_i): s_i G comes precomputed (fixed-base byte windows), e_i C_i by
// the G1 endomorphism's joint double-and-add over the host's decomposition of e_i.
__global__ void __launch_bounds__(64) k_schnorr_check(const g1a* __restrict__ sG, const g1a* __restrict__ A,
                                                      const g1a* __restrict__ C, const uint32_t* __restrict__ edec,
                                                      size_t B, uint8_t* __restrict__ ok) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= B) return;
  jac<fp> d = g1_glv_mul(jac_from_aff(C[i]), edec + 8 * i);
  d = jac_add_aff(d, A[i]);
  d = jac_add_aff(d, aff_neg(sG[i]));
  ok[i] = jac_is_identity(d) ? 1 : 0;
}

// Scratch and uploads through the stream's workspace and pinned staging (ctx.hpp).  These
// checks return their verdicts in host memory, so each call ends by waiting for its stream.
struct VScratch : CallScope {
  using CallScope::CallScope;
  template <class T> int get(T** p, size_t count) { return dev(p, count); }
};

#define V_TRY(x)          \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

static int check_args(const uint32_t* party, size_t T, size_t N, size_t L, size_t B, const uint8_t* valid,
                      std::vector<uint32_t>& slots) {
  if (B == 0 || B > L) return set_error(BTE_EINVAL, "verify_shares: InvalidBatchSize(%zu)", B);
  for (size_t j = 0; j < T; ++j)
    if (party[j] >= N) return set_error(BTE_EINVAL, "verify_shares: InvalidPartyIndex(%u)", party[j]);
  for (size_t i = 0; i < B; ++i)
    if (valid[i]) slots.push_back((uint32_t)i);
  return BTE_OK;
}

}  // namespace bte

using namespace bte;

extern "C" int bte_btx_verify_shares(const bte_p1* shares, const uint32_t* party_index, size_t n_shares,
                                     const bte_p2_affine* verification_keys, size_t n_parties, size_t max_batch,
                                     const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size,
                                     uint8_t* ok_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t T = n_shares, N = n_parties;
  if (T == 0) return BTE_OK;
  if (!shares || !party_index || !verification_keys || !ct1 || !valid || !ok_out)
    return set_error(BTE_EINVAL, "verify_shares: null pointer");
  std::vector<uint32_t> slots;
  V_TRY(check_args(party_index, T, N, max_batch, batch_size, valid, slots));
  const size_t nv = slots.size(), np = T * (nv + 1);
  std::vector<uint64_t> off(T + 1);
  for (size_t j = 0; j <= T; ++j) off[j] = j * (nv + 1);
  VScratch sc(s);
  uint32_t *d_slots, *d_party;
  uint64_t* d_off;
  g1a *sa, *P;
  g2a* Q;
  fp12 *ml, *res;
  uint8_t* d_ok;
  V_TRY(sc.upload(&d_slots, slots.data(), nv));
  V_TRY(sc.upload(&d_party, party_index, T));
  V_TRY(sc.upload(&d_off, off.data(), T + 1));
  V_TRY(sc.get(&sa, T));
  V_TRY(sc.get(&P, np));
  V_TRY(sc.get(&Q, np));
  V_TRY(sc.get(&ml, T));
  V_TRY(sc.get(&res, T));
  V_TRY(sc.get(&d_ok, T));
  V_TRY(bte_g1_batch_normalize(shares, T, (bte_p1_affine*)sa, s));
  hipLaunchKernelGGL(k_verify_pairs, dim3(blocks_for(np, 256)), dim3(256), 0, s, (const g1a*)ct1,
                     (const uint32_t*)d_slots, nv, (const g2a*)verification_keys, N, (const uint32_t*)d_party,
                     (const g1a*)sa, T, P, Q);
  V_TRY(bte_multi_miller_loop((const bte_p1_affine*)P, (const bte_p2_affine*)Q, np, d_off, T, (bte_fp12*)ml, s));
  V_TRY(bte_final_exp((const bte_fp12*)ml, T, (bte_fp12*)res, s));
  hipLaunchKernelGGL(k_fp12_is_one, dim3(blocks_for(T, 64)), dim3(64), 0, s, (const fp12*)res, T, d_ok);
  BTE_HIP(hipMemcpyAsync(ok_out, d_ok, T, hipMemcpyDeviceToHost, s));
  BTE_HIP(hipStreamSynchronize(s));
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

static bool canonical_fr(const uint8_t* b, bteh::Fr* out) {
  memcpy(out->l, b, 32);
  for (int i = 3; i >= 0; --i) {
    if (out->l[i] < bteh::FR_R[i]) return true;
    if (out->l[i] > bteh::FR_R[i]) return false;
  }
  return false;
}

extern "C" int bte_btx_verify_shares_batched(const bte_p1* shares, const uint32_t* party_index,
                                             const uint8_t* challenges, size_t n_shares,
                                             const bte_p2_affine* verification_keys, size_t n_parties,
                                             size_t max_batch, const bte_p1_affine* ct1, const uint8_t* valid,
                                             size_t batch_size, int* result, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t T = n_shares, N = n_parties;
  if (!result) return set_error(BTE_EINVAL, "verify_shares_batched: null result");
  *result = 1;
  if (T == 0) return BTE_OK;
  if (!shares || !party_index || !challenges || !verification_keys || !ct1 || !valid)
    return set_error(BTE_EINVAL, "verify_shares_batched: null pointer");
  // The folded check is sound only for nonzero challenges (decryption.rs:363-370 draws them
  // nonzero): a zero challenge would drop that party's share from both sides.
  for (size_t j = 0; j < T; ++j) {
    bteh::Fr c;
    if (!canonical_fr(challenges + 32 * j, &c) || (c.l[0] | c.l[1] | c.l[2] | c.l[3]) == 0)
      return set_error(BTE_EINVAL, "verify_shares_batched: challenge %zu is zero or not canonical", j);
  }
  std::vector<uint32_t> slots;
  V_TRY(check_args(party_index, T, N, max_batch, batch_size, valid, slots));
  const size_t nv = slots.size();
  // indexed G2 MSM terms: segment k (slot slots[k]) = sum_j rho_j vk[slots[k] N + party_j]
  std::vector<uint32_t> pidx(nv * T), sidx(nv * T);
  std::vector<uint64_t> seg(nv + 1), off2{0, nv + 1};
  for (size_t k = 0; k < nv; ++k) {
    seg[k] = k * T;
    for (size_t j = 0; j < T; ++j) {
      pidx[k * T + j] = (uint32_t)(slots[k] * N + party_index[j]);
      sidx[k * T + j] = (uint32_t)j;
    }
  }
  seg[nv] = nv * T;
  VScratch sc(s);
  uint32_t *d_slots, *d_pidx, *d_sidx;
  uint64_t *d_seg, *d_off2;
  uint8_t* d_chal;
  g1a *sa, *sig_a, *P;
  g1j* sig;
  g2j* Vj;
  g2a *Va, *Q;
  fp12 *ml, *res;
  uint8_t* d_ok;
  V_TRY(sc.upload(&d_slots, slots.data(), nv));
  V_TRY(sc.upload(&d_pidx, pidx.data(), nv * T));
  V_TRY(sc.upload(&d_sidx, sidx.data(), nv * T));
  V_TRY(sc.upload(&d_seg, seg.data(), nv + 1));
  V_TRY(sc.upload(&d_off2, off2.data(), 2));
  V_TRY(sc.upload(&d_chal, challenges, 32 * T));
  V_TRY(sc.get(&sa, T));
  V_TRY(sc.get(&sig, 1));
  V_TRY(sc.get(&sig_a, 1));
  V_TRY(sc.get(&Vj, nv));
  V_TRY(sc.get(&Va, nv));
  V_TRY(sc.get(&P, nv + 1));
  V_TRY(sc.get(&Q, nv + 1));
  V_TRY(sc.get(&ml, 1));
  V_TRY(sc.get(&res, 1));
  V_TRY(sc.get(&d_ok, 1));
  V_TRY(bte_g1_batch_normalize(shares, T, (bte_p1_affine*)sa, s));
  V_TRY(bte_g1_msm((const bte_p1_affine*)sa, d_chal, T, (bte_p1*)sig, s));
  V_TRY(bte_g1_batch_normalize((const bte_p1*)sig, 1, (bte_p1_affine*)sig_a, s));
  if (nv) {
    V_TRY(g2_msm_batched_gather(verification_keys, max_batch * N, d_pidx, d_chal, d_sidx, nv * T, d_seg, nv, 0, Vj,
                                s));
    V_TRY(bte_g2_batch_normalize((const bte_p2*)Vj, nv, (bte_p2_affine*)Va, s));
  }
  hipLaunchKernelGGL(k_verify_pairs_folded, dim3(blocks_for(nv + 1, 64)), dim3(64), 0, s, (const g1a*)ct1,
                     (const uint32_t*)d_slots, nv, (const g2a*)Va, (const g1a*)sig_a, P, Q);
  V_TRY(bte_multi_miller_loop((const bte_p1_affine*)P, (const bte_p2_affine*)Q, nv + 1, d_off2, 1, (bte_fp12*)ml, s));
  V_TRY(bte_final_exp((const bte_fp12*)ml, 1, (bte_fp12*)res, s));
  hipLaunchKernelGGL(k_fp12_is_one, dim3(1), dim3(64), 0, s, (const fp12*)res, (size_t)1, d_ok);
  uint8_t ok = 0;
  BTE_HIP(hipMemcpyAsync(&ok, d_ok, 1, hipMemcpyDeviceToHost, s));
  BTE_HIP(hipStreamSynchronize(s));
  BTE_CHECK_LAUNCH();
  *result = ok;
  return BTE_OK;
}

extern "C" int bte_btx_verify_proofs(const bte_p1_affine* commitments, const bte_p1_affine* ct1,
                                     const uint8_t* responses, const uint8_t* challenges, size_t batch_size,
                                     uint8_t* ok_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t B = batch_size;
  if (B == 0) return BTE_OK;
  if (!commitments || !ct1 || !responses || !challenges || !ok_out)
    return set_error(BTE_EINVAL, "verify_proofs: null pointer");
  std::vector<uint32_t> edec(8 * B);
  for (size_t i = 0; i < B; ++i) {
    bteh::Fr e, r;
    if (!canonical_fr(responses + 32 * i, &r) || !canonical_fr(challenges + 32 * i, &e))
      return set_error(BTE_EINVAL, "verify_proofs: scalar %zu is not canonical", i);
    glv_decompose_g1(e, &edec[8 * i]);
  }
  VScratch sc(s);
  uint8_t *d_resp, *d_ok;
  uint32_t* d_edec;
  g1a* sG;
  V_TRY(sc.upload(&d_resp, responses, 32 * B));
  V_TRY(sc.upload(&d_edec, edec.data(), 8 * B));
  V_TRY(sc.get(&sG, B));
  V_TRY(sc.get(&d_ok, B));
  V_TRY(bte_g1_generator_mul(d_resp, B, (bte_p1_affine*)sG, s));
  hipLaunchKernelGGL(k_schnorr_check, dim3(blocks_for(B, 64)), dim3(64), 0, s, (const g1a*)sG, (const g1a*)commitments,
                     (const g1a*)ct1, (const uint32_t*)d_edec, B, d_ok);
  BTE_HIP(hipMemcpyAsync(ok_out, d_ok, B, hipMemcpyDeviceToHost, s));
  BTE_HIP(hipStreamSynchronize(s));
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}
