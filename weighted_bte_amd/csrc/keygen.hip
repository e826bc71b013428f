// Key material on the GPU: the scalar tables and fixed-base G2 multiplications of the
// trusted-dealer setups of weighted_btx (setup.rs:98-233) and weighted_pfe (setup.rs:149-354),
// given the dealer's random choices (the Shamir polynomial, every party's q or rho, pfe's alpha
// and beta) so that the output is a pure function of them.
//
//   z_v = poly(x_v)                      x_v = omega^v, the first W points of the radix-2 domain
//                                        of size nextpow2(W) (setup.rs:111-131)
//   negative block i (i = 0..L-1): [z_v q_owner(v)^-(i+1)]_2        (setup.rs:166-177)
//   positive block i (i = 0..L-2): [z_v q_owner(v)^(i+1)]_2         (setup.rs:179-189)
//   verification key (i, j):       [q_j^(i+1)]_2, i = 0..L-1         (setup.rs:159-164)
//
// The polynomial is evaluated by Horner at each point (the reference uses a scalar FFT; the
// values are the same), the power tables are N x L, and the material is generated block by
// block through the fixed-base multiplier so the Jacobian scratch stays bounded.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "curve.hpp"
#include "fr.hpp"
#include "host_field.hpp"

extern "C" int bte_g2_generator_mul(const uint8_t* scalars, size_t n, bte_p2_affine* out, void* stream);

namespace bte {

__global__ void k_fr_to_mont(fr* v, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = fr_to_mont(v[i]);
}

// z[v] = sum_k c_k x_v^k (Montgomery in and out).
__global__ void __launch_bounds__(256) k_poly_eval(const fr* __restrict__ coeffs, size_t ncoef,
                                                   const fr* __restrict__ x, size_t W, fr* __restrict__ z) {
  const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (v >= W) return;
  const fr xv = x[v];
  fr acc = coeffs[ncoef - 1];
  for (size_t k = ncoef - 1; k-- > 0;) acc = fr_add(fr_mont_mul(acc, xv), coeffs[k]);
  z[v] = acc;
}

// out[i N + j] = base_j^(i+1), i < L (Montgomery).
__global__ void __launch_bounds__(256) k_fr_powers(const fr* __restrict__ base, size_t N, size_t L,
                                                   fr* __restrict__ out) {
  const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (j >= N) return;
  const fr b = base[j];
  fr acc = b;
  for (size_t i = 0; i < L; ++i) {
    out[i * N + j] = acc;
    acc = fr_mont_mul(acc, b);
  }
}

// Canonical scalars of nblk material blocks starting at power row `row0`:
// out[k W + v] = z_v * pw[(row0 + k) N + owner_v].
__global__ void __launch_bounds__(256) k_material_scalars(const fr* __restrict__ z, const uint32_t* __restrict__ owner,
                                                          size_t W, const fr* __restrict__ pw, size_t N, size_t row0,
                                                          size_t nblk, fr* __restrict__ out) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= nblk * W) return;
  const size_t k = g / W, v = g - k * W;
  out[g] = fr_from_mont(fr_mont_mul(z[v], pw[(row0 + k) * N + owner[v]]));
}

__global__ void k_from_mont_all(const fr* __restrict__ in, size_t n, fr* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = fr_from_mont(in[i]);
}

}  // namespace bte

using namespace bte;

static bool fr_canonical(const uint8_t* b) {
  bteh::Fr a;
  memcpy(a.l, b, 32);
  for (int i = 3; i >= 0; --i) {
    if (a.l[i] < bteh::FR_R[i]) return true;
    if (a.l[i] > bteh::FR_R[i]) return false;
  }
  return false;
}

extern "C" int bte_btx_keygen(const uint8_t* poly, size_t n_coeffs, const uint8_t* q, const uint64_t* party_offsets,
                              size_t n_parties, size_t max_batch, bte_p2_affine* material,
                              bte_p2_affine* verification_keys, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t N = n_parties, L = max_batch;
  if (!poly || !q || !party_offsets || !material || n_coeffs == 0 || N == 0 || L == 0)
    return set_error(BTE_EINVAL, "btx_keygen: bad arguments");
  const size_t W = party_offsets[N];
  if (party_offsets[0] != 0 || W == 0 || W > (1u << 28)) return set_error(BTE_EINVAL, "btx_keygen: bad party offsets");
  for (size_t j = 0; j < N; ++j)
    if (party_offsets[j + 1] <= party_offsets[j]) return set_error(BTE_EINVAL, "btx_keygen: party %zu has no weight", j);
  for (size_t k = 0; k < n_coeffs; ++k)
    if (!fr_canonical(poly + 32 * k)) return set_error(BTE_EINVAL, "btx_keygen: coefficient %zu not canonical", k);
  // host: domain points, owners, q^-1
  int logd = 0;
  while (((size_t)1 << logd) < W) ++logd;
  bteh::Fr w;
  memcpy(w.l, bteh::FR_ROOT_OF_UNITY, 32);
  for (int i = logd; i < bteh::FR_S; ++i) w = bteh::fr_mul(w, w);
  std::vector<bteh::Fr> xs(W), qs(N), qinv(N);
  xs[0] = bteh::fr_from_u64(1);
  for (size_t v = 1; v < W; ++v) xs[v] = bteh::fr_mul(xs[v - 1], w);
  for (size_t j = 0; j < N; ++j) {
    memcpy(qs[j].l, q + 32 * j, 32);
    if (!fr_canonical(q + 32 * j) || bteh::fr_is_zero(qs[j]))
      return set_error(BTE_EINVAL, "btx_keygen: q of party %zu must be a nonzero canonical scalar", j);
    qinv[j] = bteh::fr_inv(qs[j]);
  }
  std::vector<uint32_t> owner(W);
  for (size_t j = 0; j < N; ++j)
    for (uint64_t v = party_offsets[j]; v < party_offsets[j + 1]; ++v) owner[v] = (uint32_t)j;

  fr *d_coef, *d_x, *d_z, *d_q, *d_qi, *d_pw, *d_nw, *d_sc;
  uint32_t* d_owner;
  const size_t chunk_pts = std::max<size_t>(W, (size_t)1 << 20);  // points per generator call
  const size_t blk_per_chunk = std::max<size_t>(1, chunk_pts / W);
  std::vector<void*> bufs;
  auto get = [&](void** p, size_t bytes) -> int {
    hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, s);
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "btx_keygen: %s", hipGetErrorString(e));
    bufs.push_back(*p);
    return BTE_OK;
  };
  int rc = BTE_OK;
  do {
    if ((rc = get((void**)&d_coef, 32 * n_coeffs)) || (rc = get((void**)&d_x, 32 * W)) ||
        (rc = get((void**)&d_z, 32 * W)) || (rc = get((void**)&d_q, 32 * N)) || (rc = get((void**)&d_qi, 32 * N)) ||
        (rc = get((void**)&d_pw, 32 * N * L)) || (rc = get((void**)&d_nw, 32 * N * L)) ||
        (rc = get((void**)&d_sc, 32 * blk_per_chunk * W)) || (rc = get((void**)&d_owner, 4 * W)))
      break;
    if (hipMemcpyAsync(d_coef, poly, 32 * n_coeffs, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_x, xs.data(), 32 * W, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_q, qs.data(), 32 * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_qi, qinv.data(), 32 * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_owner, owner.data(), 4 * W, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(BTE_EDEVICE, "btx_keygen: upload failed");
      break;
    }
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(n_coeffs, 256)), dim3(256), 0, s, d_coef, n_coeffs);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(W, 256)), dim3(256), 0, s, d_x, W);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(N, 256)), dim3(256), 0, s, d_q, N);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(N, 256)), dim3(256), 0, s, d_qi, N);
    hipLaunchKernelGGL(k_poly_eval, dim3(blocks_for(W, 256)), dim3(256), 0, s, (const fr*)d_coef, n_coeffs,
                       (const fr*)d_x, W, d_z);
    hipLaunchKernelGGL(k_fr_powers, dim3(blocks_for(N, 256)), dim3(256), 0, s, (const fr*)d_q, N, L, d_pw);
    hipLaunchKernelGGL(k_fr_powers, dim3(blocks_for(N, 256)), dim3(256), 0, s, (const fr*)d_qi, N, L, d_nw);
    // negative blocks 0..L-1 (q^-(i+1)), then positive blocks 0..L-2 (q^(i+1))
    const size_t total_blocks = 2 * L - 1;
    size_t b0 = 0;
    while (b0 < total_blocks && !rc) {
      const bool neg = b0 < L;
      const size_t nb = std::min(blk_per_chunk, (neg ? L : total_blocks) - b0);  // one power table per chunk
      const fr* pw = neg ? d_nw : d_pw;
      const size_t row0 = neg ? b0 : b0 - L;
      hipLaunchKernelGGL(k_material_scalars, dim3(blocks_for(nb * W, 256)), dim3(256), 0, s, (const fr*)d_z,
                         (const uint32_t*)d_owner, W, pw, N, row0, nb, d_sc);
      rc = bte_g2_generator_mul((const uint8_t*)d_sc, nb * W, material + b0 * W, s);
      b0 += nb;
    }
    if (rc) break;
    if (verification_keys) {
      fr* d_vk;
      if ((rc = get((void**)&d_vk, 32 * N * L))) break;
      hipLaunchKernelGGL(k_from_mont_all, dim3(blocks_for(N * L, 256)), dim3(256), 0, s, (const fr*)d_pw, N * L,
                         d_vk);
      rc = bte_g2_generator_mul((const uint8_t*)d_vk, N * L, verification_keys, s);
    }
  } while (0);
  for (void* p : bufs) (void)hipFreeAsync(p, s);
  if (rc) return rc;
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}

// weighted_pfe (setup.rs:149-354): g_{j,i} = 1 / (rho_j + alpha_i), u_{j,i} = g_{j,i}^2 +
// beta_{-1} / (alpha_i + 1) + beta_0 / alpha_i (host, one batched inversion each), then on the
// device d1 block i = [z_v g_{owner(v),i}]_2 and d2 block i = [z_v u_{owner(v),i}]_2 over the W
// share points, verification keys [g_{j,i}]_2 slot-major and the global key [z (beta_{-1} +
// beta_0)]_2 with z = poly(0).
static bool batch_invert(std::vector<bteh::Fr>& v) {
  std::vector<bteh::Fr> pre(v.size());
  bteh::Fr acc = bteh::fr_from_u64(1);
  for (size_t i = 0; i < v.size(); ++i) {
    if (bteh::fr_is_zero(v[i])) return false;
    pre[i] = acc;
    acc = bteh::fr_mul(acc, v[i]);
  }
  bteh::Fr inv = bteh::fr_inv(acc);
  for (size_t i = v.size(); i-- > 0;) {
    const bteh::Fr vi = v[i];
    v[i] = bteh::fr_mul(inv, pre[i]);
    inv = bteh::fr_mul(inv, vi);
  }
  return true;
}

extern "C" int bte_pfe_keygen(const uint8_t* poly, size_t n_coeffs, const uint8_t* rho, const uint8_t* alpha,
                              const uint8_t* beta, const uint64_t* party_offsets, size_t n_parties, size_t batch_size,
                              bte_p2_affine* material, bte_p2_affine* verification_keys, bte_p2_affine* global_key,
                              uint8_t* fractions_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t N = n_parties, B = batch_size;
  if (!poly || !rho || !alpha || !beta || !party_offsets || !material || n_coeffs == 0 || N == 0 || B == 0)
    return set_error(BTE_EINVAL, "pfe_keygen: bad arguments");
  const size_t W = party_offsets[N];
  if (party_offsets[0] != 0 || W == 0 || W > (1u << 28)) return set_error(BTE_EINVAL, "pfe_keygen: bad party offsets");
  for (size_t j = 0; j < N; ++j)
    if (party_offsets[j + 1] <= party_offsets[j]) return set_error(BTE_EINVAL, "pfe_keygen: party %zu has no weight", j);
  for (size_t k = 0; k < n_coeffs; ++k)
    if (!fr_canonical(poly + 32 * k)) return set_error(BTE_EINVAL, "pfe_keygen: coefficient %zu not canonical", k);
  for (size_t j = 0; j < N; ++j)
    if (!fr_canonical(rho + 32 * j)) return set_error(BTE_EINVAL, "pfe_keygen: rho %zu not canonical", j);
  for (size_t i = 0; i < B; ++i)
    if (!fr_canonical(alpha + 32 * i)) return set_error(BTE_EINVAL, "pfe_keygen: alpha %zu not canonical", i);
  if (!fr_canonical(beta) || !fr_canonical(beta + 32)) return set_error(BTE_EINVAL, "pfe_keygen: beta not canonical");
  bteh::Fr bm1, b0;
  memcpy(bm1.l, beta, 32);
  memcpy(b0.l, beta + 32, 32);
  std::vector<bteh::Fr> al(B), g(N * B), ai(B), a1i(B);
  for (size_t i = 0; i < B; ++i) {
    memcpy(al[i].l, alpha + 32 * i, 32);
    ai[i] = al[i];
    a1i[i] = bteh::fr_add(al[i], bteh::fr_from_u64(1));
  }
  for (size_t j = 0; j < N; ++j) {
    bteh::Fr rj;
    memcpy(rj.l, rho + 32 * j, 32);
    for (size_t i = 0; i < B; ++i) g[j * B + i] = bteh::fr_add(rj, al[i]);
  }
  if (!batch_invert(g)) return set_error(BTE_EINVAL, "pfe_keygen: a label rho_j + alpha_i is zero");
  if (!batch_invert(ai) || !batch_invert(a1i)) return set_error(BTE_EINVAL, "pfe_keygen: alpha_i or alpha_i + 1 is zero");
  std::vector<bteh::Fr> gT(B * N), uT(B * N);  // slot-major: [i N + j]
  for (size_t j = 0; j < N; ++j)
    for (size_t i = 0; i < B; ++i) {
      const bteh::Fr gj = g[j * B + i];
      gT[i * N + j] = gj;
      uT[i * N + j] = bteh::fr_add(bteh::fr_mul(gj, gj),
                                   bteh::fr_add(bteh::fr_mul(bm1, a1i[i]), bteh::fr_mul(b0, ai[i])));
    }
  if (fractions_out) memcpy(fractions_out, g.data(), 32 * N * B);
  int logd = 0;
  while (((size_t)1 << logd) < W) ++logd;
  bteh::Fr w;
  memcpy(w.l, bteh::FR_ROOT_OF_UNITY, 32);
  for (int i = logd; i < bteh::FR_S; ++i) w = bteh::fr_mul(w, w);
  std::vector<bteh::Fr> xs(W);
  xs[0] = bteh::fr_from_u64(1);
  for (size_t v = 1; v < W; ++v) xs[v] = bteh::fr_mul(xs[v - 1], w);
  std::vector<uint32_t> owner(W);
  for (size_t j = 0; j < N; ++j)
    for (uint64_t v = party_offsets[j]; v < party_offsets[j + 1]; ++v) owner[v] = (uint32_t)j;
  bteh::Fr z0, gk;
  memcpy(z0.l, poly, 32);
  gk = bteh::fr_mul(z0, bteh::fr_add(bm1, b0));

  fr *d_coef, *d_x, *d_z, *d_gc, *d_gm, *d_um, *d_sc, *d_gk;
  uint32_t* d_owner;
  const size_t chunk_pts = std::max<size_t>(W, (size_t)1 << 20);
  const size_t blk_per_chunk = std::max<size_t>(1, chunk_pts / W);
  std::vector<void*> bufs;
  auto get = [&](void** p, size_t bytes) -> int {
    hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, s);
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "pfe_keygen: %s", hipGetErrorString(e));
    bufs.push_back(*p);
    return BTE_OK;
  };
  int rc = BTE_OK;
  do {
    if ((rc = get((void**)&d_coef, 32 * n_coeffs)) || (rc = get((void**)&d_x, 32 * W)) ||
        (rc = get((void**)&d_z, 32 * W)) || (rc = get((void**)&d_gc, 32 * B * N)) ||
        (rc = get((void**)&d_gm, 32 * B * N)) || (rc = get((void**)&d_um, 32 * B * N)) ||
        (rc = get((void**)&d_sc, 32 * blk_per_chunk * W)) || (rc = get((void**)&d_gk, 32)) ||
        (rc = get((void**)&d_owner, 4 * W)))
      break;
    if (hipMemcpyAsync(d_coef, poly, 32 * n_coeffs, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_x, xs.data(), 32 * W, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_gc, gT.data(), 32 * B * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_gm, gT.data(), 32 * B * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_um, uT.data(), 32 * B * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_gk, gk.l, 32, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_owner, owner.data(), 4 * W, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(BTE_EDEVICE, "pfe_keygen: upload failed");
      break;
    }
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(n_coeffs, 256)), dim3(256), 0, s, d_coef, n_coeffs);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(W, 256)), dim3(256), 0, s, d_x, W);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, d_gm, B * N);
    hipLaunchKernelGGL(k_fr_to_mont, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, d_um, B * N);
    hipLaunchKernelGGL(k_poly_eval, dim3(blocks_for(W, 256)), dim3(256), 0, s, (const fr*)d_coef, n_coeffs,
                       (const fr*)d_x, W, d_z);
    // d1 blocks 0..B-1 (g), then d2 blocks 0..B-1 (u)
    size_t b0i = 0;
    while (b0i < 2 * B && !rc) {
      const bool first = b0i < B;
      const size_t nb = std::min(blk_per_chunk, (first ? B : 2 * B) - b0i);
      hipLaunchKernelGGL(k_material_scalars, dim3(blocks_for(nb * W, 256)), dim3(256), 0, s, (const fr*)d_z,
                         (const uint32_t*)d_owner, W, (const fr*)(first ? d_gm : d_um), N, first ? b0i : b0i - B, nb,
                         d_sc);
      rc = bte_g2_generator_mul((const uint8_t*)d_sc, nb * W, material + b0i * W, s);
      b0i += nb;
    }
    if (rc) break;
    if (verification_keys && (rc = bte_g2_generator_mul((const uint8_t*)d_gc, B * N, verification_keys, s))) break;
    if (global_key && (rc = bte_g2_generator_mul((const uint8_t*)d_gk, 1, global_key, s))) break;
  } while (0);
  for (void* p : bufs) (void)hipFreeAsync(p, s);
  if (rc) return rc;
  BTE_CHECK_LAUNCH();
  return BTE_OK;
}
