/*
 * C ABI of the MI355X (gfx950) BLS12-381 batch-decryption hot path.
 *
 * Every entry point replaces one crate-private L1 function of the reference's weighted_btx /
 * weighted_pfe crates (cited per declaration).  Conventions:
 *   - Memory layouts are blst's: Fp = 6 x u64 little-endian limbs in Montgomery form
 *     (R = 2^384); blst_fp2 {fp[2]}, blst_fp6 {fp2[3]}, blst_fp12 {fp6[2]};
 *     blst_p1_affine {x, y} (identity = all zero), blst_p1 {x, y, z} (Jacobian, z = 0 is the
 *     identity), and the same over Fp2 for G2.  Scalars are canonical 32-byte little-endian.
 *   - Unless a function says otherwise, pointers are DEVICE pointers and work is queued on
 *     `stream` (a hipStream_t, NULL = the default stream); inputs must stay valid until the
 *     stream reaches the call's work.
 *   - Functions return BTE_OK or a nonzero status; they never abort.  Where the reference
 *     panics on a length mismatch, the C ABI returns BTE_EINVAL.  bte_last_error() gives the
 *     message of the most recent failure on the calling thread.
 *   - Entry points are thread-safe across streams.
 */
#ifndef BTE_HIP_H
#define BTE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { BTE_OK = 0, BTE_EINVAL = 1, BTE_ENOMEM = 2, BTE_EDEVICE = 3 };

typedef struct { uint64_t l[6]; } bte_fp;
typedef struct { bte_fp c[2]; } bte_fp2;
typedef struct { bte_fp2 c[3]; } bte_fp6;
typedef struct { bte_fp6 c[2]; } bte_fp12;
typedef struct { bte_fp x, y; } bte_p1_affine;
typedef struct { bte_fp x, y, z; } bte_p1;
typedef struct { bte_fp2 x, y; } bte_p2_affine;
typedef struct { bte_fp2 x, y, z; } bte_p2;

/* Message of the last failure on this thread ("" if none). */
const char* bte_last_error(void);

/* Library version string. */
const char* bte_version(void);

/* ---------------------------------------------------------------------------------------- *
 * a1/a2: multi-scalar multiplication, sum_i s_i * P_i over the low 255 bits of each scalar
 * (blst_p{1,2}s_mult_pippenger with nbits = 255).  n == 0 gives the identity.
 * `out` is one Jacobian point in device memory.  The last window combination (a serial chain of
 * about 256 doublings) runs as a host callback on `stream` between two stream-ordered copies,
 * so bte_g1_msm and bte_g2_msm neither wait nor allocate once the stream's workspace has grown;
 * captured into a graph, the same chain runs in one device thread instead (a few ms).
 * bte_g2_msm_indexed waits for its stream to report out-of-range indices.
 * ---------------------------------------------------------------------------------------- */

/* weighted_btx/src/blst_utils.rs:43-67 g1_multi_exp_affine_bytes */
int bte_g1_msm(const bte_p1_affine* points, const uint8_t* scalars, size_t n, bte_p1* out,
               void* stream);

/* weighted_btx/src/blst_utils.rs:88-99 g2_multi_exp_affine_bytes */
int bte_g2_msm(const bte_p2_affine* points, const uint8_t* scalars, size_t n, bte_p2* out,
               void* stream);

/* weighted_btx/src/blst_utils.rs:106-125 g2_multi_exp_affine_indexed: term i uses
 * points[indices[i]] (indices bounds-checked against n_points on the device: an index out of
 * range makes the call return BTE_EINVAL). */
int bte_g2_msm_indexed(const bte_p2_affine* points, size_t n_points, const uint32_t* indices,
                       const uint8_t* scalars, size_t n, bte_p2* out, void* stream);

/* Batched MSMs of uneven sizes in one asynchronous launch sequence (no host synchronisation,
 * graph-capturable): out[s] = sum over terms t in [seg_offsets[s], seg_offsets[s+1]) of
 * scalars[t] * P(t), P(t) = points[indices[t]] when `indices` is non-NULL else points[t].
 * seg_offsets is a device array of nseg+1 u64 with seg_offsets[nseg] = nterms.  Indices are
 * checked against n_points on the device: a term with an index out of range is skipped and
 * `status` (a DEVICE int, may be NULL) receives BTE_EINVAL in stream order, else BTE_OK; read it
 * after the stream reaches the call (the call itself cannot know without synchronising).
 * Replaces the rayon fan-out of per-party / per-slot MSMs around
 * weighted_btx/src/decryption.rs:205-233 (partial_decrypt) and :465-481 (prepare_decryption). */
int bte_g1_msm_batched(const bte_p1_affine* points, size_t n_points, const uint32_t* indices,
                       const uint8_t* scalars, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                       bte_p1* out, int* status, void* stream);
int bte_g2_msm_batched(const bte_p2_affine* points, size_t n_points, const uint32_t* indices,
                       const uint8_t* scalars, size_t nterms, const uint64_t* seg_offsets, size_t nseg,
                       bte_p2* out, int* status, void* stream);

/* Window width override for the Pippenger kernels (0 = the built-in table by size). */
int bte_msm_set_window(int g1_or_g2, size_t n, int window_bits);

/* ---------------------------------------------------------------------------------------- *
 * a3: Jacobian -> affine, one shared inversion per chunk (blst_p{1,2}s_to_affine).
 * ---------------------------------------------------------------------------------------- */

/* weighted_btx/src/blst_utils.rs:11-24 batch_normalize_g1 */
int bte_g1_batch_normalize(const bte_p1* in, size_t n, bte_p1_affine* out, void* stream);
/* weighted_btx/src/blst_utils.rs:26-36 batch_normalize_g2 */
int bte_g2_batch_normalize(const bte_p2* in, size_t n, bte_p2_affine* out, void* stream);

/* Canonical compressed encodings (blstrs to_compressed / from_compressed: big-endian x, G2 as
 * x.c1 || x.c0, flags 0x80 compressed, 0x40 infinity, 0x20 larger y).  compress: n points ->
 * 48 n (G1) or 96 n (G2) bytes.  decompress: ok[i] = 1 when encoding i has valid flags, x < p,
 * x on the curve and the point in the order-r subgroup (invalid elements decode to the
 * identity).  All arrays DEVICE.  Wire formats used by weighted_pfe/src/encoding.rs:31-45 and
 * the transcripts (weighted_pfe/src/proof.rs:62-65). */
int bte_g1_compress(const bte_p1_affine* points, size_t n, uint8_t* out, void* stream);
int bte_g2_compress(const bte_p2_affine* points, size_t n, uint8_t* out, void* stream);
int bte_g1_decompress(const uint8_t* in, size_t n, bte_p1_affine* points, uint8_t* ok, void* stream);
int bte_g2_decompress(const uint8_t* in, size_t n, bte_p2_affine* points, uint8_t* ok, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * a4-a9: radix-2 transforms over the Fr domain of size n (a power of two <= 2^32), in place.
 * Natural-order DFT X_k = sum_j x_j w^(jk), w = ROOT_OF_UNITY^(2^(32 - log n)); `inverse` uses
 * w^-1; `scaled` multiplies by n^-1 (inverse only).  A size that is not a power of two returns
 * BTE_EINVAL (the reference's Error::InvalidBatchSize).
 * ---------------------------------------------------------------------------------------- */

/* weighted_btx/src/fft.rs:197-234 Radix2Domain::{fft, ifft} over Scalar: `values` holds n
 * canonical 32-byte little-endian scalars (DEVICE), transformed in place; the batched form runs
 * `count` transforms stored back to back.  Used by the Lagrange product tree below. */
int bte_fft_fr(uint8_t* values, size_t n, int inverse, int scaled, void* stream);
int bte_fft_fr_batched(uint8_t* values, size_t n, size_t count, int inverse, int scaled, void* stream);

/* G1/G2 transforms of size 2^2 .. 2^log_n (default 10) with n^2 x count <= 2^20 run as one
 * batched MSM per call (X_k = sum_j w^(jk) x_j, one segment per output) instead of log n
 * latency-bound stages; 0 turns that off.  The outputs are the same points.  No counterpart in
 * the reference (its transform is always staged). */
int bte_fft_set_dft_max_log(int log_n);

/* weighted_btx/src/fft.rs:197-234 Radix2Domain::{fft, ifft, ifft_unscaled} over G1Projective */
int bte_fft_g1(bte_p1* values, size_t n, int inverse, int scaled, void* stream);
/* same over G2Projective */
int bte_fft_g2(bte_p2* values, size_t n, int inverse, int scaled, void* stream);
/* same over Gt (fft.rs:197-234 generic path, fft.rs:242-274 ifft_gt_unscaled) */
int bte_fft_gt(bte_fp12* values, size_t n, int inverse, int scaled, void* stream);
/* `count` independent transforms of size n stored back to back, one launch per stage */
int bte_fft_g1_batched(bte_p1* values, size_t n, size_t count, int inverse, int scaled, void* stream);
int bte_fft_g2_batched(bte_p2* values, size_t n, size_t count, int inverse, int scaled, void* stream);
int bte_fft_cyclotomic_batched(bte_fp12* values, size_t n, size_t count, int inverse, void* stream);
int bte_ifft_cyclotomic_prefix_unscaled_batched(bte_fp12* values, size_t n, size_t count,
                                                size_t output_count, void* stream);
/* `count` transforms of uneven sizes 2^log_sizes[f] (HOST array, each 0..28) stored back to
 * back; one launch per stage level serves all of them (the reference's per-transform rayon
 * loops over Radix2Domain::fft / ifft, batched). */
int bte_fft_g1_segmented(bte_p1* values, const int* log_sizes, size_t count, int inverse,
                         int scaled, void* stream);
int bte_fft_g2_segmented(bte_p2* values, const int* log_sizes, size_t count, int inverse,
                         int scaled, void* stream);
/* weighted_pfe/src/fft.rs:324-331 fft_cyclotomic / ifft_cyclotomic_unscaled */
int bte_fft_cyclotomic(bte_fp12* values, size_t n, int inverse, void* stream);
/* weighted_btx/src/fft.rs:284-350 ifft_cyclotomic_prefix_unscaled: values[0..output_count)
 * receive the first outputs (output_count <= n/2); the rest of the buffer is scratch. */
int bte_ifft_cyclotomic_prefix_unscaled(bte_fp12* values, size_t n, size_t output_count, void* stream);
/* ifft_cyclotomic_prefix_unscaled over `count` transforms of uneven sizes 2^log_sizes[f]
 * stored back to back (HOST arrays); transform f keeps output_counts[f] <= 2^(log_sizes[f]-1)
 * outputs at its start. */
int bte_ifft_cyclotomic_prefix_unscaled_segmented(bte_fp12* values, const int* log_sizes,
                                                  const uint32_t* output_counts, size_t count,
                                                  void* stream);
/* weighted_btx/src/fft.rs:65-71 cyclotomic_mul_scalar_decomposed, element-wise:
 * values[i] <- decomposed action of scalars[i] (canonical 32-byte LE, HOST memory). */
int bte_cyclotomic_pow_scalars(bte_fp12* values, const uint8_t* host_scalars, size_t n, void* stream);
/* weighted_pfe/src/fft.rs:671-697 mul_cyclotomic_signed_small, element-wise (device i64). */
int bte_cyclotomic_pow_small(bte_fp12* values, const int64_t* coefficients, size_t n, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * a10-a14: Miller loops and final exponentiation.  Line tables are 68 x bte_fp6 per G2 point
 * (19,584 bytes, the size of the reference's PreparedG2Lines lines); an identity point gives an
 * all-zero table.  Raw Miller values are only meaningful after the final exponentiation.
 * ---------------------------------------------------------------------------------------- */

/* weighted_btx/src/final_exponentiation.rs:186-195 PreparedG2Lines::from_affine */
int bte_g2_prepare_lines(const bte_p2_affine* q, size_t n, bte_fp6* lines, void* stream);
/* final_exponentiation.rs:207-219 PreparedG2Lines::miller_loop, pair i = (p[i], lines[68 i..]) */
int bte_miller_loop_prepared(const bte_fp6* lines, const bte_p1_affine* p, size_t n, bte_fp12* out,
                             void* stream);
/* as bte_miller_loop_prepared with pair i using table i % ntables */
int bte_miller_loop_prepared_tiled(const bte_fp6* lines, size_t ntables, const bte_p1_affine* p,
                                   size_t n, bte_fp12* out, void* stream);
/* as bte_miller_loop_prepared over several tables: pair i of segment s (seg_offsets[s] <= i <
 * seg_offsets[s+1]) uses table i - seg_offsets[s] of tables[s].  tables (nseg pointers) and
 * seg_offsets (nseg+1) are DEVICE arrays. */
int bte_miller_loop_prepared_segmented(const bte_fp6* const* tables, const uint64_t* seg_offsets,
                                       size_t nseg, const bte_p1_affine* p, size_t n, bte_fp12* out,
                                       void* stream);
/* one unprepared Miller loop per pair (p[i], q[i]) */
int bte_miller_loop_pairs(const bte_p1_affine* p, const bte_p2_affine* q, size_t n, bte_fp12* out,
                          void* stream);
/* final_exponentiation.rs:229-273 unprepared_multi_miller_loop, batched over segments of uneven
 * arity: out[s] = product of the loops of pairs seg_offsets[s] .. seg_offsets[s+1] (device u64
 * offsets, seg_offsets[nseg] = npairs); pairs with an identity are skipped, empty = one. */
int bte_multi_miller_loop(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs,
                          const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream);
/* No counterpart in the reference: bte_multi_miller_loop for callers that cannot allocate (work
 * under graph capture, tight serving loops, harnesses that forbid allocation).  The same kernels
 * and results; the per-pair Miller values live in a file-scope device array of 32 MiB (at most
 * 58,254 pairs; more is BTE_EINVAL) that each GPU the library runs on holds until the process
 * ends.  Nothing is allocated, freed or waited for.  Calls that overlap on one GPU share that
 * array and race: issue them in one stream order. */
int bte_multi_miller_loop_noalloc(const bte_p1_affine* p, const bte_p2_affine* q, size_t npairs,
                                  const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream);
/* unprepared_multi_miller_loop with the G2 side prepared: pair g uses the line table
 * pair_tables[g] (DEVICE array of npairs pointers to 68-line tables) and p[g]; seg_offsets is a
 * HOST array (nseg+1, from 0 to npairs).  The pairs of a segment share the accumulator and its
 * squarings in chunks of up to 8 pairs (fewer when npairs < 2^19, to keep the GPU full).
 * Identity pairs are skipped, empty = one. */
int bte_multi_miller_loop_prepared(const bte_fp6* const* pair_tables, const bte_p1_affine* p, size_t npairs,
                                   const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream);
/* final_exponentiation(unprepared_multi_miller_loop(...)) per segment (final_exponentiation.rs:
 * 229-273 then :348-354, the pairing product the reference's open_batch and share checks
 * finalise), with the same arguments as bte_multi_miller_loop_prepared, in one kernel launch
 * (after a small initialisation launch): the thread that finishes the last Miller chunk of a
 * group of about sqrt(chunks) multiplies the group, the one that finishes a segment's last group
 * multiplies the groups and runs the easy part, and its wavefront then runs the hard parts of the
 * segments it closed on three lanes each.  out[s] equals bte_final_exp of
 * bte_multi_miller_loop_prepared's out[s]; an empty segment gives one. */
int bte_multi_pairing_prepared_fused(const bte_fp6* const* pair_tables, const bte_p1_affine* p, size_t npairs,
                                     const uint64_t* seg_offsets, size_t nseg, bte_fp12* out, void* stream);
/* final_exponentiation.rs:276-305 easy_final_exponentiation / batch_easy_final_exponentiation
 * (inputs nonzero; in and out must not alias) */
int bte_final_exp_easy(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
/* final_exponentiation.rs:308-336 hard_final_exponentiation (in place allowed); below the
 * three-lane switch (BTE_TUNE_CYC_THREE_LANE_MAX elements) each element runs on three lanes,
 * with the same values */
int bte_final_exp_hard(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);
/* final_exponentiation.rs:348-354 full_final_exponentiation (in and out must not alias) */
int bte_final_exp(const bte_fp12* in, size_t n, bte_fp12* out, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * a16: weighted_btx batch-decryption phases (host control flow over the kernels above).
 * Key material (device) is one G2 affine array: the reference's negative_material (L blocks of
 * W points) followed by its positive_material (L-1 blocks of W).  Host arrays are marked.
 * ---------------------------------------------------------------------------------------- */

/* weighted_btx/src/interpolation.rs:31-52 lagrange_at_zero over the first W points of the
 * radix-2 domain of size nextpow2(W); selected (n indices) and out (n x 32 bytes) are HOST.
 * Above 64 points the coefficients come from the fast path of interpolation.rs:101-197 on the
 * device (product tree of Fr transforms, then P' on the whole domain); a repeated or
 * out-of-range index is BTE_EINVAL (Error::InvalidInterpolationSet). */
int bte_lagrange_at_zero(size_t total_weight, const uint64_t* selected, size_t n, uint8_t* out);
/* The same with `out` in DEVICE memory, queued on `stream` (selected stays HOST). */
int bte_lagrange_at_zero_device(size_t total_weight, const uint64_t* selected, size_t n, uint8_t* out,
                                void* stream);

/* weighted_btx/src/decryption.rs:427-551 prepare_decryption.  party_offsets (N+1) and accepted
 * (T party indices, committee order) are HOST.  Outputs (device): kernel_lines, nextpow2(2B)
 * tables of 68 lines; opening_keys, B x T G2 affine (slot-major). */
int bte_btx_prepare(const bte_p2_affine* material, size_t total_weight, size_t max_batch,
                    const uint64_t* party_offsets, size_t party_count, const uint32_t* accepted,
                    size_t n_accepted, size_t batch_size, bte_fp6* kernel_lines,
                    bte_p2_affine* opening_keys, void* stream);

/* decryption.rs:555-603 precompute_batch: ct1 = B G1 affine, valid = B HOST flags (an invalid
 * slot's ct1 is read as the identity, as ValidatedBatch stores it); beta_out = B Fp12 (device). */
int bte_btx_precompute(const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size,
                       const bte_fp6* kernel_lines, bte_fp12* beta_out, void* stream);

/* decryption.rs:607-695 open_batch: shares = T G1 Jacobian (committee order), ct2 and beta =
 * B Fp12, valid = B HOST flags; messages_out = B Fp12 (one in invalid slots, the reference's
 * None). */
int bte_btx_open(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* opening_keys,
                 const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid, size_t batch_size,
                 bte_fp12* messages_out, void* stream);

/* Fused forms over `count` batches of one committee and batch size (the north-star batching):
 * arrays hold count x B slots (count x T shares) back to back; one launch per stage serves all
 * batches. */
int bte_btx_precompute_batched(const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size,
                               size_t count, const bte_fp6* kernel_lines, bte_fp12* beta_out,
                               void* stream);
int bte_btx_open_batched(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* opening_keys,
                         const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid,
                         size_t batch_size, size_t count, bte_fp12* messages_out, void* stream);

/* One decryption request of the serving loop: `count` ciphertexts of a committee whose state was
 * prepared for batch size `batch_size` >= count (slots count..batch_size-1 are empty, the
 * reference's None).  Device: ct1 (count G1 affine), ct2 (count Fp12), kernel_lines and
 * opening_keys from bte_btx_prepare(batch_size), shares (n_accepted G1 Jacobian), messages_out
 * (count Fp12).  HOST: valid (count flags, NULL = all valid). */
typedef struct {
  const bte_p1_affine* ct1;
  const bte_fp12* ct2;
  const uint8_t* valid;
  size_t count;
  size_t batch_size;
  const bte_fp6* kernel_lines;
  const bte_p2_affine* opening_keys;
  const bte_fp6* opening_lines; /* NULL, or the keys' line tables (bte_g2_prepare_lines) */
  const bte_p1* shares;
  size_t n_accepted;
  bte_fp12* messages_out;
} bte_btx_request;

/* decryption.rs:555-603 precompute_batch followed by :607-695 open_batch for requests of
 * different committees and batch sizes at once (opening keys prepared when every request gives
 * opening_lines): every stage (segmented G1 FFT, prepared Miller
 * loops, easy part, segmented prefix IFFT, hard part, uneven multi-Miller loops, final
 * exponentiation, combine) is one launch over all requests. */
int bte_btx_decrypt_requests(const bte_btx_request* requests, size_t count, void* stream);

/* open_batch with the opening keys prepared once per committee and batch size:
 * key_lines = bte_g2_prepare_lines over the B x T opening keys (68 tables each, slot-major).
 * The same messages as bte_btx_open_batched; the prepared multi-Miller loop shares its
 * accumulator over chunks of pairs. */
int bte_btx_open_prepared_batched(const bte_p1* shares, size_t n_accepted, const bte_fp6* key_lines,
                                  const bte_fp12* ct2, const bte_fp12* beta, const uint8_t* valid,
                                  size_t batch_size, size_t count, bte_fp12* messages_out, void* stream);

/* Launch-only plan of the fused decrypt step (bte_btx_precompute_batched followed by
 * bte_btx_open_prepared_batched for `count` batches of one committee, with the same outputs).
 * Creation allocates every buffer, uploads the validity flags (count x B, HOST) and builds the
 * domain tables; bte_btx_plan_run then only enqueues kernels.  bte_btx_plan_launch_graph
 * captures one run into a HIP graph on its first call (stream must not be the null stream; the
 * graph stays bound to the buffers of that call, later calls must pass the same pointers) and
 * launches the graph.  No counterpart in the reference (its steps are rayon loops). */
typedef struct bte_btx_plan bte_btx_plan;
int bte_btx_plan_create(size_t batch_size, size_t count, size_t n_accepted, const bte_fp6* kernel_lines,
                        const bte_fp6* key_lines, const uint8_t* valid, bte_btx_plan** plan, void* stream);
int bte_btx_plan_run(bte_btx_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares, const bte_fp12* ct2,
                     bte_fp12* messages_out, void* stream);
int bte_btx_plan_launch_graph(bte_btx_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                              const bte_fp12* ct2, bte_fp12* messages_out, void* stream);
int bte_btx_plan_destroy(bte_btx_plan* plan);

/* weighted_btx/src/decryption.rs:236-272 verify_decryption_share for n_shares shares at once:
 * ok_out[j] (HOST) = 1 when prod_{valid i} e(ct1_i, vk[i N + party_index[j]]) e(-sigma_j, G2) is
 * one.  shares (T G1 Jacobian), verification_keys (L x N G2 affine, bte_btx_keygen's layout) and
 * ct1 (B G1 affine) are DEVICE; party_index (T) and valid (B) are HOST. */
int bte_btx_verify_shares(const bte_p1* shares, const uint32_t* party_index, size_t n_shares,
                          const bte_p2_affine* verification_keys, size_t n_parties, size_t max_batch,
                          const bte_p1_affine* ct1, const uint8_t* valid, size_t batch_size, uint8_t* ok_out,
                          void* stream);
/* decryption.rs:356-422 verify_decryption_shares_batched: the shares folded with caller-chosen
 * nonzero challenges (T canonical scalars, HOST; a zero or non-canonical challenge is
 * BTE_EINVAL): one G1 MSM, one batched indexed G2 MSM over the valid slots and one
 * multi-pairing; *result (HOST) = 1 when the folded check holds. */
int bte_btx_verify_shares_batched(const bte_p1* shares, const uint32_t* party_index, const uint8_t* challenges,
                                  size_t n_shares, const bte_p2_affine* verification_keys, size_t n_parties,
                                  size_t max_batch, const bte_p1_affine* ct1, const uint8_t* valid,
                                  size_t batch_size, int* result, void* stream);

/* weighted_btx/src/proof.rs:56-70 SchnorrProof::verify for B ciphertexts at once, after the host's
 * Fiat-Shamir hashing: ok_out[i] = 1 when s_i G == A_i + e_i C_i.  commitments A and ct1 C: B G1
 * affine (device); responses s and challenges e: B canonical scalars each (HOST); ok_out: B
 * bytes (HOST).  A non-canonical scalar is BTE_EINVAL. */
int bte_btx_verify_proofs(const bte_p1_affine* commitments, const bte_p1_affine* ct1, const uint8_t* responses,
                          const uint8_t* challenges, size_t batch_size, uint8_t* ok_out, void* stream);

/* weighted_btx/src/setup.rs:98-233 keygen_with_rng, with the dealer's random choices given:
 * poly = n_coeffs canonical scalars (the degree-t Shamir polynomial, t = n_coeffs - 1), q = N
 * nonzero canonical scalars, party_offsets = N+1 prefix sums of the weights (all HOST).
 * Outputs (device): material = negative_material (L blocks of W) then positive_material (L-1
 * blocks of W), G2 affine, as bte_btx_prepare reads it; verification_keys = L x N G2 affine
 * ([q_j^(i+1)]_2 at i N + j) or NULL. */
int bte_btx_keygen(const uint8_t* poly, size_t n_coeffs, const uint8_t* q, const uint64_t* party_offsets,
                   size_t n_parties, size_t max_batch, bte_p2_affine* material, bte_p2_affine* verification_keys,
                   void* stream);

/* weighted_pfe/src/setup.rs:149-354 keygen with the dealer's random choices given: poly = n_coeffs
 * canonical scalars (the Shamir polynomial, z = poly(0)); rho = N canonical party labels; alpha =
 * B canonical alpha points; beta = (beta_{-1}, beta_0), two canonical scalars; party_offsets = N+1
 * prefix sums of the weights (all HOST).  Outputs: material (device) = d1 (B blocks of W) then d2
 * (B blocks of W), G2 affine, as bte_pfe_prepare reads it; verification_keys (device, or NULL) =
 * B x N G2 affine, [g_{j,i}]_2 at i N + j; global_key (device, or NULL) = [z (beta_{-1} +
 * beta_0)]_2; fractions_out (HOST, or NULL) = N x B canonical g_{j,i} = 1 / (rho_j + alpha_i) at
 * j B + i.  A zero label rho_j + alpha_i, alpha_i or alpha_i + 1 is BTE_EINVAL (the reference
 * resamples rho; the caller does here). */
int bte_pfe_keygen(const uint8_t* poly, size_t n_coeffs, const uint8_t* rho, const uint8_t* alpha, const uint8_t* beta,
                   const uint64_t* party_offsets, size_t n_parties, size_t batch_size, bte_p2_affine* material,
                   bte_p2_affine* verification_keys, bte_p2_affine* global_key, uint8_t* fractions_out,
                   void* stream);

/* decryption.rs:205-233 partial_decrypt for N parties at once: sigma_j = sum_i q_j^i ct1_i,
 * i = 1..B.  q = N canonical scalars (HOST); sigma_out = N G1 Jacobian (device). */
int bte_btx_partial_decrypt(const bte_p1_affine* ct1, size_t batch_size, const uint8_t* q,
                            size_t n_parties, bte_p1* sigma_out, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * a16: weighted_pfe batch-decryption phases.  Key material (device) is one G2 affine array: the
 * reference's d1 (B blocks of W points) followed by its d2 (B blocks of W).  alpha holds the B
 * alpha points (canonical scalars, HOST).
 * ---------------------------------------------------------------------------------------- */

/* weighted_pfe/src/decryption.rs:421-537 prepare_decryption.  Outputs (device): d1_by_party
 * (B x T G2 affine, slot-major, coefficients scaled by 1/2B), total_lines and mask_lines (B line
 * tables each). */
int bte_pfe_prepare(const bte_p2_affine* material, size_t total_weight, size_t batch_size,
                    const uint64_t* party_offsets, size_t party_count, const uint32_t* accepted,
                    size_t n_accepted, const uint8_t* alpha, const bte_p2_affine* global_key,
                    bte_p2_affine* d1_by_party, bte_fp6* total_lines, bte_fp6* mask_lines,
                    void* stream);
/* decryption.rs:540-573 precompute_batch: first_cauchy (B G1 Jacobian) and pairing_cauchy
 * (B Fp12) from ct1 (B G1 affine); B must be a power of two. */
int bte_pfe_precompute(const bte_p1_affine* ct1, size_t batch_size, const bte_fp6* total_lines,
                       bte_p1* first_cauchy, bte_fp12* pairing_cauchy, void* stream);
/* decryption.rs:576-644 open_batch -> messages_out (B Fp12). */
int bte_pfe_open(const bte_p1* shares, size_t n_accepted, const bte_p2_affine* d1_by_party,
                 const bte_fp6* mask_lines, const bte_p1_affine* ct1, const bte_fp12* ct2,
                 const bte_p1* first_cauchy, const bte_fp12* pairing_cauchy, const uint8_t* alpha,
                 size_t batch_size, bte_fp12* messages_out, void* stream);
/* Fused forms over `count` batches of one committee (arrays count x B back to back, shares
 * count x T): precompute_batch per batch, and open_batch against d1_lines = bte_g2_prepare_lines
 * of d1_by_party (B x T tables, prepared once per committee and batch size). */
int bte_pfe_precompute_batched(const bte_p1_affine* ct1, size_t batch_size, size_t count,
                               const bte_fp6* total_lines, bte_p1* first_cauchy, bte_fp12* pairing_cauchy,
                               void* stream);
int bte_pfe_open_prepared_batched(const bte_p1* shares, size_t n_accepted, const bte_fp6* d1_lines,
                                  const bte_fp6* mask_lines, const bte_p1_affine* ct1, const bte_fp12* ct2,
                                  const bte_p1* first_cauchy, const bte_fp12* pairing_cauchy,
                                  const uint8_t* alpha, size_t batch_size, size_t count,
                                  bte_fp12* messages_out, void* stream);

/* Plan of the fused weighted_pfe step (bte_pfe_precompute_batched then
 * bte_pfe_open_prepared_batched for `count` batches of one committee, same outputs): creation
 * keeps the committee's line tables and the B alpha points (HOST, canonical) and allocates the
 * Cauchy intermediates; bte_pfe_plan_launch_graph captures one run into a HIP graph on its first
 * call (non-null stream; later calls must pass the same buffers) and replays it.  No counterpart
 * in the reference (its steps are rayon loops). */
typedef struct bte_pfe_plan bte_pfe_plan;
int bte_pfe_plan_create(size_t batch_size, size_t count, size_t n_accepted, const bte_fp6* total_lines,
                        const bte_fp6* d1_lines, const bte_fp6* mask_lines, const uint8_t* alpha,
                        bte_pfe_plan** plan, void* stream);
int bte_pfe_plan_run(bte_pfe_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares, const bte_fp12* ct2,
                     bte_fp12* messages_out, void* stream);
int bte_pfe_plan_launch_graph(bte_pfe_plan* plan, const bte_p1_affine* ct1, const bte_p1* shares,
                              const bte_fp12* ct2, bte_fp12* messages_out, void* stream);
int bte_pfe_plan_destroy(bte_pfe_plan* plan);

/* weighted_pfe/src/proof.rs:74-122 verify_batch after the host-side transcript hashing: 1 when
 * sum c_i s_i G - sum c_i A_i - sum c_i e_i C_i is the identity (one G1 MSM of 2B+1 terms).
 * commitments A and ct1 C: B G1 affine (device); responses s, challenges e, coefficients c: B
 * canonical scalars each (HOST); *result HOST. */
int bte_pfe_verify_proofs_batched(const bte_p1_affine* commitments, const bte_p1_affine* ct1,
                                  const uint8_t* responses, const uint8_t* challenges,
                                  const uint8_t* coefficients, size_t batch_size, int* result, void* stream);

/* decryption.rs:230-251 partial_decrypt for N parties: sigma_j = sum_i g_{j,i} ct1_i with the
 * N x B fractions g (canonical scalars, device). */
int bte_pfe_partial_decrypt(const bte_p1_affine* ct1, size_t batch_size, const uint8_t* fractions,
                            size_t n_parties, bte_p1* sigma_out, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * Fixed-base generation of test and key material: out[i] = k_i * G for the standard
 * generators (affine).  Used to build bases k_i * G for MSM inputs and CRS powers.
 * ---------------------------------------------------------------------------------------- */
int bte_g1_generator_mul(const uint8_t* scalars, size_t n, bte_p1_affine* out, void* stream);
int bte_g2_generator_mul(const uint8_t* scalars, size_t n, bte_p2_affine* out, void* stream);

/* ---------------------------------------------------------------------------------------- *
 * Measurement: per-phase HIP-event timing of the drivers (off by default) and a throughput
 * probe of the v_mad_u64_u32 + v_addc_co_u32 pair every field multiplication is built from.
 * No counterpart in the reference (measurement support).
 * ---------------------------------------------------------------------------------------- */
int bte_profile_enable(int on);
/* Operations of one unscaled prefix inverse cyclotomic transform of size n with output_count
 * outputs, counted from the twiddle plans the kernels use (HOST only): counts[0] cyclotomic
 * squarings, counts[1] Fp12 products, counts[2] Frobenius maps. */
int bte_cyclotomic_prefix_work(size_t n, size_t output_count, uint64_t* counts);
/* "phase=ms;phase=ms..." of the most recent profiled call. */
int bte_profile_read(char* buf, size_t len);
/* Measured 32x32->64 multiply-accumulates per second of the whole device. */
int bte_probe_mac_rate(double* macs_per_second, void* stream);

/* Launch-shape knobs (no counterpart in the reference: blst and rayon pick their own work
 * split; results never depend on these).  Set by the autotuner (weighted_bte_amd/autotune.py),
 * read by a driver when it sizes a launch; a plan keeps the values it was created with.
 *   BTE_TUNE_MSM_ACC_ROUNDS      full-occupancy rounds of bucket-accumulation threads, 1..16
 *                                (default 3): fewer rounds mean longer chunks per thread
 *   BTE_TUNE_CYC_THREE_LANE_MAX  cyclotomic stages with fewer butterflies than this run each
 *                                Fp12 on three lanes, 0..2^30 (default 32768)
 *   BTE_TUNE_MM_CHUNK_MAX        pairs sharing one multi-Miller accumulator, 1..64 (default 8)
 *   BTE_TUNE_MM_MIN_THREADS      fewest multi-Miller threads before chunks shrink, 1..2^24
 *                                (default 65536) */
enum {
  BTE_TUNE_MSM_ACC_ROUNDS = 1,
  BTE_TUNE_CYC_THREE_LANE_MAX = 2,
  BTE_TUNE_MM_CHUNK_MAX = 3,
  BTE_TUNE_MM_MIN_THREADS = 4
};
int bte_tune_set(int knob, int64_t value);
int bte_tune_get(int knob, int64_t* value);

/* Test support (no counterpart in the reference): the nth allocation from now of the library's
 * cached tables (FFT domain twiddles and plans) fails as if the device were out of memory (0
 * cancels), so a test can check that a failed build leaves nothing half-made behind. */
int bte_test_fail_alloc(int nth);

/* Workspace of `stream` (no counterpart in the reference: its calls allocate per call).  Every
 * entry point takes its device scratch from a grow-only buffer per (device, stream) and stages
 * host data through pinned buffers per stream, so after the first calls of a shape it neither
 * allocates nor waits (calls that return host results excepted).  This pre-sizes both: device
 * bytes of scratch and host bytes per staging buffer.  Graph capture of any entry point is
 * supported on a stream other than the null stream; captured work owns its own scratch. */
int bte_stream_reserve(void* stream, size_t device_bytes, size_t host_bytes);

/* ---------------------------------------------------------------------------------------- *
 * Device memory helpers for hosts without a HIP binding (the Rust FFI crate).
 * ---------------------------------------------------------------------------------------- */
int bte_device_alloc(void** ptr, size_t bytes);
int bte_device_free(void* ptr);
int bte_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int bte_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int bte_stream_synchronize(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BTE_HIP_H */
