/*
 * bg16 -- MI355X (gfx950) Groth16/BN254 proving backend: C ABI.
 *
 * This is the drop-in boundary for the reference prover's device-dispatched hot path
 * (bend2/prove.bend:24-49, the three `!` sites at :27, :37, :47): the Pippenger MSMs
 * of bend2/src/msm.bend and the QAP NTTs of bend2/src/ntt.bend, plus the resident-key
 * prover around them.  All entry points return an int status (0 = BG_OK, negative =
 * error class; bg_last_error() gives the message).  No exception or abort crosses the
 * ABI.  Buffers are caller-owned; inputs are read-only.
 *
 * Element encodings (flags):
 *   BG_FMT_WORDS16  the reference's Array<U32> layout (bend2/FORMAT.md:5-13): one field
 *                   element = 16 u32 words each holding a 16-bit limb, Montgomery R = 2^256
 *                   (MSM scalars: canonical).  G1 affine 32 words, G2 affine 64 words,
 *                   infinity = all zero.
 *   BG_FMT_PACKED   the same integers packed as 8 x u32 little-endian words (32 B per
 *                   element; w32[j] = w16[2j] | w16[2j+1] << 16).
 * Memory (flags):
 *   BG_MEM_HOST     pointers are host memory (copied in and out by the call)
 *   BG_MEM_DEVICE   pointers are device memory on the library's device (packed only)
 */
#ifndef BG16_H
#define BG16_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BG_OK = 0,
  BG_ERR_ARG = -1,      /* bad argument (size, flag, null pointer) */
  BG_ERR_DEVICE = -2,   /* HIP runtime error */
  BG_ERR_NOMEM = -3,    /* device or host allocation failed / over budget */
  BG_ERR_IO = -4,       /* file missing, truncated or malformed */
  BG_ERR_INTERNAL = -5
};

enum {
  BG_FMT_WORDS16 = 0,
  BG_FMT_PACKED = 1,
  BG_MEM_HOST = 0,
  BG_MEM_DEVICE = 2,
  /* bg_prove only: finish the MSMs and form the proof's combination and affine points on
   * the device (bend2/src/groth16.bend:281-308 there, no host round trip); needs a key
   * with fixed-base tables.  The default combines on the host while MSM H runs. */
  BG_COMBINE_DEVICE = 4
};

/* NTT directions: bend2/src/ntt.bend:431-442 (fft, ifft, coset_fft, coset_ifft) */
enum { BG_NTT_FFT = 0, BG_NTT_IFFT = 1, BG_NTT_COSET_FFT = 2, BG_NTT_COSET_IFFT = 3 };

/* ------------------------------------------------------------------ library */

/* Selects the HIP device (default 0) and creates the context; idempotent. */
int bg_init(int device);
/* Message of the last failing call on this process. */
const char* bg_last_error(void);
/* Enables per-kernel HIP-event timing of the hot-path kernels (0/1). */
int bg_set_timing(int on);
/* Fills names/avg_ms/counts for up to cap kernels timed since the last reset; returns
 * the number of entries.  names receives NUL-separated strings (cap * 48 bytes). */
int bg_timing_report(char* names, double* avg_ms, uint64_t* counts, int cap);
int bg_timing_reset(void);

/* ------------------------------------------------------------------ MSM
 * Replaces Msm.g1(points, scalars, n) / Msm.g2 (bend2/src/msm.bend:297, :529) and
 * Msm.g1_with / g2_with (:293, :525): out = sum_{i<n} scalars[i] * points[i].
 * points: n affine Montgomery points; scalars: n canonical Fr (any value < 2^256 is
 * read as an integer and reduced mod r, which leaves the result unchanged for points of
 * order r).  out: one affine point (same encoding as points; infinity = zeros).  The
 * reference returns Jacobian; the affine point is unique, and the reference's caller
 * normalises it (bend2/src/groth16.bend:281-293).
 * window_bits: 0 = tuned default, else the signed-digit window width c (2..22). */
int bg_msm_g1(const uint32_t* points, const uint32_t* scalars, uint64_t n, uint32_t* out,
              int flags, int window_bits);
int bg_msm_g2(const uint32_t* points, const uint32_t* scalars, uint64_t n, uint32_t* out,
              int flags, int window_bits);
/* count MSMs of one group in one launch sequence (Msm.g1 / Msm.g2 applied to count
 * independent inputs of uneven sizes): every stage -- digit decomposition, bucket sort,
 * accumulation, bucket reduction -- runs once over the whole batch.  points[i],
 * scalars[i]: ns[i] packed points / canonical scalars (host, or device with
 * BG_MEM_DEVICE); outs: count packed affine results (host, or device). */
int bg_msm_batch(int group, uint32_t count, const uint32_t* const* points, const uint32_t* const* scalars,
                 const uint64_t* ns, uint32_t* outs, int flags, int window_bits);
/* The MSM schedule chosen for n points of group (1 G1, 2 G2): info8 = window bits c,
 * windows W, buckets per window, sorted entries (W n), entries per accumulation thread,
 * accumulation threads, level-1 groups, level-2 outputs.  No counterpart in the
 * reference (its window choice is Msm.pick_c, bend2/src/msm.bend:40-44); used to state
 * the algorithmic work of a launch. */
int bg_msm_plan_info(uint64_t n, int group, int window_bits, uint64_t* info8);

/* ------------------------------------------------------------------ tuned parameters
 * Replaces the reference's fixed heuristics Msm.pick_c / Msm.pick_fd
 * (bend2/src/msm.bend:37-44) and Ntt.fd (bend2/src/ntt.bend:382-385): a table measured on
 * the device (`bench.py --mode tune`) of MSM window width, accumulation slice target and
 * reduction shape per (group, windowed / fixed-base, log2 points in half steps), and of NTT tile
 * and radix per log2 n.  The planners read it first and fall back to the cost model.
 * bg_tune_load(path): (re)loads the table (NULL: $BG16_TUNE_FILE, else tune_gfx950.txt
 * next to the library; a missing file is an empty table, a malformed one BG_ERR_IO).
 * bg_tune_lookup(kind, group, fixed, size, values4): kind 0 = MSM (size = total points;
 * values4 = c, L, m, K), kind 1 = NTT (size = log2 n; values4 = tile log2, radix); 0 = an
 * entry exists, 1 = none (values4 zero), negative = error.  0 in a value: planner's choice. */
int bg_tune_load(const char* path);
int bg_tune_lookup(int kind, int group, int fixed, uint64_t size, int* values4);

/* ------------------------------------------------------------------ NTT
 * Replaces Ntt.fft / ifft / coset_fft / coset_ifft (bend2/src/ntt.bend:431-442), in
 * place on 2^log2n Montgomery Fr elements, natural order in and out.  ifft scales by
 * n^-1; the coset generator is 5. */
int bg_ntt(uint32_t* data, int log2n, int direction, int flags);

/* Replaces Ntt.qap_h(a, b, c, log2) (bend2/src/ntt.bend:459):
 * h = coset_ifft((coset_fft(ifft a) * coset_fft(ifft b) - coset_fft(ifft c)) / (g^d - 1)),
 * d = 2^log2n Montgomery coefficients, natural order. */
int bg_qap_h(const uint32_t* a, const uint32_t* b, const uint32_t* c, int log2n, uint32_t* h,
             int flags);

/* qap_h without allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  a, b, c, h: packed Montgomery device buffers of 2^log2n
 * elements (a, b, c read-only); h receives what bg_qap_h returns for the same inputs.
 * All working memory -- copies of a, b, c and the domain's tables -- lives in one static
 * device array of about 30 MiB (domains up to 2^17: log2n 1..17, larger gives
 * BG_ERR_ARG), held by each device the library runs on until the process ends.  The
 * call makes no hipMalloc / hipFree and no other allocation; it blocks until h is
 * written.  Calls share the array: calls that overlap on one device race. */
int bg_qap_h_noalloc(const uint32_t* a, const uint32_t* b, const uint32_t* c, int log2n, uint32_t* h);

/* ------------------------------------------------------------------ resident-key prover
 * Reproduces bend2/prove.bend's Pr.run (:129-162) on the device: evaluation vectors
 * (groth16.bend:162-208), qap_h, the five MSMs in the reference's order and slots
 * (prove.bend:64-116) and the final combination (groth16.bend:298-308). */
typedef struct bg_pk bg_pk;

/* Loads a FORMAT.md key directory (header.bin, r1cs_{a,b,c}.bin, pk_*.bin) into HBM.
 * Every file is validated on the host (bend2/tools/verify_bin.py:24-64): sizes, limb
 * words below 2^16, CSR offsets and columns, coefficients below r, and every G1 / G2
 * point (pk_g1, pk_g2 and the five queries) below q and on its curve (or all-zero
 * infinity); a failing file gives BG_ERR_IO. */
int bg_pk_load(const char* dir, bg_pk** out);
/* Frees a loaded key.  A key must outlive every proof that uses it: while a bg_prove
 * call on it runs or an engine request of it is queued or running (until the proof is
 * done, i.e. until bg_engine_wait can return it), this returns BG_ERR_ARG and the key
 * stays loaded. */
int bg_pk_free(bg_pk* pk);
/* header: num_constraints, num_instance, num_witness, domain_log2, n_a, n_b, n_h, n_l */
int bg_pk_header(const bg_pk* pk, uint32_t* header8);
/* Window widths of the key's fixed-base tables, chosen at load (tuned table, else cost
 * model): c3 = G1 batch (A, B1, L), H, B2; zeros when the key keeps plain arrays. */
int bg_pk_windows(const bg_pk* pk, int* c3);
/* Device bytes held by a loaded key (the resident queries, matrices and NTT tables;
 * per-proof work arrays live in the prover slots). */
int bg_pk_device_bytes(const bg_pk* pk, uint64_t* bytes);

/* One proof.  witness: num_instance + num_witness Fr (Montgomery, FORMAT.md witness.bin
 * order), host memory in either encoding, or with BG_MEM_DEVICE a packed device buffer
 * already resident in HBM; rs: r then s (Montgomery, FORMAT.md rs.bin, host).  A host
 * witness element or r / s at or above r gives BG_ERR_ARG (a device witness is the
 * caller's to keep reduced).
 * proof_out (host): A (G1), B (G2), C (G1) affine Montgomery (the proof_ref.bin layout),
 * in the flags' encoding.
 * t_ms (optional, 9 doubles): device time of upload, evals, qap, msm_a, msm_b1, msm_b2,
 * msm_h, msm_l (the T phases of bend2/prove.bend:154-162; the MSMs run concurrently on
 * separate streams, so they overlap), then the call's wall time. */
int bg_prove(bg_pk* pk, const uint32_t* witness, const uint32_t* rs, uint32_t* proof_out,
             int flags, double* t_ms);

/* Single proofs (bg_prove, and engine requests proved alone) of domains up to 2^15
 * replay recorded HIP graphs of their device work after the upload, one linear graph per
 * stream: the second proof of a key on a slot records them, later ones launch them
 * (BG16_GRAPH=0 turns this off, =1 on for every size).  No counterpart in the reference
 * (one process per proof); the proof is the same either way.  info3: graph sets recorded,
 * proofs launched from graphs, recordings that failed (and ran eagerly), counted over
 * the process. */
int bg_graph_info(uint64_t* info3);

/* ------------------------------------------------------------------ proving engine
 * Continuous batching of proofs on one device.  The reference proves one instance per
 * process (bend2/prove.bend:129-162, run by bend2/scripts/prove.sh); the engine is the
 * batched / served form of the same Pr.run: each request is proved exactly as bg_prove
 * proves it alone, on one of `slots` worker slots (four streams each), so proofs of
 * different keys and sizes overlap on the device.  Requests are taken in arrival order
 * when their work arrays fit hbm_budget_bytes (0: 90 % of the device memory free at
 * creation; resident keys are not counted).  Thread-safe. */
typedef struct bg_engine bg_engine;
int bg_engine_create(int slots, uint64_t hbm_budget_bytes, bg_engine** out);
/* Queues one proof; witness / rs / flags as bg_prove (host inputs are copied; a device
 * witness must stay valid until bg_engine_wait returns).  Fails with BG_ERR_NOMEM when
 * one proof of this key exceeds the budget. */
int bg_engine_submit(bg_engine* e, bg_pk* pk, const uint32_t* witness, const uint32_t* rs, int flags,
                     uint64_t* ticket);
/* Blocks until the proof is done; proof_out as bg_prove.  t_ms (optional, 3 doubles):
 * time queued, time proving, submit-to-done latency. */
int bg_engine_wait(bg_engine* e, uint64_t ticket, uint32_t* proof_out, double* t_ms);
/* info: slots, device bytes reserved by the slots, budget, proofs completed. */
int bg_engine_info(bg_engine* e, uint64_t* info4);
/* Finishes every queued proof, then stops the workers. */
int bg_engine_destroy(bg_engine* e);
/* count proofs (key, witness, rs each) through a temporary engine with `slots` slots
 * (<= 0: 4); proofs[i] as bg_prove's proof_out. */
int bg_prove_batch(int count, bg_pk* const* pks, const uint32_t* const* witnesses, const uint32_t* const* rs,
                   uint32_t* const* proofs, int flags, int slots);

/* Proof as the reference prints it (bend2/FORMAT.md:33): "A x y\nB x0 x1 y0 y1\nC x y\n"
 * with canonical big-endian 0x hex; proof in BG_FMT_PACKED.  Returns bytes written. */
int bg_proof_format(const uint32_t* proof_packed, char* buf, int cap);

/* ------------------------------------------------------------------ key generation
 * Trapdoor Groth16 setup for SquareChain(2^log2) (ref/src/circuit.rs:37-73,
 * ref/src/gen.rs:65-164 restated with this library's own seeded generator): writes the
 * FORMAT.md files under dir plus trapdoor.bin (tau, alpha, beta, gamma, delta as
 * Montgomery Fr) for closed-form checks.  threads <= 0 uses all host cores. */
int bg_setup_square_chain(const char* dir, int log2, uint64_t seed, int threads);

/* k_i * G for canonical packed scalars (G the G1 or G2 generator), affine packed out
 * (host).  The closed form of an MSM over bases k_i G; no counterpart in the reference
 * (its harness builds such bases with arkworks, ref/src/bench.rs). */
int bg_fixed_base_mul(int group, const uint32_t* scalars, uint64_t n, uint32_t* out_affine, int threads);
/* n synthetic bases k_i G with k_i uniform in [0, r) from the seeded generator (the
 * primitive MSM inputs); k_out (optional) receives the k_i canonical packed. */
int bg_gen_bases(int group, uint64_t seed, uint64_t n, uint32_t* out_affine, uint32_t* k_out, int threads);

/* ------------------------------------------------------------------ verification (host)
 * Replaces ref/src/verify.rs:8-27 (arkworks prepare_verifying_key + verify_proof).
 * vk: the words of vk.bin (bg_setup_square_chain): alpha G1, beta G2, gamma G2, delta G2,
 * a u32 count, gamma_abc G1[count]; proof as bg_prove writes it; inputs: n_inputs
 * Montgomery Fr (count must be n_inputs + 1).  flags: BG_FMT_*.  Returns 1 when
 * e(A, B) = e(alpha, beta) e(IC(x), gamma) e(C, delta), 0 when not, BG_ERR_ARG when a
 * point is off its curve, B / a vk G2 point is outside the order-r subgroup, or a
 * coordinate is not below q (the reference's json.rs check_point errors on those). */
int bg_verify(const uint32_t* vk, uint64_t vk_words, const uint32_t* proof, const uint32_t* inputs,
              uint32_t n_inputs, int flags);
/* prod_{i<n} e(P_i, Q_i) == 1 (optimal ate on BN254): 1 yes, 0 no, BG_ERR_ARG for a
 * bad point.  g1s: n affine G1, g2s: n affine G2 (Montgomery, infinity = all zero). */
int bg_pairing_check(const uint32_t* g1s, const uint32_t* g2s, uint32_t n, int flags);

/* ------------------------------------------------------------------ diagnostics
 * Known-answer and element-wise kernels used by the parity tests (host buffers, packed). */

/* bend2/bench/field_mul.bend:9-31: xor of limb 0 of 2^N hash-generated Fr products. */
int bg_kat_field_mul_xor(int log2_count, uint32_t* xor_out);
/* bend2/bench/g1_add.bend: k reference-formula mixed adds acc <- acc + G from G; returns
 * limb 0 (16 bits) of the Jacobian X in Montgomery form. */
int bg_kat_g1_add_chain(uint64_t k, uint32_t* x16_out);

/* field: 0 Fr, 1 Fq, 2 Fq2.  op: 0 add, 1 sub, 2 mul, 3 neg, 4 inv, 5 sqr, 6 to_mont,
 * 7 from_mont.  n elements of a (and b for binary ops) -> out. */
int bg_field_op(int field, int op, const uint32_t* a, const uint32_t* b, uint32_t* out,
                uint64_t n);
/* group: 1 G1, 2 G2.  op: 0 add (p + q), 1 double (2p), 2 neg (-p), 3 add through XYZZ
 * full add (from_affine(p) + from_affine(q)).  Affine in, affine out. */
int bg_curve_op(int group, int op, const uint32_t* p, const uint32_t* q, uint32_t* out,
                uint64_t n);
/* Throughput probes: kind 0 = v_mad_u64_u32 issue rate (ops/s), 1 = Fq Montgomery
 * products/s, 2 = G1 XYZZ mixed adds/s.  result[0] = rate, result[1] = checksum. */
int bg_probe_rate(int kind, uint64_t iters, double* result);
/* Launch probe (no counterpart in the reference): microseconds per repetition of `kernels`
 * short kernels (`spin` dependent steps each) launched 0 eagerly on one stream, 1 from one
 * linear graph, 2 eagerly forked over four streams, 3 from one graph that forks and joins
 * them, 4 as 3 with host-visible event nodes, 5 from one linear graph per stream after an
 * eager fork: the choice behind the prover's recorded launch sequences. */
int bg_probe_launch(int kernels, int spin, int reps, double* us6);

#ifdef __cplusplus
}
#endif
#endif
