/*
 * CPU ORACLE -- test infrastructure only, never part of the product path.
 *
 * A plain-C restatement of the reference prover's arithmetic for the Groth16/BN254
 * hot path (field, curve, Pippenger MSM, radix-2 NTT, qap_h, proof combination).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker (or the timed CPU baseline), never as the thing measured
 * or shipped.
 *
 * Reference files followed (paths relative to the reference repository root):
 *   bend2/FORMAT.md:5-33            encodings (Montgomery R = 2^256, 16-bit limbs, infinity = zeros)
 *   bend2/gen/gen_field.py:47-155   16x16-bit CIOS Montgomery product, add/sub, Fermat inverse
 *   bend2/src/fq2.bend:30-50        Fq2 Karatsuba mul, complex square, inverse
 *   bend2/gen/gen_curve.py:11-231   Jacobian dbl-2009-l, add-2007-bl, madd-2007-bl + special cases
 *   bend2/src/msm.bend:33-58,94-246 Pippenger: unsigned c-bit windows, running-sum reduce, window combine
 *   bend2/gen/ref_qap.py:27-72      radix-2 DIT NTT, ifft, coset variants, qap_h
 *   bend2/src/groth16.bend:162-308  R1CS evaluation rows, canonical scalars, proof combination (combine.c9)
 *   ref/README.md:24-33             proof formula (ark-groth16 0.5 LibsnarkReduction)
 *
 * Field elements are 4 x u64 little-endian (byte-identical to 8 x u32 little-endian).
 * Every function that is checked by a known-answer test or closed form is listed in
 * the oracle's Python header (oracle/oracle.py) together with its pin.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

typedef struct {
    u64 p[4];
    u64 n0;      /* -p^-1 mod 2^64 */
    u64 one[4];  /* R mod p */
    u64 r2[4];   /* R^2 mod p */
} field_t;

/* BN254 scalar field r and base field q (bend2/FORMAT.md:6). n0 / one / r2 are filled
 * in by or_init() from the modulus alone. */
static field_t FR = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}, 0, {0}, {0}};
static field_t FQ = {{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}, 0, {0}, {0}};
static int g_inited = 0;

static inline int geq4(const u64 *a, const u64 *b) {
    for (int i = 3; i >= 0; i--) {
        if (a[i] > b[i]) return 1;
        if (a[i] < b[i]) return 0;
    }
    return 1;
}
static inline u64 sub4(u64 *r, const u64 *a, const u64 *b) {
    u64 br = 0;
    for (int i = 0; i < 4; i++) {
        u128 t = (u128)a[i] - b[i] - br;
        r[i] = (u64)t;
        br = (u64)(t >> 64) & 1;
    }
    return br;
}
static inline u64 add4(u64 *r, const u64 *a, const u64 *b) {
    u64 c = 0;
    for (int i = 0; i < 4; i++) {
        u128 t = (u128)a[i] + b[i] + c;
        r[i] = (u64)t;
        c = (u64)(t >> 64);
    }
    return c;
}

/* ---------------------------------------------------------------- prime field */

static inline void f_add(const field_t *F, u64 *r, const u64 *a, const u64 *b) {
    u64 t[4];
    u64 c = add4(t, a, b);
    if (c || geq4(t, F->p)) sub4(t, t, F->p);
    memcpy(r, t, 32);
}
static inline void f_sub(const field_t *F, u64 *r, const u64 *a, const u64 *b) {
    u64 t[4];
    if (sub4(t, a, b)) add4(t, t, F->p);
    memcpy(r, t, 32);
}
static inline void f_neg(const field_t *F, u64 *r, const u64 *a) {
    u64 z[4] = {0, 0, 0, 0};
    f_sub(F, r, z, a);
}
static inline int f_is_zero(const u64 *a) { return (a[0] | a[1] | a[2] | a[3]) == 0; }
static inline int f_eq(const u64 *a, const u64 *b) { return memcmp(a, b, 32) == 0; }

/* Montgomery product a*b*2^-256 mod p (CIOS on 64-bit words).  The product is unique
 * for inputs < p, so it equals the reference's 16x16-bit CIOS bit for bit; or_fe16_mul
 * below restates that limb schedule and the tests compare the two. */
static void f_mul(const field_t *F, u64 *r, const u64 *a, const u64 *b) {
    u64 t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c = (u128)a[j] * b[i] + t[j] + (u64)(c >> 64);
            t[j] = (u64)c;
        }
        c = (u128)t[4] + (u64)(c >> 64);
        t[4] = (u64)c;
        t[5] = (u64)(c >> 64);
        u64 m = t[0] * F->n0;
        c = (u128)m * F->p[0] + t[0];
        for (int j = 1; j < 4; j++) {
            c = (u128)m * F->p[j] + t[j] + (u64)(c >> 64);
            t[j - 1] = (u64)c;
        }
        c = (u128)t[4] + (u64)(c >> 64);
        t[3] = (u64)c;
        t[4] = t[5] + (u64)(c >> 64);
    }
    if (t[4] || geq4(t, F->p)) sub4(t, t, F->p);
    memcpy(r, t, 32);
}
static inline void f_sqr(const field_t *F, u64 *r, const u64 *a) { f_mul(F, r, a, a); }

/* a^e for a 256-bit exponent e (little-endian words), left to right */
static void f_pow(const field_t *F, u64 *r, const u64 *a, const u64 *e) {
    u64 acc[4];
    memcpy(acc, F->one, 32);
    for (int i = 255; i >= 0; i--) {
        f_sqr(F, acc, acc);
        if ((e[i >> 6] >> (i & 63)) & 1) f_mul(F, acc, acc, a);
    }
    memcpy(r, acc, 32);
}
/* Fermat inverse a^(p-2); inv(0) = 0 (gen_field.py:137-155) */
static void f_inv(const field_t *F, u64 *r, const u64 *a) {
    u64 e[4], two[4] = {2, 0, 0, 0};
    sub4(e, F->p, two);
    f_pow(F, r, a, e);
}
static void f_to_mont(const field_t *F, u64 *r, const u64 *a) { f_mul(F, r, a, F->r2); }
static void f_from_mont(const field_t *F, u64 *r, const u64 *a) {
    u64 one[4] = {1, 0, 0, 0};
    f_mul(F, r, a, one);
}

static void field_setup(field_t *F) {
    /* n0 = -p^-1 mod 2^64 by Newton iteration */
    u64 inv = 1;
    for (int i = 0; i < 7; i++) inv *= 2 - F->p[0] * inv;
    F->n0 = (u64)0 - inv;
    /* one = 2^256 mod p: double 1 256 times with reduction */
    u64 x[4] = {1, 0, 0, 0};
    for (int i = 0; i < 256; i++) f_add(F, x, x, x);
    memcpy(F->one, x, 32);
    /* r2 = 2^512 mod p: keep doubling */
    for (int i = 0; i < 256; i++) f_add(F, x, x, x);
    memcpy(F->r2, x, 32);
}

/* ----------------------------------------------- reference-shaped 16-bit CIOS
 * Restates gen_field.py:105-135 (gen_mul) and :47-60 (cond_sub_p) on 16 x 16-bit limbs
 * held in u32 with wrapping arithmetic, exactly as emitted into bend2/src/fr.bend. */
static void fe16_mul(const u32 *pl, u32 n0, u32 *out, const u32 *a, const u32 *b) {
    const int L = 16;
    u32 t[18];
    memset(t, 0, sizeof t);
    for (int i = 0; i < L; i++) {
        u32 c = 0;
        for (int j = 0; j < L; j++) {
            u32 s = t[j] + a[j] * b[i] + c;
            t[j] = s & 0xFFFF;
            c = s >> 16;
        }
        u32 s = t[L] + c;
        t[L] = s & 0xFFFF;
        t[L + 1] = s >> 16;
        u32 m = (t[0] * n0) & 0xFFFF;
        s = t[0] + m * pl[0];
        c = s >> 16;
        for (int j = 1; j < L; j++) {
            s = t[j] + m * pl[j] + c;
            t[j - 1] = s & 0xFFFF;
            c = s >> 16;
        }
        s = t[L] + c;
        t[L - 1] = s & 0xFFFF;
        c = s >> 16;
        t[L] = t[L + 1] + c;
        t[L + 1] = 0;
    }
    /* branch-free conditional subtraction of p */
    u32 d[16], br = 0;
    for (int j = 0; j < L; j++) {
        u32 s = t[j] + 65536 - pl[j] - br;
        d[j] = s & 0xFFFF;
        br = 1 - (s >> 16);
    }
    u32 ge = t[L] | (1 - br);
    u32 keep = ge - 1;
    for (int j = 0; j < L; j++) out[j] = (t[j] & keep) | (d[j] & ~keep);
}

/* ---------------------------------------------------------------- Fq2 = Fq[u]/(u^2+1) */
typedef struct { u64 c0[4], c1[4]; } fq2_t;

static inline void fq2_add(fq2_t *r, const fq2_t *a, const fq2_t *b) {
    f_add(&FQ, r->c0, a->c0, b->c0);
    f_add(&FQ, r->c1, a->c1, b->c1);
}
static inline void fq2_sub(fq2_t *r, const fq2_t *a, const fq2_t *b) {
    f_sub(&FQ, r->c0, a->c0, b->c0);
    f_sub(&FQ, r->c1, a->c1, b->c1);
}
static inline void fq2_neg(fq2_t *r, const fq2_t *a) {
    f_neg(&FQ, r->c0, a->c0);
    f_neg(&FQ, r->c1, a->c1);
}
/* Karatsuba, fq2.bend:30-37 */
static void fq2_mul(fq2_t *r, const fq2_t *a, const fq2_t *b) {
    u64 t0[4], t1[4], t2[4], sa[4], sb[4];
    f_mul(&FQ, t0, a->c0, b->c0);
    f_mul(&FQ, t1, a->c1, b->c1);
    f_add(&FQ, sa, a->c0, a->c1);
    f_add(&FQ, sb, b->c0, b->c1);
    f_mul(&FQ, t2, sa, sb);
    f_sub(&FQ, r->c0, t0, t1);
    f_sub(&FQ, t2, t2, t0);
    f_sub(&FQ, r->c1, t2, t1);
}
/* fq2.bend:39-43 */
static void fq2_sqr(fq2_t *r, const fq2_t *a) {
    u64 s[4], d[4], a2[4], c0[4], c1[4];
    f_add(&FQ, s, a->c0, a->c1);
    f_sub(&FQ, d, a->c0, a->c1);
    f_add(&FQ, a2, a->c0, a->c0);
    f_mul(&FQ, c0, s, d);
    f_mul(&FQ, c1, a2, a->c1);
    memcpy(r->c0, c0, 32);
    memcpy(r->c1, c1, 32);
}
/* fq2.bend:45-50 */
static void fq2_inv(fq2_t *r, const fq2_t *a) {
    u64 t0[4], t1[4], ni[4];
    f_sqr(&FQ, t0, a->c0);
    f_sqr(&FQ, t1, a->c1);
    f_add(&FQ, t0, t0, t1);
    f_inv(&FQ, ni, t0);
    u64 c0[4], c1[4];
    f_mul(&FQ, c0, a->c0, ni);
    f_mul(&FQ, c1, a->c1, ni);
    f_neg(&FQ, c1, c1);
    memcpy(r->c0, c0, 32);
    memcpy(r->c1, c1, 32);
}
static inline int fq2_is_zero(const fq2_t *a) { return f_is_zero(a->c0) && f_is_zero(a->c1); }

/* ----------------------------------------------------------- generic curve ops
 * The same Jacobian formulas serve G1 (over Fq) and G2 (over Fq2) as in gen_curve.py;
 * they are written once as a macro template over the coordinate field. */

typedef struct { u64 x[4], y[4]; } g1a_t;
typedef struct { u64 x[4], y[4], z[4]; } g1j_t;
typedef struct { fq2_t x[1], y[1]; } g2a_t;
typedef struct { fq2_t x[1], y[1], z[1]; } g2j_t;

#define G1_ADD(r, a, b) f_add(&FQ, (r), (a), (b))
#define G1_SUB(r, a, b) f_sub(&FQ, (r), (a), (b))
#define G1_MUL(r, a, b) f_mul(&FQ, (r), (a), (b))
#define G1_SQR(r, a) f_sqr(&FQ, (r), (a))
#define G1_ISZ(a) f_is_zero(a)
#define G1_SETONE(a) memcpy((a), FQ.one, 32)
#define G1_SETZERO(a) memset((a), 0, 32)
#define G1_EV(v) u64 v[4]

#define G2_ADD(r, a, b) fq2_add((r), (a), (b))
#define G2_SUB(r, a, b) fq2_sub((r), (a), (b))
#define G2_MUL(r, a, b) fq2_mul((r), (a), (b))
#define G2_SQR(r, a) fq2_sqr((r), (a))
#define G2_ISZ(a) fq2_is_zero(a)
#define G2_SETONE(a) do { memcpy((a)->c0, FQ.one, 32); memset((a)->c1, 0, 32); } while (0)
#define G2_SETZERO(a) memset((a), 0, sizeof(fq2_t))
#define G2_EV(v) fq2_t v[1]

/* Template body: G = g1|g2, P = G1|G2 (macro prefix), J = jacobian type, A = affine type */
#define CURVE_IMPL(G, P, J, A)                                                              \
static void G##_infinity(J *r) { P##_SETZERO(r->x); P##_SETONE(r->y); P##_SETZERO(r->z); } \
static int G##_aff_is_inf(const A *p) { return P##_ISZ(p->x) && P##_ISZ(p->y); }           \
static int G##_is_inf(const J *p) { return P##_ISZ(p->z); }                                \
static void G##_from_affine(J *r, const A *p) {                                            \
    if (G##_aff_is_inf(p)) { G##_infinity(r); return; }                                    \
    memcpy(r->x, p->x, sizeof(p->x)); memcpy(r->y, p->y, sizeof(p->y)); P##_SETONE(r->z);  \
}                                                                                          \
/* dbl-2009-l (gen_curve.py:86-102); z = 0 stays z3 = 0 */                                 \
static void G##_double(J *r, const J *p) {                                                 \
    P##_EV(a); P##_EV(b); P##_EV(c); P##_EV(d); P##_EV(e); P##_EV(f); P##_EV(t);           \
    P##_EV(x3); P##_EV(y3); P##_EV(z3); P##_EV(c4);                                        \
    P##_SQR(a, p->x); P##_SQR(b, p->y); P##_SQR(c, b);                                     \
    P##_ADD(t, p->x, b); P##_SQR(t, t); P##_SUB(t, t, a); P##_SUB(t, t, c);                \
    P##_ADD(d, t, t);                                                                      \
    P##_ADD(e, a, a); P##_ADD(e, e, a);                                                    \
    P##_SQR(f, e);                                                                         \
    P##_ADD(t, d, d); P##_SUB(x3, f, t);                                                   \
    P##_ADD(c4, c, c); P##_ADD(c4, c4, c4); P##_ADD(c4, c4, c4);                           \
    P##_SUB(t, d, x3); P##_MUL(t, e, t); P##_SUB(y3, t, c4);                               \
    P##_ADD(t, p->y, p->y); P##_MUL(z3, t, p->z);                                          \
    memcpy(r->x, x3, sizeof(r->x)); memcpy(r->y, y3, sizeof(r->y));                        \
    memcpy(r->z, z3, sizeof(r->z));                                                        \
}                                                                                          \
/* add-2007-bl with the special-case resolution of gen_curve.py:104-145 */                 \
static void G##_add(J *r, const J *p, const J *q) {                                        \
    if (G##_is_inf(p)) { *r = *q; return; }                                                \
    if (G##_is_inf(q)) { *r = *p; return; }                                                \
    P##_EV(z1z1); P##_EV(z2z2); P##_EV(u1); P##_EV(u2); P##_EV(s1); P##_EV(s2);            \
    P##_EV(h); P##_EV(i); P##_EV(j); P##_EV(rr); P##_EV(rv); P##_EV(v); P##_EV(t);         \
    P##_EV(x3); P##_EV(y3); P##_EV(z3);                                                    \
    P##_SQR(z1z1, p->z); P##_SQR(z2z2, q->z);                                              \
    P##_MUL(u1, p->x, z2z2); P##_MUL(u2, q->x, z1z1);                                      \
    P##_MUL(s1, p->y, q->z); P##_MUL(s1, s1, z2z2);                                        \
    P##_MUL(s2, q->y, p->z); P##_MUL(s2, s2, z1z1);                                        \
    P##_SUB(h, u2, u1);                                                                    \
    P##_SUB(rr, s2, s1);                                                                   \
    if (P##_ISZ(h)) {                                                                      \
        if (P##_ISZ(rr)) G##_double(r, p); else G##_infinity(r);                           \
        return;                                                                            \
    }                                                                                      \
    P##_ADD(t, h, h); P##_SQR(i, t); P##_MUL(j, h, i);                                     \
    P##_ADD(rv, rr, rr); P##_MUL(v, u1, i);                                                \
    P##_SQR(x3, rv); P##_SUB(x3, x3, j); P##_ADD(t, v, v); P##_SUB(x3, x3, t);             \
    P##_MUL(s1, s1, j); P##_ADD(s1, s1, s1);                                               \
    P##_SUB(t, v, x3); P##_MUL(y3, rv, t); P##_SUB(y3, y3, s1);                            \
    P##_ADD(t, p->z, q->z); P##_SQR(t, t); P##_SUB(t, t, z1z1); P##_SUB(t, t, z2z2);       \
    P##_MUL(z3, t, h);                                                                     \
    memcpy(r->x, x3, sizeof(r->x)); memcpy(r->y, y3, sizeof(r->y));                        \
    memcpy(r->z, z3, sizeof(r->z));                                                        \
}                                                                                          \
/* madd-2007-bl with the special cases of gen_curve.py:147-184 */                          \
static void G##_add_mixed(J *r, const J *p, const A *q) {                                  \
    if (G##_is_inf(p)) { G##_from_affine(r, q); return; }                                  \
    if (G##_aff_is_inf(q)) { *r = *p; return; }                                            \
    P##_EV(z1z1); P##_EV(u2); P##_EV(s2); P##_EV(h); P##_EV(hh); P##_EV(i); P##_EV(j);     \
    P##_EV(rr); P##_EV(rv); P##_EV(v); P##_EV(t); P##_EV(x3); P##_EV(y3); P##_EV(z3);      \
    P##_SQR(z1z1, p->z);                                                                   \
    P##_MUL(u2, q->x, z1z1);                                                               \
    P##_MUL(s2, q->y, p->z); P##_MUL(s2, s2, z1z1);                                        \
    P##_SUB(h, u2, p->x);                                                                  \
    P##_SUB(rr, s2, p->y);                                                                 \
    if (P##_ISZ(h)) {                                                                      \
        if (P##_ISZ(rr)) G##_double(r, p); else G##_infinity(r);                           \
        return;                                                                            \
    }                                                                                      \
    P##_SQR(hh, h); P##_ADD(i, hh, hh); P##_ADD(i, i, i);                                  \
    P##_MUL(j, h, i);                                                                      \
    P##_ADD(rv, rr, rr);                                                                   \
    P##_MUL(v, p->x, i);                                                                   \
    P##_SQR(x3, rv); P##_SUB(x3, x3, j); P##_ADD(t, v, v); P##_SUB(x3, x3, t);             \
    P##_MUL(t, p->y, j); P##_ADD(t, t, t);                                                 \
    P##_SUB(y3, v, x3); P##_MUL(y3, rv, y3); P##_SUB(y3, y3, t);                           \
    P##_ADD(t, p->z, h); P##_SQR(t, t); P##_SUB(t, t, z1z1); P##_SUB(z3, t, hh);           \
    memcpy(r->x, x3, sizeof(r->x)); memcpy(r->y, y3, sizeof(r->y));                        \
    memcpy(r->z, z3, sizeof(r->z));                                                        \
}                                                                                          \
static void G##_neg(J *r, const J *p);                                                     \
/* one inversion; infinity -> (0, 0) */                                                    \
static void G##_to_affine(A *r, const J *p);                                               \
/* double-and-add over the 256 bits of a canonical scalar (gen_curve.py:186-205) */        \
static void G##_mul_scalar(J *r, const J *p, const u64 *k) {                               \
    J acc; G##_infinity(&acc);                                                             \
    for (int b = 255; b >= 0; b--) {                                                       \
        G##_double(&acc, &acc);                                                            \
        if ((k[b >> 6] >> (b & 63)) & 1) G##_add(&acc, &acc, p);                           \
    }                                                                                      \
    *r = acc;                                                                              \
}

CURVE_IMPL(g1, G1, g1j_t, g1a_t)
CURVE_IMPL(g2, G2, g2j_t, g2a_t)

static void g1_neg(g1j_t *r, const g1j_t *p) { *r = *p; f_neg(&FQ, r->y, p->y); }
static void g2_neg(g2j_t *r, const g2j_t *p) { *r = *p; fq2_neg(r->y, p->y); }
static void g1_to_affine(g1a_t *r, const g1j_t *p) {
    if (g1_is_inf(p)) { memset(r, 0, sizeof *r); return; }
    u64 zi[4], zi2[4], zi3[4];
    f_inv(&FQ, zi, p->z);
    f_sqr(&FQ, zi2, zi);
    f_mul(&FQ, zi3, zi, zi2);
    f_mul(&FQ, r->x, p->x, zi2);
    f_mul(&FQ, r->y, p->y, zi3);
}
static void g2_to_affine(g2a_t *r, const g2j_t *p) {
    if (g2_is_inf(p)) { memset(r, 0, sizeof *r); return; }
    fq2_t zi, zi2, zi3;
    fq2_inv(&zi, p->z);
    fq2_sqr(&zi2, &zi);
    fq2_mul(&zi3, &zi, &zi2);
    fq2_mul(r->x, p->x, &zi2);
    fq2_mul(r->y, p->y, &zi3);
}

/* ---------------------------------------------------------------- Pippenger MSM
 * msm.bend:33-58 (windows, digit), :94-155 (bucket accumulation), :191-235 (running-sum
 * bucket reduction and window combine).  Windows are independent, so they run in
 * parallel over OpenMP threads; the final affine point is unique, so the combine order
 * does not affect the output. */

static int ceil_log2(u64 n) { int d = 0; while (((u64)1 << d) < n) d++; return d; }
static int pick_c(u64 n) { return ceil_log2(n) < 20 ? 8 : 10; }     /* msm.bend:40-41 */
static int n_windows(int c) { return 253 / c + 1; }                  /* msm.bend:33-35 */

/* bits [lo, lo + min(c, 256 - lo)) of a canonical scalar (msm.bend:52-58) */
static u32 digit(const u64 *s, int lo, int c) {
    if (lo >= 256) return 0;
    int w = c < 256 - lo ? c : 256 - lo;
    int k = lo >> 6, r = lo & 63;
    u128 v = s[k] >> r;
    if (r && k + 1 < 4) v |= (u128)s[k + 1] << (64 - r);
    return (u32)(v & (((u64)1 << w) - 1));
}

#define MSM_IMPL(G, J, A)                                                                  \
static void G##_msm(J *out, const A *pts, const u64 *sc, u64 n, int c) {                   \
    if (c <= 0) c = pick_c(n);                                                             \
    int W = n_windows(c);                                                                  \
    u64 m = (u64)1 << c;                                                                   \
    J *wsum = (J *)malloc(sizeof(J) * W);                                                  \
    _Pragma("omp parallel for schedule(dynamic, 1)")                                       \
    for (int w = 0; w < W; w++) {                                                          \
        J *b = (J *)malloc(sizeof(J) * m);                                                 \
        for (u64 u = 0; u < m; u++) G##_infinity(&b[u]);                                   \
        for (u64 i = 0; i < n; i++) {                                                      \
            u32 d = digit(sc + 4 * i, w * c, c);                                           \
            if (d) G##_add_mixed(&b[d], &b[d], &pts[i]);                                   \
        }                                                                                  \
        J run, acc; G##_infinity(&run); G##_infinity(&acc);                                \
        for (u64 u = m - 1; u >= 1; u--) {                                                 \
            G##_add(&run, &run, &b[u]);                                                    \
            G##_add(&acc, &acc, &run);                                                     \
        }                                                                                  \
        wsum[w] = acc;                                                                     \
        free(b);                                                                           \
    }                                                                                      \
    J acc; G##_infinity(&acc);                                                             \
    for (int w = W - 1; w >= 0; w--) {                                                     \
        for (int k = 0; k < c; k++) G##_double(&acc, &acc);                                \
        G##_add(&acc, &acc, &wsum[w]);                                                     \
    }                                                                                      \
    free(wsum);                                                                            \
    *out = acc;                                                                            \
}

MSM_IMPL(g1, g1j_t, g1a_t)
MSM_IMPL(g2, g2j_t, g2a_t)

/* ---------------------------------------------------------------- NTT over Fr
 * ref_qap.py:27-72 restated: natural-order in/out radix-2 DIT with bit reversal.  The
 * DFT is unique, so any correct transform gives the reference's ntt.bend output. */

static u64 ROOT28[4];   /* 2-adic root of unity of order 2^28, Montgomery */
static u64 GEN5[4];     /* coset generator 5, Montgomery */

static void fr_omega(u64 *w, int log2n, int inverse) {
    u64 x[4];
    memcpy(x, ROOT28, 32);
    if (inverse) f_inv(&FR, x, x);
    for (int i = 0; i < 28 - log2n; i++) f_sqr(&FR, x, x);
    memcpy(w, x, 32);
}

static void ntt_core(u64 *a, int log2n, const u64 *w) {
    u64 n = (u64)1 << log2n;
    for (u64 i = 1, j = 0; i < n; i++) {
        u64 bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j |= bit;
        if (i < j) {
            u64 t[4];
            memcpy(t, a + 4 * i, 32);
            memcpy(a + 4 * i, a + 4 * j, 32);
            memcpy(a + 4 * j, t, 32);
        }
    }
    /* twiddle table for the largest stage: tw[k] = w^k, k < n/2 */
    u64 *tw = (u64 *)malloc(32 * (n / 2 + 1));
    memcpy(tw, FR.one, 32);
    for (u64 k = 1; k < n / 2; k++) f_mul(&FR, tw + 4 * k, tw + 4 * (k - 1), w);
    for (u64 m = 1; m < n; m <<= 1) {
        u64 stride = n / (2 * m);
        #pragma omp parallel for schedule(static) if (n >= 4096)
        for (u64 k = 0; k < n; k += 2 * m) {
            for (u64 j = 0; j < m; j++) {
                u64 *u = a + 4 * (k + j), *v = a + 4 * (k + j + m);
                u64 t[4], x[4];
                f_mul(&FR, t, v, tw + 4 * (j * stride));
                memcpy(x, u, 32);
                f_add(&FR, u, x, t);
                f_sub(&FR, v, x, t);
            }
        }
    }
    free(tw);
}

static void scale_geometric(u64 *a, u64 n, const u64 *start, const u64 *ratio) {
    u64 v[4];
    memcpy(v, start, 32);
    for (u64 i = 0; i < n; i++) {
        f_mul(&FR, a + 4 * i, a + 4 * i, v);
        f_mul(&FR, v, v, ratio);
    }
}

/* dir: 0 fft, 1 ifft, 2 coset_fft, 3 coset_ifft (ntt.bend:431-442) */
static void ntt_dir(u64 *a, int log2n, int dir) {
    u64 n = (u64)1 << log2n, w[4];
    if (dir == 2) scale_geometric(a, n, FR.one, GEN5);
    fr_omega(w, log2n, dir == 1 || dir == 3);
    ntt_core(a, log2n, w);
    if (dir == 1 || dir == 3) {
        u64 nn[4] = {n, 0, 0, 0}, ninv[4];
        f_to_mont(&FR, nn, nn);
        f_inv(&FR, ninv, nn);
        if (dir == 1) scale_geometric(a, n, ninv, FR.one);
        else {
            u64 gi[4];
            f_inv(&FR, gi, GEN5);
            scale_geometric(a, n, ninv, gi);
        }
    }
}

/* h = coset_ifft((coset_fft(ifft a) * coset_fft(ifft b) - coset_fft(ifft c)) / (g^d - 1))
 * (ref_qap.py:68-72; ntt.bend:444-459).  a, b, c are overwritten. */
static void qap_h(u64 *h, u64 *a, u64 *b, u64 *c, int log2n) {
    u64 n = (u64)1 << log2n;
    u64 *v[3] = {a, b, c};
    for (int k = 0; k < 3; k++) {
        ntt_dir(v[k], log2n, 1);
        ntt_dir(v[k], log2n, 2);
    }
    u64 gd[4], zinv[4], e[4] = {n, 0, 0, 0};
    f_pow(&FR, gd, GEN5, e);
    f_sub(&FR, gd, gd, FR.one);
    f_inv(&FR, zinv, gd);
    for (u64 i = 0; i < n; i++) {
        u64 t[4];
        f_mul(&FR, t, a + 4 * i, b + 4 * i);
        f_sub(&FR, t, t, c + 4 * i);
        f_mul(&FR, h + 4 * i, t, zinv);
    }
    ntt_dir(h, log2n, 3);
}

/* ================================================================ exported API */

static void or_init_once(void) {
    if (g_inited) return;
    field_setup(&FR);
    field_setup(&FQ);
    /* ROOT = 19103219067921713944291392827692070036145651957329286315305642004821462161904
     * (bend2/gen/gen_ntt.py:9), little-endian 64-bit words */
    u64 root[4] = {0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL, 0x2a3c09f0a58a7e85ULL};
    f_to_mont(&FR, ROOT28, root);
    u64 five[4] = {5, 0, 0, 0};
    f_to_mont(&FR, GEN5, five);
    g_inited = 1;
}

static const field_t *fsel(int fid) { or_init_once(); return fid ? &FQ : &FR; }

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* fid: 0 = Fr, 1 = Fq.  Element pointers are 4 x u64 (Montgomery unless noted). */
void or_f_consts(int fid, u64 *p, u64 *one, u64 *r2, u64 *n0) {
    const field_t *F = fsel(fid);
    memcpy(p, F->p, 32); memcpy(one, F->one, 32); memcpy(r2, F->r2, 32); *n0 = F->n0;
}
void or_f_add(int fid, u64 *r, const u64 *a, const u64 *b) { f_add(fsel(fid), r, a, b); }
void or_f_sub(int fid, u64 *r, const u64 *a, const u64 *b) { f_sub(fsel(fid), r, a, b); }
void or_f_neg(int fid, u64 *r, const u64 *a) { f_neg(fsel(fid), r, a); }
void or_f_mul(int fid, u64 *r, const u64 *a, const u64 *b) { f_mul(fsel(fid), r, a, b); }
void or_f_inv(int fid, u64 *r, const u64 *a) { f_inv(fsel(fid), r, a); }
void or_f_to_mont(int fid, u64 *r, const u64 *a) { f_to_mont(fsel(fid), r, a); }
void or_f_from_mont(int fid, u64 *r, const u64 *a) { f_from_mont(fsel(fid), r, a); }
/* elementwise Montgomery product over n elements */
void or_f_mul_vec(int fid, u64 *r, const u64 *a, const u64 *b, u64 n) {
    const field_t *F = fsel(fid);
    #pragma omp parallel for schedule(static) if (n >= 4096)
    for (u64 i = 0; i < n; i++) f_mul(F, r + 4 * i, a + 4 * i, b + 4 * i);
}
void or_f_from_mont_vec(int fid, u64 *r, const u64 *a, u64 n) {
    const field_t *F = fsel(fid);
    #pragma omp parallel for schedule(static) if (n >= 4096)
    for (u64 i = 0; i < n; i++) f_from_mont(F, r + 4 * i, a + 4 * i);
}
void or_f_to_mont_vec(int fid, u64 *r, const u64 *a, u64 n) {
    const field_t *F = fsel(fid);
    #pragma omp parallel for schedule(static) if (n >= 4096)
    for (u64 i = 0; i < n; i++) f_to_mont(F, r + 4 * i, a + 4 * i);
}

/* reference-shaped 16-bit-limb product (gen_field.py:105-135); a, b, out are 16 words */
void or_fe16_mul(int fid, u32 *out, const u32 *a, const u32 *b) {
    const field_t *F = fsel(fid);
    u32 pl[16];
    for (int j = 0; j < 16; j++) pl[j] = (u32)(F->p[j / 4] >> (16 * (j % 4))) & 0xFFFF;
    u32 n0 = (u32)(F->n0 & 0xFFFF);
    fe16_mul(pl, n0, out, a, b);
}

/* bend2/bench/field_mul.bend:9-31: hash-generated Montgomery operands, xor of limb 0 of
 * 2^N products.  use16 = 1 runs the 16-bit restatement, 0 the 64-bit CIOS. */
static u32 kat_h(u32 i) { return ((i * 2654435761u) ^ (i >> 15)) * 2246822519u; }
void or_kat_rnd(u32 s, u32 *limbs16) {
    for (int k = 0; k < 16; k++) {
        u32 off = k == 0 ? 0 : 40503u * k + 1;
        limbs16[k] = kat_h(s + off) & (k == 15 ? 12287u : 65535u);
    }
}
u32 or_field_mul_xor(int N, int use16) {
    const field_t *F = fsel(0);
    u32 pl[16];
    for (int j = 0; j < 16; j++) pl[j] = (u32)(F->p[j / 4] >> (16 * (j % 4))) & 0xFFFF;
    u32 n0 = (u32)(F->n0 & 0xFFFF);
    u64 total = (u64)1 << N;
    u32 x = 0;
    #pragma omp parallel for reduction(^ : x) schedule(static)
    for (u64 i = 0; i < total; i++) {
        u32 a16[16], b16[16];
        or_kat_rnd((u32)(2 * i), a16);
        or_kat_rnd((u32)(2 * i + 1), b16);
        if (use16) {
            u32 o[16];
            fe16_mul(pl, n0, o, a16, b16);
            x ^= o[0];
        } else {
            u64 a[4] = {0}, b[4] = {0}, o[4];
            for (int j = 0; j < 16; j++) {
                a[j / 4] |= (u64)a16[j] << (16 * (j % 4));
                b[j / 4] |= (u64)b16[j] << (16 * (j % 4));
            }
            f_mul(F, o, a, b);
            x ^= (u32)(o[0] & 0xFFFF);
        }
    }
    return x;
}

/* bend2/bench/g1_add.bend: chain acc <- acc + G (k mixed adds starting from G, the first
 * one on the P == Q doubling path); returns limb 0 (16 bits) of the Jacobian X. */
u32 or_g1_add_chain_x16(u64 k) {
    or_init_once();
    g1a_t g;
    memcpy(g.x, FQ.one, 32);
    f_add(&FQ, g.y, FQ.one, FQ.one);
    g1j_t acc;
    g1_from_affine(&acc, &g);
    for (u64 i = 0; i < k; i++) g1_add_mixed(&acc, &acc, &g);
    return (u32)(acc.x[0] & 0xFFFF);
}
/* same chain, returning the full Jacobian point (12 words) */
void or_g1_add_chain(u64 k, u64 *jac_out) {
    or_init_once();
    g1a_t g;
    memcpy(g.x, FQ.one, 32);
    f_add(&FQ, g.y, FQ.one, FQ.one);
    g1j_t acc;
    g1_from_affine(&acc, &g);
    for (u64 i = 0; i < k; i++) g1_add_mixed(&acc, &acc, &g);
    memcpy(jac_out, &acc, sizeof acc);
}

/* Fq2 ops on 8-word elements (c0 then c1) */
void or_fq2_mul(u64 *r, const u64 *a, const u64 *b) { or_init_once(); fq2_mul((fq2_t *)r, (const fq2_t *)a, (const fq2_t *)b); }
void or_fq2_sqr(u64 *r, const u64 *a) { or_init_once(); fq2_sqr((fq2_t *)r, (const fq2_t *)a); }
void or_fq2_inv(u64 *r, const u64 *a) { or_init_once(); fq2_inv((fq2_t *)r, (const fq2_t *)a); }

/* curve ops: Jacobian in/out (G1 12 words, G2 24 words), affine (G1 8, G2 16) */
void or_g1_add(u64 *r, const u64 *p, const u64 *q) { or_init_once(); g1_add((g1j_t *)r, (const g1j_t *)p, (const g1j_t *)q); }
void or_g1_add_mixed(u64 *r, const u64 *p, const u64 *q) { or_init_once(); g1_add_mixed((g1j_t *)r, (const g1j_t *)p, (const g1a_t *)q); }
void or_g1_double(u64 *r, const u64 *p) { or_init_once(); g1_double((g1j_t *)r, (const g1j_t *)p); }
void or_g1_to_affine(u64 *r, const u64 *p) { or_init_once(); g1_to_affine((g1a_t *)r, (const g1j_t *)p); }
void or_g1_from_affine(u64 *r, const u64 *p) { or_init_once(); g1_from_affine((g1j_t *)r, (const g1a_t *)p); }
void or_g2_add(u64 *r, const u64 *p, const u64 *q) { or_init_once(); g2_add((g2j_t *)r, (const g2j_t *)p, (const g2j_t *)q); }
void or_g2_add_mixed(u64 *r, const u64 *p, const u64 *q) { or_init_once(); g2_add_mixed((g2j_t *)r, (const g2j_t *)p, (const g2a_t *)q); }
void or_g2_double(u64 *r, const u64 *p) { or_init_once(); g2_double((g2j_t *)r, (const g2j_t *)p); }
void or_g2_to_affine(u64 *r, const u64 *p) { or_init_once(); g2_to_affine((g2a_t *)r, (const g2j_t *)p); }
void or_g2_from_affine(u64 *r, const u64 *p) { or_init_once(); g2_from_affine((g2j_t *)r, (const g2a_t *)p); }

/* affine k*P for a canonical scalar k (gen_curve.py:186-205) */
void or_g1_mul_aff(u64 *r, const u64 *p_aff, const u64 *k) {
    or_init_once();
    g1j_t p, o;
    g1_from_affine(&p, (const g1a_t *)p_aff);
    g1_mul_scalar(&o, &p, k);
    g1_to_affine((g1a_t *)r, &o);
}
void or_g2_mul_aff(u64 *r, const u64 *p_aff, const u64 *k) {
    or_init_once();
    g2j_t p, o;
    g2_from_affine(&p, (const g2a_t *)p_aff);
    g2_mul_scalar(&o, &p, k);
    g2_to_affine((g2a_t *)r, &o);
}

/* MSM: affine points (Montgomery), canonical scalars (4 words each) -> affine result.
 * c <= 0 picks the reference's window (msm.bend:40-41). */
void or_msm_g1(u64 *out_aff, const u64 *pts, const u64 *sc, u64 n, int c) {
    or_init_once();
    g1j_t r;
    g1_msm(&r, (const g1a_t *)pts, sc, n, c);
    g1_to_affine((g1a_t *)out_aff, &r);
}
void or_msm_g2(u64 *out_aff, const u64 *pts, const u64 *sc, u64 n, int c) {
    or_init_once();
    g2j_t r;
    g2_msm(&r, (const g2a_t *)pts, sc, n, c);
    g2_to_affine((g2a_t *)out_aff, &r);
}

/* NTT family on Montgomery Fr (in place): dir 0 fft, 1 ifft, 2 coset_fft, 3 coset_ifft */
void or_ntt(u64 *a, int log2n, int dir) { or_init_once(); ntt_dir(a, log2n, dir); }
void or_omega(u64 *w, int log2n, int inverse) { or_init_once(); fr_omega(w, log2n, inverse); }
void or_root28(u64 *w) { or_init_once(); memcpy(w, ROOT28, 32); }
/* qap_h over d = 2^log2n (a, b, c are clobbered) */
void or_qap_h(u64 *h, u64 *a, u64 *b, u64 *c, int log2n) { or_init_once(); qap_h(h, a, b, c, log2n); }

/* ------------------------------------------------------------ full prove
 * groth16.bend:162-208 (evals with the libsnark input-consistency rows), :210-219 and
 * :263-270 (canonical scalars, zl = z[2..]), prove.bend:52-116 (MSM order and slots),
 * groth16.bend:298-308 (combine.c9).
 *
 * csr_*: row offsets (nc+1 u32), cols (nnz u32), coeffs (nnz x 4 u64 Montgomery).
 * pk arrays are affine Montgomery.  z, r, s Montgomery.  Output: A (8 words),
 * B (16 words), C (8 words) affine Montgomery.  t_ms (optional, 7 doubles): evals,
 * qap, msm_a, msm_b1, msm_b2, msm_h, msm_l wall times. */
#include <time.h>
static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static void eval_rows(u64 *ev, const u32 *off, const u32 *col, const u64 *coef, u32 nc, const u64 *z) {
    #pragma omp parallel for schedule(static) if (nc >= 4096)
    for (u32 i = 0; i < nc; i++) {
        u64 acc[4] = {0, 0, 0, 0};
        for (u32 e = off[i]; e < off[i + 1]; e++) {
            u64 t[4];
            f_mul(&FR, t, coef + 4 * (u64)e, z + 4 * (u64)col[e]);
            f_add(&FR, acc, acc, t);
        }
        memcpy(ev + 4 * (u64)i, acc, 32);
    }
}

/* one matrix's evaluation rows: ev[i] = sum_e coef[e] z[col[e]] for rows i < nc (Montgomery) */
void or_eval_rows(u64 *ev, const u32 *off, const u32 *col, const u64 *coef, u32 nc, const u64 *z) {
    or_init_once();
    eval_rows(ev, off, col, coef, nc, z);
}

int or_prove(const u32 *hdr,
             const u64 *z,
             const u32 *a_off, const u32 *a_col, const u64 *a_coef,
             const u32 *b_off, const u32 *b_col, const u64 *b_coef,
             const u32 *c_off, const u32 *c_col, const u64 *c_coef,
             const u64 *pk_g1, const u64 *pk_g2,
             const u64 *pa, const u64 *pb1, const u64 *pb2, const u64 *ph, const u64 *pl,
             const u64 *r, const u64 *s,
             u64 *A_out, u64 *B_out, u64 *C_out, double *t_ms) {
    or_init_once();
    u32 nc = hdr[0], ni = hdr[1], nw = hdr[2], lg = hdr[3];
    u32 na = hdr[4], nb = hdr[5], nh = hdr[6], nl = hdr[7];
    u64 d = (u64)1 << lg;
    if ((u64)nc + ni > d || na > ni + nw || nb > ni + nw || nh > d || nl + 2 > ni + nw) return -1;
    double t0 = now_ms();
    u64 *ea = (u64 *)calloc(d, 32), *eb = (u64 *)calloc(d, 32), *ec = (u64 *)calloc(d, 32);
    u64 *h = (u64 *)calloc(d, 32);
    eval_rows(ea, a_off, a_col, a_coef, nc, z);
    eval_rows(eb, b_off, b_col, b_coef, nc, z);
    eval_rows(ec, c_off, c_col, c_coef, nc, z);
    for (u32 i = 0; i < ni; i++) memcpy(ea + 4 * (u64)(nc + i), z + 4 * (u64)i, 32);
    u64 nz = (u64)ni + nw;
    u64 *zc = (u64 *)malloc(32 * nz);
    for (u64 i = 0; i < nz; i++) f_from_mont(&FR, zc + 4 * i, z + 4 * i);
    double t1 = now_ms();
    qap_h(h, ea, eb, ec, lg);
    for (u64 i = 0; i < d; i++) f_from_mont(&FR, h + 4 * i, h + 4 * i);
    double t2 = now_ms();
    g1j_t mA, mB1, mH, mL;
    g2j_t mB2;
    g1_msm(&mA, (const g1a_t *)pa, zc, na, 0);
    double t3 = now_ms();
    g1_msm(&mB1, (const g1a_t *)pb1, zc, nb, 0);
    double t4 = now_ms();
    g2_msm(&mB2, (const g2a_t *)pb2, zc, nb, 0);
    double t5 = now_ms();
    g1_msm(&mH, (const g1a_t *)ph, h, nh, 0);
    double t6 = now_ms();
    g1_msm(&mL, (const g1a_t *)pl, zc + 8, nl, 0);
    double t7 = now_ms();

    /* slots as prove.bend accumulates them: to_affine(add_mixed(m, slot)) */
    g1a_t aA, aB1, aLH, t;
    g2a_t aB2;
    g1_to_affine(&aA, &mA);
    g1_to_affine(&aB1, &mB1);
    g2_to_affine(&aB2, &mB2);
    g1_to_affine(&t, &mH);
    {
        g1j_t x;
        g1_add_mixed(&x, &mL, &t);
        g1_to_affine(&aLH, &x);
    }
    /* combine.c9 (groth16.bend:298-308) */
    const g1a_t *alpha = (const g1a_t *)pk_g1, *beta = (const g1a_t *)pk_g1 + 1, *delta = (const g1a_t *)pk_g1 + 2;
    const g2a_t *beta2 = (const g2a_t *)pk_g2, *delta2 = (const g2a_t *)pk_g2 + 1;
    u64 rc[4], sc_[4], rs[4], rsc[4];
    f_from_mont(&FR, rc, r);
    f_from_mont(&FR, sc_, s);
    f_mul(&FR, rs, r, s);
    f_from_mont(&FR, rsc, rs);
    g1j_t dj, x, y, A, B1, C;
    g2j_t x2, y2, B2;
    g1_from_affine(&dj, delta);
    g1_from_affine(&x, &aA);
    g1_mul_scalar(&y, &dj, rc);
    g1_add(&x, &x, &y);
    g1_add_mixed(&A, &x, alpha);
    g1_from_affine(&x, &aB1);
    g1_mul_scalar(&y, &dj, sc_);
    g1_add(&x, &x, &y);
    g1_add_mixed(&B1, &x, beta);
    g2_from_affine(&x2, &aB2);
    g2_from_affine(&y2, delta2);
    g2_mul_scalar(&y2, &y2, sc_);
    g2_add(&x2, &x2, &y2);
    g2_add_mixed(&B2, &x2, beta2);
    g1j_t sA, rB, rsd;
    g1_mul_scalar(&sA, &A, sc_);
    g1_mul_scalar(&rB, &B1, rc);
    g1_mul_scalar(&rsd, &dj, rsc);
    g1_neg(&rsd, &rsd);
    g1_add(&C, &sA, &rB);
    g1_add(&C, &C, &rsd);
    g1_from_affine(&x, &aLH);
    g1_add(&C, &C, &x);
    g1_to_affine((g1a_t *)A_out, &A);
    g2_to_affine((g2a_t *)B_out, &B2);
    g1_to_affine((g1a_t *)C_out, &C);
    if (t_ms) {
        t_ms[0] = t1 - t0; t_ms[1] = t2 - t1; t_ms[2] = t3 - t2; t_ms[3] = t4 - t3;
        t_ms[4] = t5 - t4; t_ms[5] = t6 - t5; t_ms[6] = t7 - t6;
    }
    free(ea); free(eb); free(ec); free(h); free(zc);
    return 0;
}

/* ------------------------------------------------------------ fixed-base batches
 * out[i] = k_i * G (affine, Montgomery) for the standard generators, via 8-bit windowed
 * tables T[w][d] = d * 2^(8w) * G and one batched inversion per chunk (Montgomery's
 * trick).  Used to build large MSM inputs with known discrete logs (P_i = k_i G), so
 * that MSM = (sum a_i k_i mod r) G can be checked with a single scalar multiplication. */
static void g1_gen(g1a_t *g) { memcpy(g->x, FQ.one, 32); f_add(&FQ, g->y, FQ.one, FQ.one); }
static void g2_gen(g2a_t *g) {
    /* standard BN254 G2 generator (x = x0 + x1 u, y = y0 + y1 u), canonical words */
    static const u64 X0[4] = {0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL};
    static const u64 X1[4] = {0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL};
    static const u64 Y0[4] = {0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL};
    static const u64 Y1[4] = {0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL};
    f_to_mont(&FQ, g->x->c0, X0); f_to_mont(&FQ, g->x->c1, X1);
    f_to_mont(&FQ, g->y->c0, Y0); f_to_mont(&FQ, g->y->c1, Y1);
}

#define FIXED_BASE_IMPL(G, J, A, E, EINV, EMUL, ESQR, GENF)                                 \
void or_##G##_gen_mul_batch(u64 *out, const u64 *ks, u64 n) {                              \
    or_init_once();                                                                        \
    A gen; GENF(&gen);                                                                     \
    J *tab = (J *)malloc(sizeof(J) * 32 * 256);                                            \
    J base; G##_from_affine(&base, &gen);                                                  \
    for (int w = 0; w < 32; w++) {                                                         \
        G##_infinity(&tab[w * 256]);                                                       \
        for (int d = 1; d < 256; d++) G##_add(&tab[w * 256 + d], &tab[w * 256 + d - 1], &base); \
        for (int k = 0; k < 8; k++) G##_double(&base, &base);                              \
    }                                                                                      \
    A *out_a = (A *)out;                                                                   \
    _Pragma("omp parallel")                                                                \
    {                                                                                      \
        const u64 CH = 256;                                                                \
        J *acc = (J *)malloc(sizeof(J) * CH);                                              \
        E *pre = (E *)malloc(sizeof(E) * CH);                                              \
        _Pragma("omp for schedule(dynamic, 1)")                                            \
        for (u64 c0 = 0; c0 < n; c0 += CH) {                                               \
            u64 m = n - c0 < CH ? n - c0 : CH;                                             \
            for (u64 i = 0; i < m; i++) {                                                  \
                const u64 *k = ks + 4 * (c0 + i);                                          \
                G##_infinity(&acc[i]);                                                     \
                for (int w = 0; w < 32; w++) {                                             \
                    unsigned d = (unsigned)((k[w / 8] >> (8 * (w % 8))) & 0xFF);           \
                    if (d) G##_add(&acc[i], &acc[i], &tab[w * 256 + d]);                   \
                }                                                                          \
            }                                                                              \
            /* batched inversion of the z coordinates (zero z -> infinity) */              \
            E run; memset(&run, 0, sizeof run); E one_; memset(&one_, 0, sizeof one_);      \
            memcpy(&one_, FQ.one, 32); run = one_;                                         \
            for (u64 i = 0; i < m; i++) {                                                  \
                pre[i] = run;                                                              \
                if (!G##_is_inf(&acc[i])) EMUL(&run, &run, (const E *)acc[i].z);           \
            }                                                                              \
            E inv; EINV(&inv, &run);                                                       \
            for (u64 i = m; i-- > 0;) {                                                    \
                A *o = &out_a[c0 + i];                                                     \
                if (G##_is_inf(&acc[i])) { memset(o, 0, sizeof *o); continue; }            \
                E zi, zi2, zi3;                                                            \
                EMUL(&zi, &inv, &pre[i]);                                                  \
                EMUL(&inv, &inv, (const E *)acc[i].z);                                     \
                ESQR(&zi2, &zi); EMUL(&zi3, &zi2, &zi);                                    \
                EMUL((E *)o->x, (const E *)acc[i].x, &zi2);                                \
                EMUL((E *)o->y, (const E *)acc[i].y, &zi3);                                \
            }                                                                              \
        }                                                                                  \
        free(acc); free(pre);                                                              \
    }                                                                                      \
    free(tab);                                                                             \
}

typedef struct { u64 v[4]; } fqe_t;
static void fqe_mul(fqe_t *r, const fqe_t *a, const fqe_t *b) { f_mul(&FQ, r->v, a->v, b->v); }
static void fqe_sqr(fqe_t *r, const fqe_t *a) { f_mul(&FQ, r->v, a->v, a->v); }
static void fqe_inv(fqe_t *r, const fqe_t *a) { f_inv(&FQ, r->v, a->v); }
static void fq2e_mul(fq2_t *r, const fq2_t *a, const fq2_t *b) { fq2_mul(r, a, b); }
static void fq2e_sqr(fq2_t *r, const fq2_t *a) { fq2_sqr(r, a); }
static void fq2e_inv(fq2_t *r, const fq2_t *a) { fq2_inv(r, a); }

FIXED_BASE_IMPL(g1, g1j_t, g1a_t, fqe_t, fqe_inv, fqe_mul, fqe_sqr, g1_gen)
FIXED_BASE_IMPL(g2, g2j_t, g2a_t, fq2_t, fq2e_inv, fq2e_mul, fq2e_sqr, g2_gen)

/* sum_i a_i * k_i mod r for canonical a_i, k_i (4 words each) -> canonical (4 words) */
void or_fr_dot(u64 *out, const u64 *a, const u64 *k, u64 n) {
    or_init_once();
    u64 acc[4] = {0, 0, 0, 0};
    #pragma omp parallel
    {
        u64 loc[4] = {0, 0, 0, 0};
        #pragma omp for schedule(static)
        for (u64 i = 0; i < n; i++) {
            u64 x[4], y[4], t[4];
            /* a, k canonical: (a R)(k R) R^-1 = a k R, so convert one operand */
            f_to_mont(&FR, x, a + 4 * i);
            memcpy(y, k + 4 * i, 32);
            if (geq4(y, FR.p)) { /* reduce scalars >= r (at most 5 subtractions) */
                while (geq4(y, FR.p)) sub4(y, y, FR.p);
            }
            if (geq4(x, FR.p)) sub4(x, x, FR.p);
            f_mul(&FR, t, x, y);   /* = a k (canonical) */
            f_add(&FR, loc, loc, t);
        }
        #pragma omp critical
        f_add(&FR, acc, acc, loc);
    }
    memcpy(out, acc, 32);
}

/* Horner evaluation of a Montgomery-coefficient polynomial at Montgomery x */
void or_poly_eval(u64 *out, const u64 *coeffs, u64 n, const u64 *x) {
    or_init_once();
    u64 acc[4] = {0, 0, 0, 0};
    for (u64 i = n; i-- > 0;) {
        f_mul(&FR, acc, acc, x);
        f_add(&FR, acc, acc, coeffs + 4 * i);
    }
    memcpy(out, acc, 32);
}
