/* The CPU oracle (oracle/bn254_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: its known-answer checks (the reference's field_mul xor and
 * g1_add chain values), MSMs against the closed form, NTT round trips, qap_h, and the full
 * prove of a FORMAT.md key read here, equal to the proof the oracle made in-process.
 *
 * usage: oracle_checks <key dir (SquareChain 2^5)> <proof.bin (64 packed words)> */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;

u32 or_field_mul_xor(int N, int use16);
u32 or_g1_add_chain_x16(u64 k);
void or_msm_g1(u64 *out_aff, const u64 *pts, const u64 *sc, u64 n, int c);
void or_msm_g2(u64 *out_aff, const u64 *pts, const u64 *sc, u64 n, int c);
void or_g1_gen_mul_batch(u64 *out, const u64 *ks, u64 n);
void or_g2_gen_mul_batch(u64 *out, const u64 *ks, u64 n);
void or_fr_dot(u64 *out, const u64 *a, const u64 *k, u64 n);
void or_ntt(u64 *a, int log2n, int dir);
void or_qap_h(u64 *h, u64 *a, u64 *b, u64 *c, int log2n);
void or_set_threads(int n);
int or_prove(const u32 *hdr, const u64 *z, const u32 *a_off, const u32 *a_col, const u64 *a_coef, const u32 *b_off,
             const u32 *b_col, const u64 *b_coef, const u32 *c_off, const u32 *c_col, const u64 *c_coef,
             const u64 *pk_g1, const u64 *pk_g2, const u64 *pa, const u64 *pb1, const u64 *pb2, const u64 *ph,
             const u64 *pl, const u64 *r, const u64 *s, u64 *A_out, u64 *B_out, u64 *C_out, double *t_ms);

static int g_fail = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      g_fail++;                                                     \
    }                                                               \
  } while (0)

static u64 g_state = 0x9E3779B97F4A7C15ULL;
static u64 rnd(void) {  /* splitmix64 */
  u64 z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
/* uniform-ish value below r (top word below r's top word 0x30644e72e131a029) */
static void rand_fr(u64 *x) {
  x[0] = rnd();
  x[1] = rnd();
  x[2] = rnd();
  x[3] = rnd() % 0x30644e72e131a029ULL;
}

static u32 *read_file(const char *dir, const char *name, size_t *words) {
  char path[1024];
  if (dir)
    snprintf(path, sizeof path, "%s/%s", dir, name);
  else
    snprintf(path, sizeof path, "%s", name);
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  u32 *w = (u32 *)malloc(sz > 0 ? (size_t)sz : 4);
  size_t got = fread(w, 1, (size_t)sz, f);
  fclose(f);
  if (got != (size_t)sz) {
    free(w);
    return NULL;
  }
  *words = (size_t)sz / 4;
  return w;
}

/* 16-bit-limb words -> 64-bit limbs (4 words of 16 bits per limb) */
static u64 *pack16(const u32 *w16, size_t nwords) {
  u64 *out = (u64 *)calloc(nwords / 4 + 1, 8);
  for (size_t i = 0; i < nwords / 4; i++)
    out[i] = (u64)w16[4 * i] | (u64)w16[4 * i + 1] << 16 | (u64)w16[4 * i + 2] << 32 | (u64)w16[4 * i + 3] << 48;
  return out;
}

struct Csr {
  u32 *off, *col;
  u64 *coef;
};
static int read_csr(const char *dir, const char *name, u32 nc, struct Csr *c) {
  size_t n = 0;
  u32 *w = read_file(dir, name, &n);
  if (!w || n < (size_t)nc + 2) return -1;
  u32 nnz = w[0];
  if (n != 2 + (size_t)nc + 17 * (size_t)nnz) return -1;
  c->off = (u32 *)malloc(4 * ((size_t)nc + 1));
  memcpy(c->off, w + 1, 4 * ((size_t)nc + 1));
  c->col = (u32 *)malloc(4 * (size_t)(nnz ? nnz : 1));
  u32 *coef16 = (u32 *)malloc(64 * (size_t)(nnz ? nnz : 1));
  for (u32 e = 0; e < nnz; e++) {
    c->col[e] = w[2 + nc + 17 * (size_t)e];
    memcpy(coef16 + 16 * (size_t)e, w + 2 + nc + 17 * (size_t)e + 1, 64);
  }
  c->coef = pack16(coef16, 16 * (size_t)nnz);
  free(coef16);
  free(w);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: oracle_checks <key dir> <proof.bin>\n");
    return 2;
  }
  or_set_threads(2);
  /* the reference's committed known answers (bend2/probe/gpu_rounds.out.txt:3,5,7,9) */
  CHECK(or_field_mul_xor(18, 0) == 36312);
  CHECK(or_field_mul_xor(18, 1) == 36312);
  CHECK(or_g1_add_chain_x16(1u << 14) == 41096);

  /* MSMs against the closed form (sum a_i k_i) G */
  for (int g = 1; g <= 2; g++) {
    const u64 n = g == 1 ? 300 : 40;
    const int w = g == 1 ? 8 : 16;  /* affine words */
    u64 *k = malloc(32 * n), *a = malloc(32 * n), *pts = malloc(8 * w * n), dot[4], want[16], got[16];
    for (u64 i = 0; i < n; i++) {
      rand_fr(k + 4 * i);
      rand_fr(a + 4 * i);
    }
    if (g == 1) or_g1_gen_mul_batch(pts, k, n); else or_g2_gen_mul_batch(pts, k, n);
    or_fr_dot(dot, a, k, n);
    if (g == 1) or_g1_gen_mul_batch(want, dot, 1); else or_g2_gen_mul_batch(want, dot, 1);
    for (int c = 0; c <= 5; c += 5) {
      if (g == 1) or_msm_g1(got, pts, a, n, c); else or_msm_g2(got, pts, a, n, c);
      CHECK(memcmp(got, want, 8 * w) == 0);
    }
    free(k);
    free(a);
    free(pts);
  }

  /* NTT round trips in all four directions and qap_h on random inputs */
  {
    const int lg = 10;
    const u64 n = 1u << lg;
    u64 *x = malloc(32 * n), *y = malloc(32 * n), *b = malloc(32 * n), *c = malloc(32 * n), *h = malloc(32 * n);
    for (u64 i = 0; i < n; i++) rand_fr(x + 4 * i);
    memcpy(y, x, 32 * n);
    or_ntt(y, lg, 0);
    CHECK(memcmp(x, y, 32 * n) != 0);
    or_ntt(y, lg, 1);
    CHECK(memcmp(x, y, 32 * n) == 0);
    or_ntt(y, lg, 2);
    or_ntt(y, lg, 3);
    CHECK(memcmp(x, y, 32 * n) == 0);
    for (u64 i = 0; i < n; i++) {
      rand_fr(b + 4 * i);
      rand_fr(c + 4 * i);
    }
    or_qap_h(h, y, b, c, lg);
    free(x);
    free(y);
    free(b);
    free(c);
    free(h);
  }

  /* full prove of the key, equal to the oracle's own proof written by the test */
  {
    const char *dir = argv[1];
    size_t nh = 0, n;
    u32 *hdr = read_file(dir, "header.bin", &nh);
    CHECK(hdr && nh == 8);
    if (hdr && nh == 8) {
      const u32 nc = hdr[0];
      struct Csr A, B, C;
      CHECK(read_csr(dir, "r1cs_a.bin", nc, &A) == 0);
      CHECK(read_csr(dir, "r1cs_b.bin", nc, &B) == 0);
      CHECK(read_csr(dir, "r1cs_c.bin", nc, &C) == 0);
      const char *names[] = {"witness.bin", "pk_g1.bin", "pk_g2.bin", "pk_a.bin", "pk_b1.bin",
                             "pk_b2.bin", "pk_h.bin", "pk_l.bin", "rs.bin"};
      u64 *v[9];
      for (int i = 0; i < 9; i++) {
        u32 *w = read_file(dir, names[i], &n);
        CHECK(w != NULL);
        v[i] = w ? pack16(w, n) : NULL;
        free(w);
      }
      u64 out[32];
      CHECK(or_prove(hdr, v[0], A.off, A.col, A.coef, B.off, B.col, B.coef, C.off, C.col, C.coef, v[1], v[2], v[3],
                     v[4], v[5], v[6], v[7], v[8], v[8] + 4, out, out + 8, out + 24, NULL) == 0);
      u32 *proof = read_file(NULL, argv[2], &n);
      CHECK(proof && n == 64 && memcmp(proof, out, 256) == 0);
      free(proof);
      for (int i = 0; i < 9; i++) free(v[i]);
      free(A.off); free(A.col); free(A.coef);
      free(B.off); free(B.col); free(B.coef);
      free(C.off); free(C.col); free(C.coef);
    }
    free(hdr);
  }

  if (g_fail) {
    fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  printf("oracle checks OK\n");
  return 0;
}
