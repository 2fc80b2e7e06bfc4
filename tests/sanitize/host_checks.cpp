// The library's host-side code under AddressSanitizer and UndefinedBehaviorSanitizer, on
// a machine without a GPU: key generation (bg_setup.cpp) on one and several threads, the
// FORMAT.md key reader and its validation (bg_pk_load's host half, its thread pool and the
// on-curve checks), the pairing verifier (bg_verify.cpp) on an oracle-made proof, the
// host scalar multiplications, the tuning-table parser (bg_tune.cpp), proof formatting,
// and every entry point's error path when no device is present.
//
// usage: host_checks <work dir> <proof.bin of key5 (64 packed words)> <tune table>
// where <work dir>/key5 holds a SquareChain 2^5 key made with seed 5.
#include "bg16.h"
#include "bg_host_ec.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

static int g_fail = 0;
#define CHECK(c)                                                                            \
  do {                                                                                      \
    if (!(c)) {                                                                             \
      std::fprintf(stderr, "check failed: %s (line %d): %s\n", #c, __LINE__, bg_last_error()); \
      g_fail++;                                                                             \
    }                                                                                       \
  } while (0)

static std::vector<uint32_t> read_words(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.good()) return {};
  std::vector<uint32_t> w((size_t)f.tellg() / 4);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(w.data()), (std::streamsize)(w.size() * 4));
  return w;
}
static void write_words(const std::string& path, const std::vector<uint32_t>& w) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(w.size() * 4));
}
static void copy_key(const std::string& from, const std::string& to) {
  std::string cmd = "rm -rf '" + to + "' && cp -r '" + from + "' '" + to + "'";
  CHECK(std::system(cmd.c_str()) == 0);
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: host_checks <work dir> <proof.bin> <tune table>\n");
    return 2;
  }
  const std::string work = argv[1], key = work + "/key5";
  const char* names[] = {"header.bin", "r1cs_a.bin", "r1cs_b.bin", "r1cs_c.bin", "pk_g1.bin", "pk_g2.bin",
                         "pk_a.bin",   "pk_b1.bin",  "pk_b2.bin",  "pk_h.bin",   "pk_l.bin",  "vk.bin",
                         "witness.bin", "rs.bin"};

  // key generation: the same key on one and on four threads
  CHECK(std::system(("mkdir -p '" + work + "/s1' '" + work + "/s4'").c_str()) == 0);
  CHECK(bg_setup_square_chain((work + "/s1").c_str(), 5, 5, 1) == BG_OK);
  CHECK(bg_setup_square_chain((work + "/s4").c_str(), 5, 5, 4) == BG_OK);
  for (const char* n : names) {
    auto a = read_words(work + "/s1/" + n), b = read_words(work + "/s4/" + n), c = read_words(key + "/" + n);
    CHECK(!a.empty() && a == b && a == c);
  }
  CHECK(bg_setup_square_chain((work + "/bad").c_str(), 0, 5, 1) == BG_ERR_ARG);

  // key reader: a valid key passes every host check and stops at the missing device;
  // malformed files stop at the reader
  bg_pk* pk = nullptr;
  CHECK(bg_pk_load(key.c_str(), &pk) == BG_ERR_DEVICE && pk == nullptr);
  enum Edit { ADD_TO_LIMB, SET, TRUNCATE };
  struct Case {
    const char* file;
    size_t word;
    uint32_t value;
    Edit edit;
  } cases[] = {{"pk_l.bin", 20, 1, ADD_TO_LIMB},    // a y limb of the first L point: off the curve
               {"pk_b2.bin", 3, 7, ADD_TO_LIMB},    // an x limb of the first B2 point
               {"pk_g2.bin", 40, 3, ADD_TO_LIMB},   // beta2's y
               {"pk_a.bin", 5, 0x10000, SET},       // a limb word >= 2^16
               {"header.bin", 6, 1u << 20, SET},    // n_h beyond the domain
               {"r1cs_b.bin", 0, 0, TRUNCATE}};
  for (const Case& cs : cases) {
    const std::string d = work + "/mal";
    copy_key(key, d);
    auto w = read_words(d + "/" + cs.file);
    if (cs.edit == TRUNCATE)
      w.pop_back();
    else if (cs.edit == SET)
      w[cs.word] = cs.value;
    else
      w[cs.word] = (w[cs.word] + cs.value) & 0xFFFFu;
    write_words(d + "/" + cs.file, w);
    bg_pk* bad = nullptr;
    CHECK(bg_pk_load(d.c_str(), &bad) == BG_ERR_IO && bad == nullptr);
  }

  // the pairing verifier on the oracle's proof of key5 (vk.bin and the witness are FORMAT.md
  // words, so the proof goes in as words too)
  auto vk = read_words(key + "/vk.bin");
  auto proof = read_words(argv[2]);
  auto z16 = read_words(key + "/witness.bin");
  CHECK(proof.size() == 64 && z16.size() >= 32);
  std::vector<uint32_t> p16(128);
  for (int j = 0; j < 64; j++) {
    p16[2 * j] = proof[j] & 0xFFFFu;
    p16[2 * j + 1] = proof[j] >> 16;
  }
  const uint32_t* x = z16.data() + 16;  // the public input
  CHECK(bg_verify(vk.data(), vk.size(), p16.data(), x, 1, BG_FMT_WORDS16) == 1);
  auto swapped = p16;
  for (int j = 0; j < 32; j++) std::swap(swapped[j], swapped[96 + j]);  // A <-> C
  CHECK(bg_verify(vk.data(), vk.size(), swapped.data(), x, 1, BG_FMT_WORDS16) == 0);
  auto off = p16;
  off[16] ^= 1;  // A.y changed: off the curve
  CHECK(bg_verify(vk.data(), vk.size(), off.data(), x, 1, BG_FMT_WORDS16) == BG_ERR_ARG);
  char text[1024];
  const int n = bg_proof_format(proof.data(), text, sizeof text);
  CHECK(n > 0 && std::strncmp(text, "A 0x", 4) == 0);
  CHECK(bg_proof_format(proof.data(), text, 10) == BG_ERR_ARG);

  // bilinearity: e(aP, bQ) e(-ab P, Q) = 1 with host scalar multiplications
  using bgh::Fr;
  const Fr a = Fr::to_mont(Fr::raw(0x1234567890abcdefULL, 77, 5)), b = Fr::to_mont(Fr::raw(0x5555, 0, 0, 9));
  const Fr ab = -(a * b);
  uint32_t sc[3][8];
  a.from_mont().to_u32(sc[0]);
  ab.from_mont().to_u32(sc[1]);
  uint32_t one[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  uint32_t g1s[2 * 16], g2s[2 * 32];
  CHECK(bg_fixed_base_mul(1, sc[0], 1, g1s, 2) == BG_OK);
  CHECK(bg_fixed_base_mul(1, sc[1], 1, g1s + 16, 2) == BG_OK);
  b.from_mont().to_u32(sc[2]);
  CHECK(bg_fixed_base_mul(2, sc[2], 1, g2s, 2) == BG_OK);
  CHECK(bg_fixed_base_mul(2, one, 1, g2s + 32, 1) == BG_OK);
  CHECK(bg_pairing_check(g1s, g2s, 2, BG_FMT_PACKED) == 1);
  uint32_t g1bad[2 * 16];
  std::memcpy(g1bad, g1s, sizeof g1bad);
  std::memcpy(g1bad + 16, g1s, 64);  // e(aP, bQ) e(aP, Q) != 1
  CHECK(bg_pairing_check(g1bad, g2s, 2, BG_FMT_PACKED) == 0);

  // seeded synthetic bases equal the fixed-base products of their scalars
  std::vector<uint32_t> bases(100 * 16), ks(100 * 8), again(100 * 16);
  CHECK(bg_gen_bases(1, 9, 100, bases.data(), ks.data(), 3) == BG_OK);
  CHECK(bg_fixed_base_mul(1, ks.data(), 100, again.data(), 1) == BG_OK);
  CHECK(bases == again);

  // tuning table
  CHECK(bg_tune_load(argv[3]) == BG_OK);
  int v4[4];
  CHECK(bg_tune_lookup(1, 0, 0, 19, v4) == 0 && v4[0] >= 9 && v4[0] <= 11);
  {
    const std::string bad = work + "/bad_table.txt";
    std::ofstream(bad) << "msm 1 windowed 14 c=99\n";
    CHECK(bg_tune_load(bad.c_str()) == BG_ERR_IO);
  }

  // device entry points without a device: an error, not a crash
  uint32_t buf[64] = {0};
  CHECK(bg_init(0) == BG_ERR_DEVICE && std::strlen(bg_last_error()) > 0);
  CHECK(bg_msm_g1(buf, buf, 1, buf, BG_FMT_PACKED, 0) == BG_ERR_DEVICE);
  CHECK(bg_ntt(buf, 3, BG_NTT_FFT, BG_FMT_PACKED) == BG_ERR_DEVICE);
  bg_engine* e = nullptr;
  CHECK(bg_engine_create(2, 0, &e) == BG_ERR_DEVICE && e == nullptr);
  CHECK(bg_qap_h_noalloc(buf, buf, buf, 30, buf) == BG_ERR_ARG);

  if (g_fail) {
    std::fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("host checks OK\n");
  return 0;
}
