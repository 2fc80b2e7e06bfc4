// bg16_prove: command-line prover mirroring bend2/prove.bend (main :164-168, Pr.run
// :129-162).  The key directory comes from argv[1] or $BENDG_DATA (FORMAT.md files);
// stdout is "A x y", "B x0 x1 y0 y1", "C x y" (canonical big-endian hex) followed by
// "T <phase> <ms>" lines for load evals qap msm_a msm_b1 msm_b2 msm_h msm_l total.
// With --verify the proof is also checked against <key dir>/vk.bin with the public input
// z_1 (bg_verify): "V ok" on stdout, or exit status 1.
#include "../../include/bg16.h"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

static std::vector<uint32_t> read_all(const std::string& p) {
  std::ifstream f(p, std::ios::binary | std::ios::ate);
  if (!f.good()) {
    fprintf(stderr, "bg16_prove: cannot open %s\n", p.c_str());
    exit(1);
  }
  std::streamsize n = f.tellg();
  std::vector<uint32_t> w((size_t)n / 4);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(w.data()), n);
  return w;
}

static void die(const char* what, int rc) {
  fprintf(stderr, "bg16_prove: %s failed (%d): %s\n", what, rc, bg_last_error());
  exit(1);
}

int main(int argc, char** argv) {
  const char* dir = getenv("BENDG_DATA");
  bool verify = false;
  for (int i = 1; i < argc; i++) {
    if (std::string(argv[i]) == "--verify")
      verify = true;
    else
      dir = argv[i];
  }
  if (!dir) {
    fprintf(stderr, "usage: bg16_prove [--verify] <key dir>   (or set BENDG_DATA)\n");
    return 2;
  }
  auto t0 = std::chrono::steady_clock::now();
  int rc = bg_init(0);
  if (rc) die("bg_init", rc);
  bg_pk* pk = nullptr;
  if ((rc = bg_pk_load(dir, &pk))) die("bg_pk_load", rc);
  uint32_t hdr[8];
  bg_pk_header(pk, hdr);
  std::string d(dir);
  auto z = read_all(d + "/witness.bin");
  auto rs = read_all(d + "/rs.bin");
  if (z.size() != 16ull * (hdr[1] + hdr[2]) || rs.size() != 32) {
    fprintf(stderr, "bg16_prove: witness.bin / rs.bin size mismatch\n");
    return 1;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::vector<uint32_t> proof16(128);
  double t[9];
  if ((rc = bg_prove(pk, z.data(), rs.data(), proof16.data(), BG_FMT_WORDS16 | BG_MEM_HOST, t))) die("bg_prove", rc);
  auto t2 = std::chrono::steady_clock::now();
  uint32_t packed[64];
  for (int i = 0; i < 64; i++) packed[i] = proof16[2 * i] | (proof16[2 * i + 1] << 16);
  char buf[1024];
  int n = bg_proof_format(packed, buf, sizeof buf);
  if (n < 0) die("bg_proof_format", n);
  fputs(buf, stdout);
  double load = std::chrono::duration<double, std::milli>(t1 - t0).count();
  double total = std::chrono::duration<double, std::milli>(t2 - t0).count();
  printf("T load %.3f\nT evals %.3f\nT qap %.3f\nT msm_a %.3f\nT msm_b1 %.3f\nT msm_b2 %.3f\nT msm_h %.3f\n"
         "T msm_l %.3f\nT prove %.3f\nT total %.3f\n",
         load, t[0] + t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8], total);
  bg_pk_free(pk);
  if (verify) {
    auto vk = read_all(d + "/vk.bin");
    rc = bg_verify(vk.data(), vk.size(), proof16.data(), z.data() + 16, hdr[1] - 1, BG_FMT_WORDS16);
    if (rc < 0) die("bg_verify", rc);
    if (rc != 1) {
      fprintf(stderr, "bg16_prove: the proof does not verify against vk.bin\n");
      return 1;
    }
    printf("V ok\n");
  }
  return 0;
}
