// bin/kmerDist -- drop-in for the reference's kmerDist executable
// (kmerMain.cpp:12-93): same flags, defaults, limits and output files
// (<in>.kpair: int32 (i, j) pairs, i < j; <in>.kdist: float32, with -d).
// The distances are computed on the GPU through the C ABI.  Pairs are
// written in (i, j) order, the order the reference's CPU path produces with
// one thread.
#include <sys/time.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "crispy_hip.h"

static void help() {
  std::printf("<-i inFileName>: FASTA file \n");
  std::printf("[-k K]         : k-mer length, 1 to 8, default: 6\n");
  std::printf("[-t threshold] : value range: 0.0 to 1.0, default: 2/log2(numReads)\n");
  std::printf("[-n numThreads]: host threads (accepted for compatibility)\n");
  std::printf("[-d]           : output kmer distances \n");
  std::printf("[-c]           : CPU path (not part of this build: it runs on the GPU)\n");
  std::printf("[-h]           : help\n");
}

static double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

int main(int argc, char *argv[]) {
  const double t0 = now_s();
  std::string in;
  float threshold = -1.f;
  bool has_dist = false, use_gpu = true;
  long nprocs = sysconf(_SC_NPROCESSORS_ONLN);
  int num_threads = nprocs > 2 ? (int)(nprocs / 2) : 1;  // kmerMain.cpp:232-236
  int K = 6;
  // kmerMain.cpp:229-272: every flag is matched anywhere in argv
  for (int i = 0; i < argc; ++i) {
    const bool has_arg = i + 1 < argc;
    if (!std::strcmp("-i", argv[i])) {
      if (!has_arg || std::strlen(argv[i + 1]) < 2) {
        help();
        return 255;
      }
      in = argv[i + 1];
    }
    if (!std::strcmp("-t", argv[i]) && has_arg) threshold = (float)std::atof(argv[i + 1]);
    if (!std::strcmp("-n", argv[i]) && has_arg) {
      num_threads = std::atoi(argv[i + 1]);
      if (num_threads < 1 || num_threads > nprocs)
        std::printf("Warning: invalid number of threads (-n option). Set to %d.\n", num_threads);
    }
    if (!std::strcmp("-k", argv[i]) && has_arg) K = std::atoi(argv[i + 1]);
    if (!std::strcmp("-d", argv[i])) has_dist = true;
    if (!std::strcmp("-c", argv[i])) use_gpu = false;
    if (!std::strcmp("-h", argv[i])) {
      help();
      return 255;
    }
  }
  if (!use_gpu) {
    std::fprintf(stderr, "Error: -c selects the reference's CPU path, which this build does not ship "
                         "(it computes on the GPU). Exit...\n");
    return 2;
  }
  if (K < 1 || K > 8) {
    std::printf("Error: unsupported K: %d (1 to 8). Exit...\n", K);
    return 255;
  }
  char *bases = nullptr;
  int64_t *offs = nullptr;
  int32_t maxlen = 0;
  const int32_t n = crispy_read_fasta(in.c_str(), CRISPY_READS_KMER, &bases, &offs, &maxlen);
  if (n == CRISPY_ERR_IO) {
    std::printf("Error: cannot open read file: %s Exit...\n", in.c_str());
    return 255;
  }
  if (n < 0) {
    std::printf("Error: read file is not in FASTA Format. Exit...\n");
    return 255;
  }
  if (n > CRISPY_MAX_NUM_READS || maxlen > CRISPY_MAX_READ_LEN) {
    std::printf("Error: unsupported numReads: %d or maxLen: %d. Exit...", n, maxlen);
    return 255;
  }
  for (int32_t r = 0; r < n; ++r) {
    if (offs[r + 1] - offs[r] < K) {
      std::printf("Error: read %d is shorter than K = %d. Exit...\n", r, K);
      return 255;
    }
  }
  if (threshold < 0 || threshold > 1) threshold = crispy_kmer_default_threshold(n);
  std::printf("\n----------------------------------------------------------------------\n");
  std::printf("                       COMPUTE KMER DISTANCES                           \n");
  std::printf("File name: %s. K = %d\n\n", in.c_str(), K);
  std::printf("numReads: %d, maxLen: %d, threshold: %.2f\n  ", n, maxlen, threshold);

  const std::string pair_name = in + ".kpair", dist_name = in + ".kdist";
  FILE *pf = std::fopen(pair_name.c_str(), "wb");
  FILE *df = has_dist ? std::fopen(dist_name.c_str(), "wb") : nullptr;
  if (!pf || (has_dist && !df)) {
    std::printf("Error: cannot open output files. Exit...\n");
    return 255;
  }
  crispy_ctx *ctx = nullptr;
  int rc = crispy_ctx_create(0, &ctx);
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: no usable HIP device (code %d). Exit...\n", rc);
    return 255;
  }
  std::printf("USE GPU (HIP, gfx950). tile slots: %lld (%dx%d reads each)\n",
              (long long)crispy_kmer_num_tiles(n), 128, 128);
  int64_t count = 0;
  rc = crispy_reads_upload(ctx, CRISPY_READS_KMER, bases, offs, n);
  if (rc == CRISPY_OK) rc = crispy_kmer_run(ctx, K, threshold, 0, -1, &count);
  const int64_t block = 1 << 20;
  std::vector<int32_t> pairs;
  std::vector<float> dists;
  for (int64_t first = 0; rc == CRISPY_OK && first < count; first += block) {
    const int64_t m = count - first < block ? count - first : block;
    pairs.resize(2 * m);
    dists.resize(m);
    rc = crispy_kmer_fetch(ctx, first, m, pairs.data(), has_dist ? dists.data() : nullptr);
    if (rc != CRISPY_OK) break;
    if (std::fwrite(pairs.data(), sizeof(int32_t), 2 * m, pf) != (size_t)(2 * m) ||
        (has_dist && std::fwrite(dists.data(), sizeof(float), m, df) != (size_t)m)) {
      std::printf("Error: write failed. Exit...\n");
      return 255;
    }
  }
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: %s (code %d). Exit...\n", crispy_ctx_error(ctx), rc);
    return 255;
  }
  crispy_ctx_destroy(ctx);
  std::fclose(pf);
  if (df) std::fclose(df);
  crispy_free(bases);
  crispy_free(offs);
  std::printf("totalNumPairs: %llu\n", (unsigned long long)count);
  std::printf("Time taken: %.3f s\n", now_s() - t0);
  std::printf("\n----------------------------------------------------------------------\n");
  return 0;
}
