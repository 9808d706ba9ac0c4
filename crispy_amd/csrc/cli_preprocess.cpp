// bin/preprocess -- drop-in for the reference's preprocess executable on the
// path crispy.sh runs (preprocess.cpp:603-726): same flags and output files
// (<out> = <in>_Clean by default: "> label" + sequence per unique read, in
// (length, strcmp) order; <out>.frq: "label frequency" per line).  The
// ordering and the exact-duplicate merge run on the GPU through the C ABI.
// The primer filter (-p, a Smith-Waterman search for the primer at the start
// of each read, -m mismatches allowed) runs on the GPU too.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "crispy_hip.h"

static void help() {
  std::printf("<-i inFileName>\n");
  std::printf("[-o outFileName, default: <inFileName>_Clean]\n");
  std::printf("[-f input format a(FASTA) q(FASTQ), default: a\n");
  std::printf("[-p primerFileName]\n");
  std::printf("[-m number of mismatches to primers allowed, default 2]\n");
  std::printf("[-s standard deviation from average length allowed, default 1.0]\n");
  std::printf("[-h  help]\n");
}

static double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

int main(int argc, char *argv[]) {
  const double t0 = now_s();
  std::string in, out, primer_file;
  int fastq = 0, length_check = 0, primer_check = 0, max_mismatches = 2;
  double sd_allowed = 1.0;
  // preprocess.cpp:613-684: every flag is matched anywhere in argv
  for (int i = 0; i < argc; ++i) {
    const bool has_arg = i + 1 < argc;
    if (!std::strcmp(argv[i], "-i")) {
      if (!has_arg || std::strlen(argv[i + 1]) < 2) {
        help();
        return 255;
      }
      in = argv[i + 1];
    }
    if (!std::strcmp(argv[i], "-o") && has_arg) out = argv[i + 1];
    if (!std::strcmp(argv[i], "-f") && has_arg && !std::strcmp(argv[i + 1], "q")) fastq = 1;
    if (!std::strcmp(argv[i], "-p") && has_arg) {
      primer_file = argv[i + 1];
      primer_check = 1;
    }
    if (!std::strcmp(argv[i], "-m") && has_arg) max_mismatches = std::atoi(argv[i + 1]);
    if (!std::strcmp(argv[i], "-s") && has_arg) {
      sd_allowed = std::atof(argv[i + 1]);
      length_check = 1;
    }
    if (!std::strcmp(argv[i], "-h")) {
      help();
      return 255;
    }
  }
  if (in.empty()) {
    help();
    return 255;
  }
  if (out.length() < 2) out = in + "_Clean";
  std::string primer;
  if (primer_check) {  // preprocess.cpp:672-683: the first line, trailing whitespace removed
    std::ifstream pf(primer_file);
    if (!pf) {
      std::printf("Error: cannot open primer file. EXIT...");
      return 255;
    }
    std::getline(pf, primer);
    const size_t last = primer.find_last_not_of("\n\r \t\v\f\b");
    if (last != std::string::npos) primer.resize(last + 1);
  }

  std::printf("\n--------------------------------------------------------------------\n");
  std::printf("\n                            PREPROCESS\n");
  std::printf("Format: %s.\n", fastq ? "FASTQ" : "FASTA");
  char *bases = nullptr, *labels = nullptr;
  int64_t *offs = nullptr, *loffs = nullptr;
  int32_t *freq = nullptr;
  const int32_t n = crispy_preprocess_read(in.c_str(), fastq, &bases, &offs, &labels, &loffs, &freq);
  if (n < 0) {
    if (n == CRISPY_ERR_ARG)
      std::printf("Error: the read file is not in FASTA format. EXIT...\n");
    else
      std::printf("Can not open read file.\n");
    return 255;
  }
  std::printf("File name: %s.\n", in.c_str());
  int32_t deleted = 0;
  if (length_check) deleted += crispy_preprocess_length_filter(offs, bases, n, sd_allowed, freq);

  crispy_ctx *ctx = nullptr;
  int rc = crispy_ctx_create(0, &ctx);
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: no usable HIP device (code %d)\n", rc);
    return 255;
  }
  if (primer_check) {
    const int32_t d = crispy_preprocess_primer_filter(ctx, bases, offs, n, primer.data(), (int32_t)primer.size(),
                                                      max_mismatches, freq);
    if (d < 0) {
      std::fprintf(stderr, "Error: %s\n", crispy_ctx_error(ctx));
      crispy_ctx_destroy(ctx);
      return 255;
    }
    deleted += d;
  }
  std::vector<int32_t> order(n > 0 ? n : 1), ofreq(n > 0 ? n : 1);
  int32_t count = 0;
  rc = crispy_preprocess_order(ctx, bases, offs, freq, n, order.data(), ofreq.data(), &count);
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: %s\n", crispy_ctx_error(ctx));
    crispy_ctx_destroy(ctx);
    return 255;
  }
  int32_t live = 0;
  for (int32_t r = 0; r < n; ++r) live += freq[r] != 0;
  // the reference counts the merges of reads with frequency 1 or more
  deleted += live - count;
  std::printf("Total numReads: %d. Preprocessed numReads: %d\n", n, n - deleted);
  rc = crispy_preprocess_write(out.c_str(), fastq, bases, offs, labels, loffs, order.data(),
                               ofreq.data(), count);
  crispy_ctx_destroy(ctx);
  crispy_free(bases);
  crispy_free(offs);
  crispy_free(labels);
  crispy_free(loffs);
  crispy_free(freq);
  if (rc != CRISPY_OK) {
    std::printf("Error: cannot open output files. EXIT...");
    return 255;
  }
  std::printf("\nTime taken: %.3f secs\n", now_s() - t0);
  std::printf("\n--------------------------------------------------------------------\n");
  return 1;  // the reference's main returns 1 (preprocess.cpp:725)
}
