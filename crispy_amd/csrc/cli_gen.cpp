// bin/genDist -- drop-in for the reference's genDist executable
// (genMain.cpp:11-115): same flags and defaults (-b absent or <= 0 -> 2,
// -b 1 -> full alignment), reads <in>.kpair in chunks of 589,824 pairs
// (genMain.cpp:129-139) and writes <in>.npair_<f> (uint32 pairs) and
// <in>.ndist_<f> (float32), f = 0, 1, ..., each ascending by distance.  A
// file is flushed when at least 147,456 passing pairs have accumulated after
// a chunk, and the rest forms the last file (computeGenDist_CPU_band,
// genCPU.cpp:484-507), so the file split equals the reference CPU path's.
// Within a file, ties are ordered by (i, j).
#include <sys/time.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "crispy_hip.h"

static const int64_t kChunkPairs = 576 * 256 * 4;  // NUM_PAIRS, genMain.h:24-27
static const int64_t kFlushPairs = 576 * 256;      // genCPU.cpp:339,484
static const int kDefaultBand = 2;                 // BAND, genMain.h:37

static void help() {
  std::printf("<-i inFileName>: FASTA file. Require: output from kmerDist (inFileName.kpair) \n");
  std::printf("[-t threshold] : value range: 0.0 to 1.0, default: 1/log2(numReads)\n");
  std::printf("[-n numThreads]: host threads (accepted for compatibility)\n");
  std::printf("[-b]           : use banded alignment, default: 2 i.e. 1/2 band width; 1 = full\n");
  std::printf("[-c]           : CPU path (not part of this build: it runs on the GPU)\n");
  std::printf("[-h]           : help\n");
}

static double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static int write_file(crispy_ctx *ctx, const std::string &pname, const std::string &dname, int fid,
                      std::vector<int32_t> &pairs, std::vector<float> &dists) {
  const int64_t n = (int64_t)dists.size();
  int rc = crispy_sort_by_dist(ctx, pairs.data(), dists.data(), n);
  if (rc != CRISPY_OK) return rc;
  char buf[32];
  std::snprintf(buf, sizeof buf, "_%d", fid);
  FILE *pf = std::fopen((pname + buf).c_str(), "wb");
  if (!pf) {
    std::printf("cannot open pairFile: %s\n", (pname + buf).c_str());
    return CRISPY_ERR_IO;
  }
  FILE *df = std::fopen((dname + buf).c_str(), "wb");
  if (!df) {
    std::printf("cannot open distFile: %s\n", (dname + buf).c_str());
    std::fclose(pf);
    return CRISPY_ERR_IO;
  }
  const bool ok = std::fwrite(pairs.data(), sizeof(uint32_t), 2 * n, pf) == (size_t)(2 * n) &&
                  std::fwrite(dists.data(), sizeof(float), n, df) == (size_t)n;
  std::fclose(pf);
  std::fclose(df);
  pairs.clear();
  dists.clear();
  return ok ? CRISPY_OK : CRISPY_ERR_IO;
}

int main(int argc, char *argv[]) {
  const double t0 = now_s();
  std::string in;
  float threshold = -1.f;
  int band = kDefaultBand;
  bool use_gpu = true;
  int num_threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  (void)num_threads;
  // genMain.cpp:334-373
  for (int i = 0; i < argc; ++i) {
    const bool has_arg = i + 1 < argc;
    if (!std::strcmp("-i", argv[i])) {
      if (!has_arg || std::strlen(argv[i + 1]) < 2) {
        help();
        return 0;
      }
      in = argv[i + 1];
    }
    if (!std::strcmp("-t", argv[i]) && has_arg) threshold = (float)std::atof(argv[i + 1]);
    if (!std::strcmp("-n", argv[i]) && has_arg) {
      num_threads = std::atoi(argv[i + 1]);
      if (num_threads < 1)
        std::printf("Warning: invalid number of threads (-n option). Set to %d.\n", num_threads);
    }
    if (!std::strcmp("-b", argv[i]) && has_arg) {
      band = std::atoi(argv[i + 1]);
      if (band <= 0) {
        std::printf("Warning: invalid band (-b option). Set to %d.\n", kDefaultBand);
        band = kDefaultBand;
      }
    }
    if (!std::strcmp("-c", argv[i])) use_gpu = false;
    if (!std::strcmp("-h", argv[i])) {
      help();
      return 0;
    }
  }
  if (!use_gpu) {
    std::fprintf(stderr, "Error: -c selects the reference's CPU path, which this build does not ship "
                         "(it computes on the GPU). Exit...\n");
    return 2;
  }
  char *bases = nullptr;
  int64_t *offs = nullptr;
  int32_t maxlen = 0;
  const int32_t n = crispy_read_fasta(in.c_str(), CRISPY_READS_NW, &bases, &offs, &maxlen);
  if (n == CRISPY_ERR_IO) {
    std::printf("Error: cannot open read file! EXIT...\n");
    return 255;
  }
  if (n < 0) {
    std::printf("Error: the read file is not in FASTA format! EXIT...\n");
    return 0;
  }
  if (n > CRISPY_MAX_NUM_READS || maxlen > CRISPY_MAX_READ_LEN) {
    std::printf("Error: unsupported numReads or maxLen. Exit...%d\t%d\n", n, maxlen);
    return 255;
  }
  if (threshold < 0 || threshold > 1) threshold = crispy_nw_default_threshold(n);
  std::printf("\n----------------------------------------------------------------------\n");
  std::printf("                    COMPUTE GENETIC DISTANCES                          \n");
  std::printf("File name: %s.\n", in.c_str());
  std::printf("numReads: %d. band: %d. threshold: %f\n\n", n, band, threshold);
  const std::string kpair = in + ".kpair";
  FILE *inp = std::fopen(kpair.c_str(), "rb");
  if (!inp) {
    std::printf("Error: cannot open pair file: %s. Exit...\n", kpair.c_str());
    return 255;
  }
  crispy_ctx *ctx = nullptr;
  int rc = crispy_ctx_create(0, &ctx);
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: no usable HIP device (code %d). Exit...\n", rc);
    return 255;
  }
  std::printf("USE GPU (HIP, gfx950).\n");
  rc = crispy_reads_upload(ctx, CRISPY_READS_NW, bases, offs, n);
  const std::string pname = in + ".npair", dname = in + ".ndist";
  std::vector<int32_t> chunk(2 * kChunkPairs);
  std::vector<int32_t> acc_pairs, got_pairs;
  std::vector<float> acc_dists, got_dists;
  int fid = 0;
  unsigned long long total = 0;
  bool eof = false;
  while (rc == CRISPY_OK && !eof) {
    const size_t got = std::fread(chunk.data(), sizeof(int32_t), 2 * kChunkPairs, inp);
    if (got < (size_t)(2 * kChunkPairs)) eof = true;
    const int64_t np = (int64_t)(got / 2);
    int64_t c = 0;
    rc = crispy_nw_run(ctx, chunk.data(), np, band, threshold, &c);
    if (rc != CRISPY_OK) break;
    got_pairs.resize(2 * c);
    got_dists.resize(c);
    rc = crispy_nw_fetch(ctx, 0, c, got_pairs.data(), got_dists.data());
    if (rc != CRISPY_OK) break;
    acc_pairs.insert(acc_pairs.end(), got_pairs.begin(), got_pairs.end());
    acc_dists.insert(acc_dists.end(), got_dists.begin(), got_dists.end());
    if ((int64_t)acc_dists.size() >= kFlushPairs) {
      total += acc_dists.size();
      rc = write_file(ctx, pname, dname, fid++, acc_pairs, acc_dists);
    }
  }
  if (rc == CRISPY_OK && !acc_dists.empty()) {
    total += acc_dists.size();
    rc = write_file(ctx, pname, dname, fid++, acc_pairs, acc_dists);
  }
  if (rc != CRISPY_OK) {
    std::fprintf(stderr, "Error: %s (code %d). Exit...\n", crispy_ctx_error(ctx), rc);
    return 255;
  }
  std::fclose(inp);
  crispy_ctx_destroy(ctx);
  crispy_free(bases);
  crispy_free(offs);
  std::printf("%llu\n", total);
  std::printf("Time taken: %.3f s\n", now_s() - t0);
  std::printf("\n----------------------------------------------------------------------\n");
  return 0;
}
