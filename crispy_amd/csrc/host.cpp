// Host-side parts of the drop-in boundary: the reference's FASTA reader,
// default thresholds, the filter predicate and the band-size formula.  No
// device code here.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "crispy_hip.h"

namespace {

constexpr int kLineBuf = 4096;  // fgets buffer of the reference (kmerMain.h:24)

// string::find_last_not_of(" \n\r\t\v\f\b") + substr: strips trailing
// whitespace, but leaves an all-whitespace line untouched (npos).
void trim_trailing_space(std::string &s) {
  const size_t k = s.find_last_not_of(" \n\r\t\v\f\b");
  if (k != std::string::npos) s.resize(k + 1);
}

// string::find_last_of(keep) + substr: cuts after the last kept base; a
// line with none is left untouched.
void trim_after_last(std::string &s, const char *keep) {
  const size_t k = s.find_last_of(keep);
  if (k != std::string::npos) s.resize(k + 1);
}

}  // namespace

extern "C" {

// Reader of kmerMain.cpp:298-397 (variant CRISPY_READS_KMER) and
// genMain.cpp:184-277 (CRISPY_READS_NW).  The line loop keeps the
// reference's behaviour on odd input: lines are read in 4095-byte pieces,
// a line starting with '$' ends the file, an all-whitespace line is kept
// (and lands inside the read until the final trim), and '>' lines with no
// sequence before them are skipped.
int32_t crispy_read_fasta(const char *path, int variant, char **bases_out, int64_t **offsets_out,
                          int32_t *maxlen_out) {
  const char *keep = variant == CRISPY_READS_KMER ? "ATGCatgc" : "nNATGCatgc";
  FILE *f = std::fopen(path, "rb");
  if (!f) return CRISPY_ERR_IO;
  char buffer[kLineBuf];
  std::vector<int64_t> offs{0};
  std::string all, seq, line;
  int32_t maxlen = 0;
  auto add_read = [&]() {
    trim_after_last(seq, keep);
    all += seq;
    offs.push_back((int64_t)all.size());
    if ((int32_t)seq.size() > maxlen) maxlen = (int32_t)seq.size();
    seq.clear();
  };
  if (!std::fgets(buffer, kLineBuf, f)) buffer[0] = 0;
  if (buffer[0] != '>') {
    std::fclose(f);
    return CRISPY_ERR_ARG;
  }
  bool stop = false;
  while (!std::feof(f)) {
    buffer[0] = '$';
    if (!std::fgets(buffer, kLineBuf, f)) {
      // end of file: buffer[0] stays '$'
    }
    line.assign(buffer);
    trim_trailing_space(line);
    const char c0 = line.empty() ? '\0' : line[0];
    if (c0 != '>' && c0 != '$') {
      trim_after_last(line, keep);
      seq += line;
    } else if (seq.empty()) {
      if (buffer[0] == '$') {
        stop = true;
        break;
      }
    } else {
      add_read();
      if (buffer[0] == '$') {
        stop = true;
        break;
      }
    }
  }
  if (!stop && !seq.empty()) add_read();
  std::fclose(f);
  const size_t n = offs.size() - 1;
  char *b = (char *)std::malloc(all.size() + 1);
  int64_t *o = (int64_t *)std::malloc(sizeof(int64_t) * offs.size());
  if (!b || !o) {
    std::free(b);
    std::free(o);
    return CRISPY_ERR_NOMEM;
  }
  std::memcpy(b, all.data(), all.size());
  std::memcpy(o, offs.data(), sizeof(int64_t) * offs.size());
  *bases_out = b;
  *offsets_out = o;
  if (maxlen_out) *maxlen_out = maxlen;
  return (int32_t)n;
}

void crispy_free(void *p) { std::free(p); }

// kmerMain.cpp:43-48
float crispy_kmer_default_threshold(int32_t n) {
  const double t = n > 500 ? 2 / std::log2((double)n) : 2 / std::log10((double)n);
  return (float)(t < 1.0 ? t : 1.0);
}

// genMain.cpp:42-47
float crispy_nw_default_threshold(int32_t n) {
  const double t = n > 500 ? 1 / std::log2((double)n) : 1 / std::log10((double)n);
  return (float)(t < 1.0 ? t : 1.0);
}

// kmerMain.cpp:180: fp32 difference, compared with the double EPSILON 1e-5
int crispy_pass(float dist, float threshold) {
  const float diff = dist - threshold;
  return dist < threshold || (double)std::fabs(diff) < 0.00001;
}

}  // extern "C"
