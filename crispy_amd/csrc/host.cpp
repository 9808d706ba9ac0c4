// Host-side parts of the drop-in boundary: the reference's FASTA reader,
// default thresholds, the filter predicate and the band-size formula.  No
// device code here.
#include <cctype>
#include <climits>
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

// ------------------------------------------------------------ input side
namespace {

struct PreRecords {
  std::string bases, labels;
  std::vector<int64_t> boff{0}, loff{0};
  std::vector<int32_t> freq;
  void add(const std::string &seq, const std::string &label, int32_t f) {
    bases += seq;
    boff.push_back((int64_t)bases.size());
    // READ::initialize drops every whitespace character of the label
    for (char c : label)
      if (!std::isspace((unsigned char)c)) labels.push_back(c);
    loff.push_back((int64_t)labels.size());
    freq.push_back(f);
  }
};

// preprocess.cpp:74-240, FASTA branch (format 'a')
int read_pre_fasta(FILE *f, PreRecords &R) {
  char buffer[kLineBuf];
  std::string line, seq, label;
  if (!std::fgets(buffer, kLineBuf, f)) buffer[0] = 0;
  if (buffer[0] != '>') return CRISPY_ERR_ARG;
  line.assign(buffer);
  trim_trailing_space(line);
  label = line;
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
      trim_after_last(line, "nNATGCatgc");
      seq += line;
    } else if (seq.empty()) {
      // a header with no sequence before it: skipped, and the label is kept
      if (buffer[0] == '$') {
        stop = true;
        break;
      }
    } else {
      trim_after_last(seq, "nNATGCatgc");
      R.add(seq, label, 1);
      seq.clear();
      if (buffer[0] == '$') {
        stop = true;
        break;
      }
      label = line;
    }
  }
  if (!stop && !seq.empty()) {
    trim_after_last(seq, "nNATGCatgc");
    R.add(seq, label, 1);
  }
  return CRISPY_OK;
}

// preprocess.cpp:74-240, FASTQ branch (format 'q'): label, sequence, '+',
// quality; a record with an empty sequence is kept with frequency 0.
int read_pre_fastq(FILE *f, PreRecords &R) {
  char buffer[kLineBuf];
  std::string line, seq, label;
  auto next = [&]() -> bool {  // false at the end (feof or nothing read)
    buffer[0] = '$';
    if (!std::fgets(buffer, kLineBuf, f)) {
    }
    return !(std::feof(f) || buffer[0] == '$');
  };
  while (!std::feof(f)) {
    if (!next()) break;
    line.assign(buffer);
    trim_trailing_space(line);
    label = line;
    if (!next()) break;
    line.assign(buffer);
    trim_trailing_space(line);
    seq = line;
    R.add(seq, label, seq.empty() ? 0 : 1);
    if (!next()) break;
    if (!next()) break;
  }
  return CRISPY_OK;
}

template <class T>
T *copy_out(const T *src, size_t n) {
  T *p = (T *)std::malloc(sizeof(T) * (n ? n : 1));
  if (p && n) std::memcpy(p, src, sizeof(T) * n);
  return p;
}

}  // namespace

extern "C" {

int32_t crispy_preprocess_read(const char *path, int fastq, char **bases, int64_t **offsets,
                               char **labels, int64_t **label_offsets, int32_t **freq) {
  if (!path || !bases || !offsets || !labels || !label_offsets || !freq) return CRISPY_ERR_ARG;
  FILE *f = std::fopen(path, "rb");
  if (!f) return CRISPY_ERR_IO;
  PreRecords R;
  const int rc = fastq ? read_pre_fastq(f, R) : read_pre_fasta(f, R);
  std::fclose(f);
  if (rc != CRISPY_OK) return rc;
  const size_t n = R.freq.size();
  if (n > (size_t)INT32_MAX) return CRISPY_ERR_ARG;
  *bases = copy_out(R.bases.data(), R.bases.size());
  *offsets = copy_out(R.boff.data(), R.boff.size());
  *labels = copy_out(R.labels.data(), R.labels.size());
  *label_offsets = copy_out(R.loff.data(), R.loff.size());
  *freq = copy_out(R.freq.data(), n);
  if (!*bases || !*offsets || !*labels || !*label_offsets || !*freq) {
    std::free(*bases);
    std::free(*offsets);
    std::free(*labels);
    std::free(*label_offsets);
    std::free(*freq);
    return CRISPY_ERR_NOMEM;
  }
  return (int32_t)n;
}

// preprocess.cpp:322-392 (-s): over the records with frequency > 0, the mean
// length and population SD; records outside [max(min, (int)(mean - s SD)),
// min(max, (int)(mean + s SD))] or containing 'n'/'N' get frequency 0.
// Returns the number discarded.
int32_t crispy_preprocess_length_filter(const int64_t *offsets, const char *bases, int32_t n,
                                        double sd_allowed, int32_t *freq) {
  if (n < 0 || (n > 0 && (!offsets || !bases || !freq))) return CRISPY_ERR_ARG;
  double sum = 0, cnt = 0;
  for (int32_t r = 0; r < n; ++r)
    if (freq[r] != 0) {
      sum += (double)(offsets[r + 1] - offsets[r]);
      cnt += 1;
    }
  if (cnt < 1) return 0;
  const double ave = sum / cnt;
  double sq = 0;
  int minLen = 65535, maxLen = 0;
  for (int32_t r = 0; r < n; ++r)
    if (freq[r] != 0) {
      const int len = (int)(offsets[r + 1] - offsets[r]);
      const double d = len - ave;
      sq += d * d;
      minLen = len < minLen ? len : minLen;
      maxLen = len > maxLen ? len : maxLen;
    }
  const double sd = std::sqrt(sq / cnt);
  const int lo = (int)(ave - sd_allowed * sd), hi = (int)(ave + sd_allowed * sd);
  minLen = minLen > lo ? minLen : lo;
  maxLen = maxLen < hi ? maxLen : hi;
  int32_t discarded = 0;
  for (int32_t r = 0; r < n; ++r) {
    if (freq[r] == 0) continue;
    const int64_t len = offsets[r + 1] - offsets[r];
    if (len < minLen || len > maxLen) {
      freq[r] = 0;
      ++discarded;
      continue;
    }
    for (int64_t q = offsets[r]; q < offsets[r + 1]; ++q)
      if (bases[q] == 'n' || bases[q] == 'N') {
        freq[r] = 0;
        ++discarded;
        break;
      }
  }
  return discarded;
}

// preprocess.cpp:562-601: "> label-without-its-first-character" + sequence
// per read (FASTQ: "> label"), and "label freq" lines in <out>.frq.
int crispy_preprocess_write(const char *out_path, int fastq, const char *bases,
                            const int64_t *offsets, const char *labels,
                            const int64_t *label_offsets, const int32_t *order,
                            const int32_t *freq, int32_t count) {
  if (!out_path || count < 0 || (count > 0 && (!bases || !offsets || !labels || !label_offsets ||
                                               !order || !freq)))
    return CRISPY_ERR_ARG;
  std::string frq_path = std::string(out_path) + ".frq";
  FILE *out = std::fopen(out_path, "wb");
  FILE *fq = std::fopen(frq_path.c_str(), "wb");
  if (!out || !fq) {
    if (out) std::fclose(out);
    if (fq) std::fclose(fq);
    return CRISPY_ERR_IO;
  }
  for (int32_t k = 0; k < count; ++k) {
    const int32_t r = order[k];
    const char *lab = labels + label_offsets[r];
    int64_t llen = label_offsets[r + 1] - label_offsets[r];
    if (!fastq && llen > 0) {  // label.substr(1)
      ++lab;
      --llen;
    }
    std::fputs("> ", out);
    std::fwrite(lab, 1, (size_t)llen, out);
    std::fputc('\n', out);
    std::fwrite(bases + offsets[r], 1, (size_t)(offsets[r + 1] - offsets[r]), out);
    std::fputc('\n', out);
    std::fwrite(lab, 1, (size_t)llen, fq);
    std::fprintf(fq, " %d\n", freq[k]);
  }
  const bool ok = std::fclose(out) == 0;
  return (std::fclose(fq) == 0 && ok) ? CRISPY_OK : CRISPY_ERR_IO;
}

}  // extern "C"
