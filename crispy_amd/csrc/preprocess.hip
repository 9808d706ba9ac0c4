// Input side on gfx950: the ordering and exact-duplicate merge of the
// reference's preprocess stage (preprocess.cpp:244-280, READ::operator<
// :42-49).  Reads are ordered by (length, strcmp) with a stable sort
// (std::list::sort); each run of identical sequences collapses onto its last
// read, which carries the run's summed frequency.  Reads with frequency 0
// (discarded by the length filter, empty FASTQ records) take no part, as in
// the reference's merge loop.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace crispy {

// 0 when reads a and b are identical, < 0 / > 0 like (length, strcmp) order
__device__ __forceinline__ int pre_cmp(const uint8_t *__restrict__ bases, const int64_t *__restrict__ off,
                                       int32_t a, int32_t b) {
  const int64_t la = off[a + 1] - off[a], lb = off[b + 1] - off[b];
  if (la != lb) return la < lb ? -1 : 1;
  const uint8_t *pa = bases + off[a], *pb = bases + off[b];
  for (int64_t q = 0; q < la; ++q)
    if (pa[q] != pb[q]) return pa[q] < pb[q] ? -1 : 1;  // strcmp compares unsigned chars
  return 0;
}

// Strict total order: (length, bytes), then the read index, so any sort gives
// the stable order of the reference.
struct PreLess {
  const uint8_t *bases;
  const int64_t *off;
  __device__ bool operator()(int32_t a, int32_t b) const {
    const int c = pre_cmp(bases, off, a, b);
    return c != 0 ? c < 0 : a < b;
  }
};

// flags[k] = 1 when sorted position k ends a run of identical reads
__global__ void pre_run_end_kernel(const uint8_t *__restrict__ bases, const int64_t *__restrict__ off,
                                   const int32_t *__restrict__ keys, int32_t m,
                                   uint8_t *__restrict__ flags) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
       k += (int64_t)gridDim.x * blockDim.x)
    flags[k] = (k == m - 1 || pre_cmp(bases, off, keys[k], keys[k + 1]) != 0) ? 1 : 0;
}

__global__ void pre_sorted_freq_kernel(const int32_t *__restrict__ keys, int32_t m,
                                       const int32_t *__restrict__ freq, int64_t *__restrict__ f) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
       k += (int64_t)gridDim.x * blockDim.x)
    f[k] = freq[keys[k]];
}

// run t ends at sorted position sel[t]: its read, and its frequency as a
// difference of the inclusive prefix sums
__global__ void pre_runs_kernel(const uint32_t *__restrict__ sel, const unsigned int *__restrict__ nsel,
                                const int32_t *__restrict__ keys, const int64_t *__restrict__ csum,
                                int32_t *__restrict__ out_order, int32_t *__restrict__ out_freq) {
  const int64_t nr = *nsel;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nr;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t e = sel[t];
    out_order[t] = keys[e];
    out_freq[t] = (int32_t)(csum[e] - (t > 0 ? csum[sel[t - 1]] : 0));
  }
}

// processPrimer + SWLocalAlignment (preprocess.cpp:394-560), one lane per
// read: local alignment of the primer (rows i) against the read's first
// min(2 |primer|, L) bytes (columns j), +4 / -1 / gap -2, floored at 0; per
// cell the primer start and the mismatch count of the path, ties resolved 0,
// diagonal, up, left; the best cell is the first in row-major order with the
// highest score and, among those, the fewest mismatches; unaligned primer ends
// count as mismatches.  A read fails when that count exceeds max_mm or the
// alignment is empty; its frequency becomes 0.  The lane keeps one DP row
// (score, start, mismatches) in a lane-interleaved scratch (columns x lanes),
// the diagonal and the left cell in registers.
__global__ __launch_bounds__(256) void pre_primer_kernel(const uint8_t *__restrict__ bases,
                                                        const int64_t *__restrict__ off, int32_t n,
                                                        const uint8_t *__restrict__ primer, int lenA,
                                                        int max_mm, int32_t *__restrict__ freq,
                                                        int *__restrict__ scratch,
                                                        unsigned int *__restrict__ discarded) {
  const int64_t lanes = (int64_t)gridDim.x * blockDim.x;
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int W = 2 * lenA + 1;
  int *S = scratch + lane, *ST = scratch + (int64_t)W * lanes + lane, *MM = scratch + 2 * (int64_t)W * lanes + lane;
  unsigned int mine = 0;
  for (int64_t r = lane; r < n; r += lanes) {
    if (freq[r] == 0) continue;
    const int64_t L = off[r + 1] - off[r];
    const int lenB = (int)(L < 2 * (int64_t)lenA ? L : 2 * (int64_t)lenA);
    const uint8_t *B = bases + off[r];
    for (int j = 0; j <= lenB; ++j) {  // row 0
      S[j * lanes] = 0;
      ST[j * lanes] = 0;
      MM[j * lanes] = 0;
    }
    int best = -9999, bmm = 999, bstart = 0, bend = 0;
    for (int i = 1; i <= lenA; ++i) {
      const uint8_t a = primer[i - 1];
      int dS = S[0], dST = ST[0], dMM = MM[0];  // (i - 1, 0)
      int lS = 0, lST = i, lMM = 0;             // (i, 0)
      S[0] = 0;
      ST[0] = i;
      MM[0] = 0;
      for (int j = 1; j <= lenB; ++j) {
        const int uS = S[j * lanes], uST = ST[j * lanes], uMM = MM[j * lanes];  // (i - 1, j)
        const int al = a == B[j - 1] ? 4 : -1;
        const int t0 = dS + al, t1 = uS - 2, t2 = lS - 2;
        const int v = max(max(max(t0, t1), t2), 0);
        int st, m;
        if (v == 0) {
          st = i - 1;
          m = 0;
        } else if (v == t0) {
          st = dST;
          m = dMM + (al != 4);
        } else if (v == t1) {
          st = uST;
          m = uMM + 1;
        } else {
          st = lST;
          m = lMM + 1;
        }
        S[j * lanes] = v;
        ST[j * lanes] = st;
        MM[j * lanes] = m;
        if (best < v || (best == v && bmm > m)) {
          best = v;
          bstart = st;
          bend = i;
          bmm = m;
        }
        dS = uS;
        dST = uST;
        dMM = uMM;
        lS = v;
        lST = st;
        lMM = m;
      }
    }
    bmm += bstart + (lenA > bend ? lenA - bend : bend - lenA);
    if (bmm > max_mm || bstart == bend) {
      freq[r] = 0;
      ++mine;
    }
  }
  if (mine) atomicAdd(discarded, mine);
}

static inline int pre_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

}  // namespace crispy

using namespace crispy;

extern "C" {

int crispy_preprocess_order(crispy_ctx *ctx, const char *bases, const int64_t *offsets,
                            const int32_t *freq, int32_t n, int32_t *order, int32_t *out_freq,
                            int32_t *count) {
  if (!ctx || !count || n < 0 || (n > 0 && (!bases || !offsets || !freq || !order || !out_freq)))
    return CRISPY_ERR_ARG;
  *count = 0;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<int32_t> keys;
  keys.reserve(n);
  for (int32_t r = 0; r < n; ++r) {
    if (offsets[r + 1] < offsets[r]) return set_err(ctx, CRISPY_ERR_ARG, "offsets decrease at read %d", r);
    if (freq[r] < 0) return set_err(ctx, CRISPY_ERR_ARG, "negative frequency at read %d", r);
    if (freq[r] != 0) keys.push_back(r);
  }
  const int32_t m = (int32_t)keys.size();
  if (m == 0) return CRISPY_OK;
  PreState &P = ctx->pre;
  const int64_t nb = offsets[n] - offsets[0];
  CRISPY_TRY(ensure(ctx, P.bases, (size_t)nb + 16));
  CRISPY_TRY(ensure(ctx, P.off, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, P.freq, sizeof(int32_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, P.keys, sizeof(int32_t) * m));
  CRISPY_TRY(ensure(ctx, P.keys_alt, sizeof(int32_t) * m));
  CRISPY_TRY(ensure(ctx, P.flags, (size_t)m));
  CRISPY_TRY(ensure(ctx, P.sel, sizeof(uint32_t) * m));
  CRISPY_TRY(ensure(ctx, P.nsel, sizeof(unsigned int)));
  CRISPY_TRY(ensure(ctx, P.csum, sizeof(int64_t) * m));
  CRISPY_TRY(ensure(ctx, P.out_order, sizeof(int32_t) * m));
  CRISPY_TRY(ensure(ctx, P.out_freq, sizeof(int32_t) * m));
  // offsets relative to the first read
  std::vector<int64_t> off0(offsets, offsets + n + 1);
  for (auto &o : off0) o -= offsets[0];
  CRISPY_HIP(ctx, hipMemcpyAsync(P.bases.p, bases + offsets[0], (size_t)nb, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(P.off.p, off0.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(P.freq.p, freq, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(P.keys.p, keys.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, ctx->stream));

  const uint8_t *d_bases = P.bases.as<uint8_t>();
  const int64_t *d_off = P.off.as<int64_t>();
  const PreLess less{d_bases, d_off};
  hipEvent_t ev;
  timer_begin(ctx, "pre_sort", &ev);
  size_t tb = 0;
  CRISPY_HIP(ctx, rocprim::merge_sort(nullptr, tb, P.keys.as<int32_t>(), P.keys_alt.as<int32_t>(),
                                      (size_t)m, less, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tb));
  CRISPY_HIP(ctx, rocprim::merge_sort(ctx->temp.p, tb, P.keys.as<int32_t>(), P.keys_alt.as<int32_t>(),
                                      (size_t)m, less, ctx->stream));
  timer_end(ctx, "pre_sort", ev, (double)m);
  const int32_t *sorted = P.keys_alt.as<int32_t>();

  timer_begin(ctx, "pre_merge", &ev);
  pre_run_end_kernel<<<pre_grid(m), 256, 0, ctx->stream>>>(d_bases, d_off, sorted, m, P.flags.as<uint8_t>());
  CRISPY_LAUNCHED(ctx);
  int64_t *f64 = P.csum.as<int64_t>();
  pre_sorted_freq_kernel<<<pre_grid(m), 256, 0, ctx->stream>>>(sorted, m, P.freq.as<int32_t>(), f64);
  CRISPY_LAUNCHED(ctx);
  tb = 0;
  CRISPY_HIP(ctx, rocprim::inclusive_scan(nullptr, tb, f64, f64, (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tb));
  CRISPY_HIP(ctx, rocprim::inclusive_scan(ctx->temp.p, tb, f64, f64, (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
  rocprim::counting_iterator<uint32_t> it(0);
  tb = 0;
  CRISPY_HIP(ctx, rocprim::select(nullptr, tb, it, P.flags.as<uint8_t>(), P.sel.as<uint32_t>(),
                                  P.nsel.as<unsigned int>(), (size_t)m, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tb));
  CRISPY_HIP(ctx, rocprim::select(ctx->temp.p, tb, it, P.flags.as<uint8_t>(), P.sel.as<uint32_t>(),
                                  P.nsel.as<unsigned int>(), (size_t)m, ctx->stream));
  pre_runs_kernel<<<pre_grid(m), 256, 0, ctx->stream>>>(P.sel.as<uint32_t>(), P.nsel.as<unsigned int>(),
                                                       sorted, f64, P.out_order.as<int32_t>(),
                                                       P.out_freq.as<int32_t>());
  CRISPY_LAUNCHED(ctx);
  timer_end(ctx, "pre_merge", ev, (double)m);
  unsigned int nr = 0;
  CRISPY_HIP(ctx, hipMemcpyAsync(&nr, P.nsel.p, sizeof nr, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nr > 0) {
    CRISPY_HIP(ctx, hipMemcpyAsync(order, P.out_order.p, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipMemcpyAsync(out_freq, P.out_freq.p, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  *count = (int32_t)nr;
  return CRISPY_OK;
}

int32_t crispy_preprocess_primer_filter(crispy_ctx *ctx, const char *bases, const int64_t *offsets,
                                        int32_t n, const char *primer, int32_t primer_len,
                                        int32_t max_mismatches, int32_t *freq) {
  if (!ctx || n < 0 || primer_len < 0 || (primer_len > 0 && !primer) ||
      (n > 0 && (!bases || !offsets || !freq)))
    return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  int32_t live = 0;
  for (int32_t r = 0; r < n; ++r) {
    if (offsets[r + 1] < offsets[r]) return set_err(ctx, CRISPY_ERR_ARG, "offsets decrease at read %d", r);
    live += freq[r] != 0;
  }
  if (live == 0) return 0;
  PreState &P = ctx->pre;
  const int64_t nb = offsets[n] - offsets[0];
  // lanes: one per read up to 64 Ki, fewer when the DP rows would pass 256 MiB
  const int64_t row_bytes = 3 * sizeof(int) * (2 * (int64_t)primer_len + 1);
  int64_t lanes = std::min<int64_t>(((int64_t)n + 255) / 256 * 256, 65536);
  while (lanes > 256 && lanes * row_bytes > (256ll << 20)) lanes -= 256;
  CRISPY_TRY(ensure(ctx, P.bases, (size_t)nb + 16));
  CRISPY_TRY(ensure(ctx, P.off, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, P.freq, sizeof(int32_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, P.primer, (size_t)primer_len + 16));
  CRISPY_TRY(ensure(ctx, P.scratch, (size_t)(lanes * row_bytes)));
  CRISPY_TRY(ensure(ctx, P.nsel, sizeof(unsigned int)));
  std::vector<int64_t> off0(offsets, offsets + n + 1);
  for (auto &o : off0) o -= offsets[0];
  CRISPY_HIP(ctx, hipMemcpyAsync(P.bases.p, bases + offsets[0], (size_t)nb, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(P.off.p, off0.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(P.freq.p, freq, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
  if (primer_len)
    CRISPY_HIP(ctx, hipMemcpyAsync(P.primer.p, primer, (size_t)primer_len, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemsetAsync(P.nsel.p, 0, sizeof(unsigned int), ctx->stream));
  hipEvent_t ev;
  timer_begin(ctx, "pre_primer", &ev);
  pre_primer_kernel<<<(unsigned)(lanes / 256), 256, 0, ctx->stream>>>(
      P.bases.as<uint8_t>(), P.off.as<int64_t>(), n, P.primer.as<uint8_t>(), primer_len, max_mismatches,
      P.freq.as<int32_t>(), P.scratch.as<int>(), P.nsel.as<unsigned int>());
  CRISPY_LAUNCHED(ctx);
  timer_end(ctx, "pre_primer", ev, (double)live);
  unsigned int d = 0;
  CRISPY_HIP(ctx, hipMemcpyAsync(freq, P.freq.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(&d, P.nsel.p, sizeof d, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return (int32_t)d;
}

}  // extern "C"
