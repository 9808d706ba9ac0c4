// Context, device buffers, timing, the read store and the self checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace crispy {

int set_err(crispy_ctx *ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

// Grow-only device buffer.
int ensure(crispy_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.bytes >= bytes) return CRISPY_OK;
  if (ctx->mem_limit > 0 && ctx->dev_bytes - (int64_t)b.bytes + (int64_t)bytes > ctx->mem_limit)
    return set_err(ctx, CRISPY_ERR_NOMEM, "growing a buffer to %zu bytes would take the context past its "
                   "device memory limit (%lld of %lld bytes held)", bytes, (long long)ctx->dev_bytes,
                   (long long)ctx->mem_limit);
  if (b.p) {
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CRISPY_HIP(ctx, hipFree(b.p));
    ctx->dev_bytes -= (int64_t)b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return set_err(ctx, CRISPY_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  ctx->dev_bytes += (int64_t)bytes;
  return CRISPY_OK;
}

int ensure_host(crispy_ctx *ctx, size_t bytes) {
  CRISPY_HIP(ctx, hipEventSynchronize(ctx->stage_ev));
  if (ctx->h_stage_bytes >= bytes) return CRISPY_OK;
  if (ctx->h_stage) {
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CRISPY_HIP(ctx, hipHostFree(ctx->h_stage));
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
  }
  bytes = std::max<size_t>(bytes, 64 << 10);
  if (hipHostMalloc(&ctx->h_stage, bytes, hipHostMallocDefault) != hipSuccess) {
    ctx->h_stage = nullptr;
    (void)hipGetLastError();
    return set_err(ctx, CRISPY_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
  }
  ctx->h_stage_bytes = bytes;
  return CRISPY_OK;
}

void release(crispy_ctx *ctx, DevBuf &b) {
  if (b.p) {
    (void)hipFree(b.p);
    ctx->dev_bytes -= (int64_t)b.bytes;
  }
  b.p = nullptr;
  b.bytes = 0;
}

static hipEvent_t get_event(crispy_ctx *ctx) {
  if (!ctx->pool.empty()) {
    hipEvent_t e = ctx->pool.back();
    ctx->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timer_begin(crispy_ctx *ctx, const char *, hipEvent_t *a, hipStream_t s) {
  *a = nullptr;
  if (!ctx->timing) return;
  *a = get_event(ctx);
  if (*a) (void)hipEventRecord(*a, s ? s : ctx->stream);
}

void timer_end(crispy_ctx *ctx, const char *name, hipEvent_t a, double units, hipStream_t s) {
  if (!ctx->timing || !a) return;
  hipEvent_t b = get_event(ctx);
  if (!b) return;
  (void)hipEventRecord(b, s ? s : ctx->stream);
  ctx->pending.push_back(PendingEvent{name, a, b, units});
}

// Work queued on `stream` so far happens before anything later queued on the
// workers.
int fork_workers(crispy_ctx *ctx) {
  CRISPY_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
  for (int w = 0; w < kNumWorkers; ++w) CRISPY_HIP(ctx, hipStreamWaitEvent(ctx->workers[w], ctx->fork_ev, 0));
  return CRISPY_OK;
}

// Everything queued on the workers happens before anything later queued on
// `stream`.
int join_workers(crispy_ctx *ctx) {
  for (int w = 0; w < kNumWorkers; ++w) {
    CRISPY_HIP(ctx, hipEventRecord(ctx->join_ev[w], ctx->workers[w]));
    CRISPY_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->join_ev[w], 0));
  }
  return CRISPY_OK;
}

static void drain_timers(crispy_ctx *ctx) {
  for (auto &p : ctx->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KStat &s = ctx->stats[p.name];
      s.ms += ms;
      s.launches += 1;
      s.units += p.units;
    }
    ctx->pool.push_back(p.a);
    ctx->pool.push_back(p.b);
  }
  ctx->pending.clear();
}

// One block per read: 2-bit codes (A0 G1 T2 C3, other bytes 0 as in
// kmerMain.cpp:137-152), 16 bases per word, and the all-ACGT flag.
__global__ __launch_bounds__(256) void pack_reads_kernel(const uint8_t *__restrict__ bases,
                                                         const int64_t *__restrict__ off,
                                                         const int64_t *__restrict__ poff,
                                                         uint32_t *__restrict__ pack,
                                                         uint8_t *__restrict__ clean) {
  const int r = blockIdx.x;
  const int64_t b0 = off[r];
  const int L = (int)(off[r + 1] - b0);
  const int64_t w0 = poff[r];
  const int nw = (int)(poff[r + 1] - w0);
  int bad = 0;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    uint32_t word = 0;
    const int p0 = w * 16;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int p = p0 + t;
      if (p < L) {
        const uint8_t c = bases[b0 + p];
        const uint32_t code = c == 'G' ? 1u : c == 'T' ? 2u : c == 'C' ? 3u : 0u;
        bad |= !(c == 'A' || c == 'C' || c == 'G' || c == 'T');
        word |= code << (2 * t);
      }
    }
    pack[w0 + w] = word;
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) clean[r] = bad ? 0 : 1;
}

__global__ void selftest_div_kernel(int maxv, float *out_kmer, float *out_nw) {
  const int64_t total = (int64_t)(maxv + 1) * (maxv + 1);
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(k / (maxv + 1));
    const int m = (int)(k % (maxv + 1));
    if (a == 0) {
      out_kmer[k] = 0.f;
      out_nw[k] = 0.f;
      continue;
    }
    // the same expressions as the kmer tile and NW epilogues
    out_kmer[k] = 1.0f - (float)m / (float)a;
    out_nw[k] = (float)m / (float)a;
  }
}

}  // namespace crispy

using namespace crispy;

extern "C" {

int crispy_ctx_create(int device, crispy_ctx **out) {
  if (!out) return CRISPY_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return CRISPY_ERR_HIP;
  }
  if (device < 0 || device >= ndev) return CRISPY_ERR_ARG;
  crispy_ctx *ctx = new crispy_ctx();
  ctx->device = device;
  bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreate(&ctx->stream) == hipSuccess &&
            hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&ctx->stage_ev, hipEventDisableTiming) == hipSuccess &&
            hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
            hipHostMalloc(&ctx->h_hist, 64, hipHostMallocDefault) == hipSuccess;
  for (int w = 0; ok && w < kNumWorkers; ++w)
    ok = hipStreamCreate(&ctx->workers[w]) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->join_ev[w], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    crispy_ctx_destroy(ctx);
    return CRISPY_ERR_HIP;
  }
  *out = ctx;
  return CRISPY_OK;
}

static void release_all(crispy_ctx *ctx) {
  release_graphs(ctx);
  DevBuf *bufs[] = {&ctx->km.bitmap,   &ctx->km.T,         &ctx->km.rcnt,
                    &ctx->km.roff,     &ctx->km.rlist,     &ctx->km.excess,    &ctx->km.loff,
                    &ctx->km.roffc,    &ctx->km.rlistc,
                    &ctx->km.lists,    &ctx->km.cand,      &ctx->km.keys,      &ctx->km.keys_alt,
                    &ctx->km.dists,    &ctx->km.dists_alt, &ctx->km.pairs,     &ctx->km.counter,
                    &ctx->nw.in_pairs, &ctx->nw.keys,      &ctx->nw.keys_alt,  &ctx->nw.idx,
                    &ctx->nw.order,    &ctx->nw.dist,      &ctx->nw.hist,      &ctx->nw.cells,
                    &ctx->nw.flags,    &ctx->nw.sel,       &ctx->nw.nsel,      &ctx->nw.out_pairs,
                    &ctx->nw.out_dists, &ctx->nw.scratch,  &ctx->nw.plan,      &ctx->nw.stat,
                    &ctx->temp,
                    &ctx->pre.bases,   &ctx->pre.off,      &ctx->pre.freq,     &ctx->pre.keys,
                    &ctx->pre.keys_alt, &ctx->pre.flags,   &ctx->pre.sel,      &ctx->pre.nsel,
                    &ctx->pre.csum,    &ctx->pre.out_order, &ctx->pre.out_freq, &ctx->pre.primer, &ctx->pre.scratch,
                    &ctx->km.cand_alt,  &ctx->nw.csr_rows,
                    &ctx->nw.csr_cols,  &ctx->km.hbitmap,   &ctx->km.hexcess,   &ctx->km.cand_n,
                    &ctx->km.seg,       &ctx->km.seg_stat,  &ctx->nw.sort_k0,   &ctx->nw.sort_k1,
                    &ctx->nw.sort_b0,   &ctx->nw.sort_b1,   &ctx->nw.sort_pairs, &ctx->nw.sort_dists};
  for (DevBuf *b : bufs) release(ctx, *b);
  for (auto &r : ctx->reads) {
    DevBuf *rb[] = {&r.bases, &r.off, &r.len, &r.pack, &r.poff, &r.clean};
    for (DevBuf *b : rb) release(ctx, *b);
  }
}

static void sync_all(crispy_ctx *ctx) {
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int w = 0; w < kNumWorkers; ++w)
    if (ctx->workers[w]) (void)hipStreamSynchronize(ctx->workers[w]);
}

void crispy_ctx_destroy(crispy_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  sync_all(ctx);
  drain_timers(ctx);
  for (auto e : ctx->pool) (void)hipEventDestroy(e);
  release_all(ctx);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  for (int w = 0; w < kNumWorkers; ++w) {
    if (ctx->workers[w]) (void)hipStreamDestroy(ctx->workers[w]);
    if (ctx->join_ev[w]) (void)hipEventDestroy(ctx->join_ev[w]);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->stage_ev) (void)hipEventDestroy(ctx->stage_ev);
  if (ctx->h_hist) (void)hipHostFree(ctx->h_hist);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int crispy_ctx_trim(crispy_ctx *ctx) {
  if (!ctx) return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  sync_all(ctx);
  release_all(ctx);
  for (auto &r : ctx->reads) {
    const int g = r.generation;
    r = ReadStore();
    r.generation = g + 1;
  }
  ctx->km = KmerState();
  ctx->nw = NwState();
  ctx->pre = PreState();
  return CRISPY_OK;
}

const char *crispy_ctx_error(const crispy_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int crispy_host_alloc(int64_t bytes, void **out) {
  if (!out || bytes < 0) return CRISPY_ERR_ARG;
  *out = nullptr;
  if (bytes == 0) return CRISPY_OK;
  return hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? CRISPY_OK : CRISPY_ERR_NOMEM;
}

int crispy_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
  return CRISPY_OK;
}

int crispy_ctx_sync(crispy_ctx *ctx) {
  if (!ctx) return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int64_t crispy_ctx_device_bytes(const crispy_ctx *ctx) { return ctx ? ctx->dev_bytes : 0; }

int crispy_ctx_set_mem_limit(crispy_ctx *ctx, int64_t bytes) {
  if (!ctx || bytes < 0) return CRISPY_ERR_ARG;
  ctx->mem_limit = bytes;
  return CRISPY_OK;
}

int64_t crispy_band_cells(int32_t len1, int32_t len2, int32_t band) {
  if (len1 <= 0 || len2 <= 0) return 0;
  if (band <= 1) return (int64_t)len1 * len2;
  return band_cells(len1, len2, (len1 + len2) / (2 * band));
}

int crispy_reads_upload(crispy_ctx *ctx, int slot, const char *bases, const int64_t *offsets,
                        int32_t n) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (slot != CRISPY_READS_KMER && slot != CRISPY_READS_NW)
    return set_err(ctx, CRISPY_ERR_ARG, "invalid read slot %d", slot);
  if (n < 0 || n > CRISPY_MAX_NUM_READS || (n > 0 && (!bases || !offsets)))
    return set_err(ctx, CRISPY_ERR_ARG, "unsupported numReads: %d", n);
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  ReadStore &rs = ctx->reads[slot];
  rs.n = n;
  rs.h_len.assign(n, 0);
  std::vector<int64_t> poff(n + 1, 0);
  int32_t maxlen = 0, minlen = CRISPY_MAX_READ_LEN;
  if (n > 0 && offsets[0] != 0) return set_err(ctx, CRISPY_ERR_ARG, "offsets[0] must be 0");
  for (int32_t r = 0; r < n; ++r) {
    const int64_t L = offsets[r + 1] - offsets[r];
    if (L < 1 || L > CRISPY_MAX_READ_LEN)
      return set_err(ctx, CRISPY_ERR_ARG, "unsupported read length %lld (read %d)", (long long)L, r);
    rs.h_len[r] = (int32_t)L;
    maxlen = std::max<int32_t>(maxlen, (int32_t)L);
    minlen = std::min<int32_t>(minlen, (int32_t)L);
    poff[r + 1] = poff[r] + (L + 15) / 16 + 1;  // one zero pad word per read
  }
  rs.nbases = n ? offsets[n] : 0;
  rs.maxlen = maxlen;
  rs.minlen = n ? minlen : 0;
  const int64_t nwords = poff[n];
  CRISPY_TRY(ensure(ctx, rs.bases, (size_t)rs.nbases + 16));
  CRISPY_TRY(ensure(ctx, rs.off, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, rs.len, sizeof(int32_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, rs.poff, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, rs.pack, sizeof(uint32_t) * (nwords + 4)));
  CRISPY_TRY(ensure(ctx, rs.clean, (size_t)n + 1));
  if (n > 0) {
    CRISPY_HIP(ctx, hipMemcpyAsync(rs.bases.p, bases, rs.nbases, hipMemcpyHostToDevice, ctx->stream));
    CRISPY_HIP(ctx, hipMemcpyAsync(rs.off.p, offsets, sizeof(int64_t) * (n + 1),
                                   hipMemcpyHostToDevice, ctx->stream));
    CRISPY_HIP(ctx, hipMemcpyAsync(rs.len.p, rs.h_len.data(), sizeof(int32_t) * n,
                                   hipMemcpyHostToDevice, ctx->stream));
    CRISPY_HIP(ctx, hipMemcpyAsync(rs.poff.p, poff.data(), sizeof(int64_t) * (n + 1),
                                   hipMemcpyHostToDevice, ctx->stream));
    pack_reads_kernel<<<n, 256, 0, ctx->stream>>>(rs.bases.as<uint8_t>(), rs.off.as<int64_t>(),
                                                  rs.poff.as<int64_t>(), rs.pack.as<uint32_t>(),
                                                  rs.clean.as<uint8_t>());
    CRISPY_LAUNCHED(ctx);
    rs.h_clean.assign(n, 0);
    CRISPY_HIP(ctx, hipMemcpyAsync(rs.h_clean.data(), rs.clean.p, n, hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  rs.all_clean = 1;
  for (int32_t r = 0; r < n; ++r) rs.all_clean &= rs.h_clean[r];
  rs.generation += 1;
  if (slot == CRISPY_READS_KMER) ctx->km.valid = false;
  ctx->nw.valid = false;
  return CRISPY_OK;
}

int crispy_timing_enable(crispy_ctx *ctx, int enable) {
  if (!ctx) return CRISPY_ERR_ARG;
  ctx->timing = enable != 0;
  return CRISPY_OK;
}

int crispy_timing_reset(crispy_ctx *ctx) {
  if (!ctx) return CRISPY_ERR_ARG;
  (void)hipStreamSynchronize(ctx->stream);
  drain_timers(ctx);
  ctx->stats.clear();
  return nw_timer_reset(ctx);
}

int crispy_timing_query(crispy_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches,
                        double *work_units) {
  if (!ctx || !kernel) return CRISPY_ERR_ARG;
  drain_timers(ctx);
  auto it = ctx->stats.find(kernel);
  const KStat s = it == ctx->stats.end() ? KStat{} : it->second;
  if (total_ms) *total_ms = s.ms;
  if (launches) *launches = s.launches;
  double units = s.units;
  CRISPY_TRY(nw_timer_units(ctx, kernel, &units));
  if (work_units) *work_units = units;
  return CRISPY_OK;
}

int crispy_selftest_division(crispy_ctx *ctx, int32_t maxv, float *out_kmer, float *out_nw) {
  if (!ctx || maxv < 1 || maxv > 8192 || !out_kmer || !out_nw) return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  const size_t total = (size_t)(maxv + 1) * (maxv + 1);
  DevBuf a, b;
  CRISPY_TRY(ensure(ctx, a, total * sizeof(float)));
  int rc = ensure(ctx, b, total * sizeof(float));
  if (rc != CRISPY_OK) {
    release(ctx, a);
    return rc;
  }
  selftest_div_kernel<<<1024, 256, 0, ctx->stream>>>(maxv, a.as<float>(), b.as<float>());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out_kmer, a.p, total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out_nw, b.p, total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release(ctx, a);
  release(ctx, b);
  if (e != hipSuccess) return set_err(ctx, CRISPY_ERR_HIP, "selftest: %s", hipGetErrorString(e));
  return CRISPY_OK;
}

}  // extern "C"
