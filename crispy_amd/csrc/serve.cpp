// Served form: a job queue in front of `slots` worker threads, each owning a
// context (its own streams and resident device buffers).  A job (one read
// set through kmerDist -> genDist) is admitted when a slot is free and its
// device-memory estimate fits the HBM budget next to the running jobs; jobs
// admitted while others run overlap on the GPU.  Results are kept on the host
// until fetched.  Host code only: it drives the library through its C ABI.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "crispy_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

// Page-locked host block (crispy_host_alloc), move-only, freed with its
// owner: a device-to-host copy into it runs at the bus's rate and touches no
// fresh pages.
struct PinnedBuf {
  void *p = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  PinnedBuf(PinnedBuf &&o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr, o.cap = 0; }
  PinnedBuf &operator=(PinnedBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p, cap = o.cap;
      o.p = nullptr, o.cap = 0;
    }
    return *this;
  }
  ~PinnedBuf() { release(); }
  void release() {
    if (p) crispy_host_free(p);
    p = nullptr, cap = 0;
  }
};

struct Job {
  int64_t id = 0;
  // the reads (bases in page-locked blocks from the server's pool, so the
  // uploads run at the bus's rate), until the job has run
  PinnedBuf kbases, nbases;
  std::vector<int64_t> koff, noff;
  const char *kb() const { return kbases.p ? static_cast<const char *>(kbases.p) : ""; }
  const char *nb() const { return nbases.p ? static_cast<const char *>(nbases.p) : ""; }
  bool same_reads = true;
  int32_t n = 0;
  int K = 6, band = 10;
  float kthr = 0.f, nthr = 0.f;
  int64_t est = 0;
  int64_t held = 0;  // device bytes the slot's context held after the job
  // outgrew its charge while other jobs ran: queued again at the head,
  // charged the whole budget, so it runs alone
  bool solo = false;
  bool done = false;
  bool consumed = false;  // fetched: results handed out, the job forgotten
  int status = CRISPY_OK;
  std::string error;
  int64_t kmer_pairs = 0;
  // the result: `count` pairs (int32 i, j) then `count` distances, in
  // page-locked memory from the server's pool
  PinnedBuf res;
  int64_t count = 0;
  const int32_t *pairs() const { return static_cast<const int32_t *>(res.p); }
  const float *dists() const { return reinterpret_cast<const float *>(static_cast<const char *>(res.p) + 8 * count); }
  std::vector<int64_t> row_ptr;  // CSR rows of the result (built on the device)
  Clock::time_point t_submit, t_start, t_end;
};

}  // namespace

struct crispy_server {
  int device = 0;
  int64_t budget = 0;
  // Device bytes charged to the slots: an idle slot is charged what its
  // context holds, a busy one max(held, job estimate).  `reserved` is the sum.
  int64_t reserved = 0;
  int64_t peak_reserved = 0;
  std::vector<int64_t> charge;
  int running = 0;
  bool stop = false;
  int64_t next_id = 0;
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  // Jobs are shared with the threads waiting on them, so a fetch that forgets
  // a job cannot free it under another waiter.
  std::map<int64_t, std::shared_ptr<Job>> jobs;
  // Error messages handed out by crispy_server_job_error stay valid until the
  // server is destroyed.
  std::map<int64_t, std::string> errors;
  std::vector<crispy_ctx *> ctxs;
  std::vector<std::thread> workers;
  // Result blocks of fetched jobs, reused by the next jobs (pool_mu).
  std::mutex pool_mu;
  std::vector<PinnedBuf> pool;

  // A block of at least `bytes`: the smallest free one that fits, else the
  // largest grown to fit, else a new one.
  int get_block(size_t bytes, PinnedBuf &out) {
    if (bytes == 0) return CRISPY_OK;
    {
      std::lock_guard<std::mutex> lk(pool_mu);
      int pick = -1, largest = -1;
      for (int k = 0; k < (int)pool.size(); ++k) {
        if (pool[k].cap >= bytes && (pick < 0 || pool[k].cap < pool[pick].cap)) pick = k;
        if (largest < 0 || pool[k].cap > pool[largest].cap) largest = k;
      }
      if (pick < 0) pick = largest;
      if (pick >= 0) {
        out = std::move(pool[pick]);
        pool.erase(pool.begin() + pick);
      }
    }
    if (out.cap >= bytes) return CRISPY_OK;
    out.release();
    void *p = nullptr;
    const int rc = crispy_host_alloc((int64_t)bytes, &p);
    if (rc != CRISPY_OK) return rc;
    out.p = p;
    out.cap = bytes;
    return CRISPY_OK;
  }
  void put_block(PinnedBuf &&b) {
    if (!b.p) return;
    std::lock_guard<std::mutex> lk(pool_mu);
    pool.push_back(std::move(b));
  }
};

namespace {

void run_job(crispy_server *srv, crispy_ctx *ctx, Job &job) {
  int rc = crispy_reads_upload(ctx, CRISPY_READS_KMER, job.kb(), job.koff.data(), job.n);
  if (rc == CRISPY_OK) {
    rc = job.same_reads
             ? crispy_reads_upload(ctx, CRISPY_READS_NW, job.kb(), job.koff.data(), job.n)
             : crispy_reads_upload(ctx, CRISPY_READS_NW, job.nb(), job.noff.data(), job.n);
  }
  if (rc == CRISPY_OK) rc = crispy_kmer_run(ctx, job.K, job.kthr, 0, -1, &job.kmer_pairs);
  int64_t c = 0;
  if (rc == CRISPY_OK) rc = crispy_nw_run_on_kmer(ctx, job.band, job.nthr, &c);
  if (rc == CRISPY_OK) rc = srv->get_block((size_t)c * 12, job.res);
  if (rc == CRISPY_OK) {
    job.count = c;
    rc = crispy_nw_fetch(ctx, 0, c, static_cast<int32_t *>(job.res.p),
                         reinterpret_cast<float *>(static_cast<char *>(job.res.p) + 8 * c));
  }
  if (rc == CRISPY_OK) {
    job.row_ptr.resize((size_t)job.n + 1);
    rc = crispy_nw_csr(ctx, job.n, job.row_ptr.data(), nullptr, nullptr);
  }
  job.status = rc;
  if (rc != CRISPY_OK) job.error = crispy_ctx_error(ctx);
}

void worker_loop(crispy_server *srv, int slot) {
  crispy_ctx *ctx = srv->ctxs[slot];
  for (;;) {
    std::shared_ptr<Job> job;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      for (;;) {
        if (srv->stop && srv->queue.empty()) return;
        if (!srv->queue.empty()) {
          // FIFO: the head job starts when it fits the budget next to what the
          // other slots hold
          const int64_t held = srv->charge[slot];
          const int64_t need = std::max(held, srv->queue.front()->est);
          if (srv->reserved - held + need <= srv->budget) {
            job = srv->queue.front();
            srv->queue.pop_front();
            srv->reserved += need - held;
            srv->peak_reserved = std::max(srv->peak_reserved, srv->reserved);
            srv->charge[slot] = need;
            srv->running += 1;
            if (!job->solo) job->t_start = Clock::now();
            // the job's buffers may not outgrow what it is charged
            crispy_ctx_set_mem_limit(ctx, need);
            break;
          }
          if (held > 0) {
            // give this idle slot's buffers back, then look again
            lk.unlock();
            crispy_ctx_trim(ctx);
            lk.lock();
            srv->reserved -= held - crispy_ctx_device_bytes(ctx);
            srv->charge[slot] = crispy_ctx_device_bytes(ctx);
            srv->cv_work.notify_all();
            continue;
          }
        }
        srv->cv_work.wait(lk);
      }
    }
    run_job(srv, ctx, *job);
    (void)crispy_ctx_sync(ctx);  // no copy from the job's host blocks is in flight when they go back
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      const int64_t held = crispy_ctx_device_bytes(ctx);
      srv->reserved += held - srv->charge[slot];
      srv->peak_reserved = std::max(srv->peak_reserved, srv->reserved);
      srv->charge[slot] = held;
      srv->running -= 1;
      if (job->status == CRISPY_ERR_NOMEM && !job->solo) {
        // it needed more than its estimate: run it again alone, with the
        // whole budget, before the jobs behind it
        job->solo = true;
        job->est = srv->budget;
        job->status = CRISPY_OK;
        job->error.clear();
        srv->put_block(std::move(job->res));
        job->count = 0;
        job->row_ptr.clear();
        srv->queue.push_front(job);
        job.reset();
      } else {
        job->t_end = Clock::now();
        job->done = true;
        job->held = held;
        // the input copies are no longer needed
        srv->put_block(std::move(job->kbases));
        srv->put_block(std::move(job->nbases));
      }
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
}

// Widest half band kB = (L1 + L2) / (2 band) the register and sliced NW
// kernels take (64 lanes x 24 cells per pair); wider pairs use the generic one.
constexpr int64_t kSlicedMaxHalfBand = (64 * 24 - 1) / 2;

int64_t words_per_read(int K) {
  const int KP = K <= 6 ? K : 6;
  const int64_t w = (int64_t(1) << (2 * KP)) / 32;
  return w < 16 ? 16 : w;
}

}  // namespace

extern "C" {

// Device bytes a context holds for a job: both read slots (bytes, 2-bit
// packing, per-read arrays), the k-mer profiles (bitmap, repeat lists; the
// K-mer lists and candidate buffer for K > 6), the k-mer result buffers at
// their first size (the pairs in the tile range, at most 4 Mi entries), the
// sort/NW/result buffers for up to 256 passing k-mer pairs per read, and the
// per-lane DP scratch of the generic NW kernel when some pair needs it (full
// alignment, a non-ACGT byte, or a band wider than the sliced kernel's).  A job
// with more passing pairs grows its buffers past the estimate.
int64_t crispy_server_job_bytes(const char *bases, const int64_t *offsets, int32_t n, int K,
                                int band) {
  if (n <= 0 || !bases || !offsets) return 0;
  const int64_t nbases = offsets[n] - offsets[0];
  int64_t maxlen = 0;
  for (int32_t r = 0; r < n; ++r) maxlen = std::max(maxlen, offsets[r + 1] - offsets[r]);
  bool clean = true;
  for (int64_t q = offsets[0]; q < offsets[n] && clean; ++q) {
    const char c = bases[q];
    clean = c == 'A' || c == 'C' || c == 'G' || c == 'T';
  }
  const int64_t nb = (n + 127) / 128;
  const int64_t tiles = nb * (nb + 1) / 2;
  const int64_t guess = std::min<int64_t>(int64_t(1) << 22, std::max<int64_t>(4096, tiles * 128 * 128));
  const int64_t all_pairs = (int64_t)n * (n - 1) / 2;
  const int64_t pairs = std::min<int64_t>(all_pairs, 256 * (int64_t)n);
  const int64_t reads = 2 * (nbases + nbases / 4 + 48 * (int64_t)n);
  int64_t prof = (int64_t)n * (4 * words_per_read(K) + 48) + 6 * nbases;
  // K > 6: K-mer lists, the hashed k = 6 bitmaps (at most 2,048 bits) and their
  // excess, candidates + their refined copy, the 256 x 4 KB segment counters
  if (K > 6) prof += 2 * nbases + 16 * (int64_t)n + 260 * (int64_t)n + 16 * (std::max(guess, pairs) + 256) + (1 << 20);
  int64_t est = reads + prof + 12 * guess + 72 * pairs + (int64_t)(8 << 20);
  const bool generic = band <= 1 || !clean || (band > 0 && maxlen / band > kSlicedMaxHalfBand);
  if (generic) est += 8 * std::min<int64_t>(pairs, 32768) * (maxlen + 1);
  return est;
}

int crispy_server_create(int device, int slots, int64_t hbm_budget_bytes, crispy_server **out) {
  if (!out || slots < 1 || slots > 64 || hbm_budget_bytes <= 0) return CRISPY_ERR_ARG;
  *out = nullptr;
  auto *srv = new crispy_server();
  srv->device = device;
  srv->budget = hbm_budget_bytes;
  srv->charge.assign(slots, 0);
  for (int s = 0; s < slots; ++s) {
    crispy_ctx *c = nullptr;
    const int rc = crispy_ctx_create(device, &c);
    if (rc != CRISPY_OK) {
      for (auto *x : srv->ctxs) crispy_ctx_destroy(x);
      delete srv;
      return rc;
    }
    srv->ctxs.push_back(c);
  }
  for (int s = 0; s < slots; ++s) srv->workers.emplace_back(worker_loop, srv, s);
  *out = srv;
  return CRISPY_OK;
}

void crispy_server_destroy(crispy_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto &t : srv->workers) t.join();
  for (auto *c : srv->ctxs) crispy_ctx_destroy(c);
  delete srv;
}

int64_t crispy_server_submit(crispy_server *srv, const char *kmer_bases, const int64_t *kmer_offsets,
                             const char *nw_bases, const int64_t *nw_offsets, int32_t n, int K,
                             float kmer_threshold, int band, float nw_threshold) {
  if (!srv || n < 0 || n > CRISPY_MAX_NUM_READS || (n > 0 && (!kmer_bases || !kmer_offsets)) ||
      K < 1 || K > 8)
    return CRISPY_ERR_ARG;
  auto job = std::make_shared<Job>();
  job->n = n;
  job->K = K;
  job->kthr = kmer_threshold;
  job->band = band;
  job->nthr = nw_threshold;
  job->same_reads = nw_bases == nullptr;
  if (n == 0) {
    job->koff.assign(1, 0);
    job->noff.assign(1, 0);
  } else {
    if (kmer_offsets[0] != 0 || (!job->same_reads && (!nw_offsets || nw_offsets[0] != 0)))
      return CRISPY_ERR_ARG;
    job->koff.assign(kmer_offsets, kmer_offsets + n + 1);
    const size_t kn = (size_t)job->koff[n];
    if (srv->get_block(kn, job->kbases) != CRISPY_OK) return CRISPY_ERR_NOMEM;
    if (kn) std::memcpy(job->kbases.p, kmer_bases, kn);
    if (!job->same_reads) {
      job->noff.assign(nw_offsets, nw_offsets + n + 1);
      const size_t nn = (size_t)job->noff[n];
      if (srv->get_block(nn, job->nbases) != CRISPY_OK) return CRISPY_ERR_NOMEM;
      if (nn) std::memcpy(job->nbases.p, nw_bases, nn);
    }
  }
  job->est = crispy_server_job_bytes(job->kb(), job->koff.data(), n, K, band);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (job->est > srv->budget) return CRISPY_ERR_NOMEM;  // would never be admitted
  job->id = srv->next_id++;
  job->t_submit = Clock::now();
  srv->jobs[job->id] = job;
  srv->queue.push_back(job);
  srv->cv_work.notify_all();
  return job->id;
}

int crispy_server_wait(crispy_server *srv, int64_t job_id, int64_t *kmer_pairs, int64_t *nw_pairs,
                       double *queue_ms, double *run_ms) {
  if (!srv) return CRISPY_ERR_ARG;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end()) return CRISPY_ERR_ARG;
  const std::shared_ptr<Job> job = it->second;
  srv->cv_done.wait(lk, [&] { return job->done; });
  if (job->consumed) return CRISPY_ERR_STATE;  // fetched meanwhile by another thread
  if (kmer_pairs) *kmer_pairs = job->kmer_pairs;
  if (nw_pairs) *nw_pairs = job->count;
  using ms = std::chrono::duration<double, std::milli>;
  if (queue_ms) *queue_ms = ms(job->t_start - job->t_submit).count();
  if (run_ms) *run_ms = ms(job->t_end - job->t_start).count();
  return job->status;
}

// Submission, start and end of a finished job in ms since the server was
// created.
int crispy_server_job_times(crispy_server *srv, int64_t job_id, double *submit_ms, double *start_ms,
                            double *end_ms, int64_t *est_bytes, int64_t *held_bytes) {
  if (!srv) return CRISPY_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end() || !it->second->done) return CRISPY_ERR_STATE;
  using ms = std::chrono::duration<double, std::milli>;
  const auto t0 = srv->t0;
  if (submit_ms) *submit_ms = ms(it->second->t_submit - t0).count();
  if (start_ms) *start_ms = ms(it->second->t_start - t0).count();
  if (end_ms) *end_ms = ms(it->second->t_end - t0).count();
  if (est_bytes) *est_bytes = it->second->est;
  if (held_bytes) *held_bytes = it->second->held;
  return CRISPY_OK;
}

const char *crispy_server_job_error(crispy_server *srv, int64_t job_id) {
  if (!srv) return "";
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end() || !it->second->done) return "";
  std::string &kept = srv->errors[job_id];
  kept = it->second->error;
  return kept.c_str();
}

int crispy_server_fetch(crispy_server *srv, int64_t job_id, int32_t *pairs, float *dists) {
  if (!srv) return CRISPY_ERR_ARG;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end()) return CRISPY_ERR_ARG;
  const std::shared_ptr<Job> job = it->second;
  srv->cv_done.wait(lk, [&] { return job->done; });
  if (job->consumed) return CRISPY_ERR_STATE;  // fetched meanwhile by another thread
  const int status = job->status;
  if (status == CRISPY_OK && job->count > 0) {
    if (pairs) std::memcpy(pairs, job->pairs(), (size_t)job->count * 2 * sizeof(int32_t));
    if (dists) std::memcpy(dists, job->dists(), (size_t)job->count * sizeof(float));
  }
  if (!job->error.empty()) srv->errors[job_id] = job->error;
  job->consumed = true;
  srv->put_block(std::move(job->res));
  srv->jobs.erase(job_id);
  srv->cv_done.notify_all();
  return status;
}

int crispy_server_fetch_csr(crispy_server *srv, int64_t job_id, int32_t n_rows, int64_t *row_ptr,
                            int32_t *cols, float *vals) {
  if (!srv || !row_ptr) return CRISPY_ERR_ARG;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end()) return CRISPY_ERR_ARG;
  const std::shared_ptr<Job> job = it->second;
  // row_ptr holds n_rows + 1 entries: it must be the job's read count
  if (n_rows != job->n) return CRISPY_ERR_ARG;
  srv->cv_done.wait(lk, [&] { return job->done; });
  if (job->consumed) return CRISPY_ERR_STATE;  // fetched meanwhile by another thread
  const int status = job->status;
  if (status == CRISPY_OK) {
    std::memcpy(row_ptr, job->row_ptr.data(), job->row_ptr.size() * sizeof(int64_t));
    const size_t c = (size_t)job->count;
    if (cols)
      for (size_t k = 0; k < c; ++k) cols[k] = job->pairs()[2 * k + 1];
    if (vals && c) std::memcpy(vals, job->dists(), c * sizeof(float));
  }
  if (!job->error.empty()) srv->errors[job_id] = job->error;
  job->consumed = true;
  srv->put_block(std::move(job->res));
  srv->jobs.erase(job_id);
  srv->cv_done.notify_all();
  return status;
}

int crispy_server_job_solo(crispy_server *srv, int64_t job_id) {
  if (!srv) return CRISPY_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end() || it->second->consumed) return CRISPY_ERR_ARG;
  return it->second->solo ? 1 : 0;
}

int crispy_server_status(crispy_server *srv, int *running, int *queued, int64_t *reserved_bytes,
                         int64_t *peak_reserved_bytes) {
  if (!srv) return CRISPY_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (running) *running = srv->running;
  if (queued) *queued = (int)srv->queue.size();
  if (reserved_bytes) *reserved_bytes = srv->reserved;
  if (peak_reserved_bytes) *peak_reserved_bytes = srv->peak_reserved;
  return CRISPY_OK;
}

}  // extern "C"
