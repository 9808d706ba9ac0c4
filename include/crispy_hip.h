/*
 * crispy_hip.h -- C ABI of the MI355X (gfx950) kmerDist / genDist library.
 *
 * Plain C types only (no HIP, torch or C++ types), so the reference's C++
 * drivers, the drop-in executables bin/kmerDist and bin/genDist, Python
 * (ctypes) and any other FFI can bind it.
 *
 * Replaces these reference seams:
 *   computeKmerDist_CUDA   kmerMain.h:73, kmerCUDA.cu:14-126
 *   launchKmerKernel       kmerKernel.h:41, kmerKernel.cu:53-58
 *   computeGenDist_CUDA    genMain.h:77,  genCUDA.cu:83-270
 *   launchGenKernel_band   genKernel.h:48 (and _full, :46)
 *   writeVectorToFile_GPU  genCUDA.cu:17-81 (its thrust::sort_by_key)
 *   readFile / thresholds  kmerMain.cpp:298-397,43-53; genMain.cpp:184-277,42-47
 *   preprocess (input side) preprocess.cpp:74-280,322-392,562-601
 *
 * Errors: every entry point returns CRISPY_OK (0) or a negative code and
 * never exits the process (the reference calls exit(-1) from
 * checkCudaErrors, kmerKernel.h:17-25).  crispy_ctx_error() returns the
 * message of the last failure on that context.
 *
 * Threading: a context owns one HIP stream and its device buffers; calls on
 * one context must not overlap.  Different contexts are independent (the
 * reference's module-global textures, kmerKernel.cu:12 / genKernel.cu:11,
 * made it non-reentrant).
 */
#ifndef CRISPY_HIP_H
#define CRISPY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRISPY_OK 0
#define CRISPY_ERR_ARG (-1)     /* invalid argument or input outside the reference's limits */
#define CRISPY_ERR_HIP (-2)     /* HIP runtime error (includes "no device")                 */
#define CRISPY_ERR_NOMEM (-3)   /* device or host allocation failed                         */
#define CRISPY_ERR_STATE (-4)   /* call order (e.g. fetch before run)                       */
#define CRISPY_ERR_IO (-5)      /* file could not be opened / read / written                */

/* Limits of the reference (kmerMain.h:29-30, genMain.h:29-30). */
#define CRISPY_MAX_READ_LEN 2048
#define CRISPY_MAX_NUM_READS (1024 * 1024)

/* Read-store slots: kmerDist and genDist trim reads differently
 * (kmerMain.cpp:289-294 keeps [ATGCatgc], genMain.cpp:175-181 [nNATGCatgc]). */
#define CRISPY_READS_KMER 0
#define CRISPY_READS_NW 1

typedef struct crispy_ctx crispy_ctx;

/* ---------------------------------------------------------------- context */
/* Opens `device` (HIP ordinal) and creates the context's stream. */
int crispy_ctx_create(int device, crispy_ctx **out);
void crispy_ctx_destroy(crispy_ctx *ctx);
const char *crispy_ctx_error(const crispy_ctx *ctx);
/* Blocks until the context's stream is idle. */
int crispy_ctx_sync(crispy_ctx *ctx);
/* Device bytes currently held by the context. */
int64_t crispy_ctx_device_bytes(const crispy_ctx *ctx);
/* Caps the device bytes the context's buffers may grow to (0 = no cap): a
 * call that would grow them past it fails with CRISPY_ERR_NOMEM and leaves
 * them as they were.  No counterpart in the reference: the job server charges
 * each running job against its HBM budget with it. */
int crispy_ctx_set_mem_limit(crispy_ctx *ctx, int64_t bytes);
/* Frees every device buffer of the context (reads, profiles, results); the
 * streams stay.  Buffers otherwise only grow, so a long-lived context holds
 * the largest job it has run. */
int crispy_ctx_trim(crispy_ctx *ctx);
/* Page-locked host memory (hipHostMalloc) for the results' copies to the
 * host: a fetch into it runs at the bus's rate, into pageable memory it is
 * staged through a bounce buffer.  *out = NULL for bytes == 0.  Release with
 * crispy_host_free.  Replaces the reference's page-locked pair and distance
 * arrays (cudaMallocHost, genCUDA.cu:112-120). */
int crispy_host_alloc(int64_t bytes, void **out);
int crispy_host_free(void *ptr);

/* ------------------------------------------------------------- host side */
/* FASTA reader of the reference (kmerMain.cpp:298-397 for variant
 * CRISPY_READS_KMER, genMain.cpp:184-277 for CRISPY_READS_NW): reads become
 * concatenated bytes plus n+1 offsets, allocated with malloc and released
 * with crispy_free.  Returns the number of reads, or CRISPY_ERR_IO /
 * CRISPY_ERR_ARG (first line not '>').  *maxlen is the true maximum length. */
int32_t crispy_read_fasta(const char *path, int variant, char **bases, int64_t **offsets,
                          int32_t *maxlen);
void crispy_free(void *p);

/* Default thresholds used when -t is absent or outside [0, 1]
 * (kmerMain.cpp:43-48, genMain.cpp:42-47). */
float crispy_kmer_default_threshold(int32_t num_reads);
float crispy_nw_default_threshold(int32_t num_reads);
/* The filter predicate dist < thr || |dist - thr| < 1e-5 (kmerMain.cpp:180). */
int crispy_pass(float dist, float threshold);
/* In-band DP cells of one pair: #{1<=i<=L2, 1<=j<=L1 : |i-j| <= (L1+L2)/(2 band)},
 * or L1*L2 for band <= 1 (full alignment). */
int64_t crispy_band_cells(int32_t len1, int32_t len2, int32_t band);

/* ----------------------------------------------------------- read store */
/* Uploads n reads (raw bytes, offsets[n+1]) into `slot`, then packs them on
 * the device: 2-bit codes A0 G1 T2 C3 (any other byte codes 0, as in
 * kmerMain.cpp:137-152), 16 bases per 32-bit word, plus a per-read flag set
 * when every byte is one of 'A','C','G','T'.  Replaces the tupleSet /
 * texture upload (kmerCUDA.cu:35-72) and convertToBinary + texture upload
 * (genMain.cpp:280-321, genCUDA.cu:94-144).  Reads stay resident until the
 * slot is re-uploaded or the context is destroyed. */
int crispy_reads_upload(crispy_ctx *ctx, int slot, const char *bases, const int64_t *offsets,
                        int32_t n);

/* ------------------------------------------------------------- kmerDist */
/* All-pairs k-mer distance over the reads of slot CRISPY_READS_KMER:
 * dist(i,j) = 1 - matches/min(T_i, T_j), T = L-K+1, matches exactly as the
 * reference's merge over signed-sorted 16-bit codes (kmerCPU.cpp:29-43,
 * kmerMain.cpp:156-159) -- including its K = 8 ordering -- and keeps pairs
 * i < j that pass `threshold`.  The result stays on the device in (i, j)
 * order; *count is the number of passing pairs.
 * tile_begin/tile_end select a range of tile slots (0 and -1 = all), which
 * is how the work shards across devices: the upper triangle of 128x128-read
 * tiles in 16x16-tile super-tiles of 256 slots each (slots below the diagonal
 * or past the last read are empty).
 * Requires 1 <= K <= 8 and every read length >= K. */
int crispy_kmer_run(crispy_ctx *ctx, int K, float threshold, int64_t tile_begin,
                    int64_t tile_end, int64_t *count);
/* Number of tile slots for n reads (the shard index space). */
int64_t crispy_kmer_num_tiles(int32_t num_reads);
/* Where K <= 6 pairs that the tile's bound cannot reject get their exact
 * count: 0 = in the tile kernel itself; 1 = the tile writes them out as
 * candidates and a second kernel counts them (for K = 6 the candidate tile
 * runs on the 4,096 codes hashed into 1,280, 1,536 or 2,048 buckets, refined by the
 * exact bitmaps); 2 = as 1 without the hashing; -1 (the default) = form 1 from
 * 10^9 read pairs up, form 0 below.  K = 7, 8 always take candidates: hashed
 * unless the mode is 2.  Results do not depend on it.  No counterpart in the
 * reference: a tuning and test switch. */
int crispy_kmer_set_exact_pass(crispy_ctx *ctx, int mode);
/* Buckets of the hashed k = 6 candidate filter: 1280, 1536, 2048, or 0 for the
 * default (1280).  Fewer buckets: less tile work, more candidates for the
 * exact passes.  Results do not depend on it.  No counterpart in the
 * reference: a tuning and test switch. */
int crispy_kmer_set_hash_buckets(crispy_ctx *ctx, int buckets);
/* Times crispy_kmer_run over the whole tile range under each form of
 * crispy_kmer_set_exact_pass that applies to K (0, 1, 2 for K = 6; form 1
 * with each bucket count of crispy_kmer_set_hash_buckets for K >= 6),
 * `repeats` runs each after one warm-up run, keeps the fastest for the
 * context and leaves its result (the same under every form); *chosen = the
 * mode (crispy_kmer_tile_bits tells the bucket count).  No counterpart in the
 * reference (its tile shapes are fixed, kmerMain.cpp:50-53). */
int crispy_kmer_autotune(crispy_ctx *ctx, int K, float threshold, int repeats, int *chosen);
/* Presence bits per read that the last crispy_kmer_run's tile pass multiplied
 * (FP4 k-slots per pair): 4^min(K, 6) padded to at least 512, or the bucket
 * count when the k = 6 candidate filter ran on bucket-hashed bitmaps (K > 6,
 * or K = 6 with the exact counts in a pass of their own); 0 before any run.  No counterpart
 * in the reference (measurement). */
int crispy_kmer_tile_bits(crispy_ctx *ctx);
/* Copies result entries [first, first+count) to host: pairs as int32 (i, j),
 * dists as float32 (either pointer may be NULL). */
int crispy_kmer_fetch(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs,
                      float *dists);

/* -------------------------------------------------------------- genDist */
/* Banded (band >= 2) or full (band <= 1) Needleman-Wunsch distance of each
 * pair (i, j) of reads in slot CRISPY_READS_NW: read i on the column axis,
 * read j on the row axis, scoring and tie rules of computeGenDist_CPU_band
 * / _full (genCPU.cpp:290-464, :81-225); dist = ML/AL.  Keeps the pairs
 * passing `threshold` in input order; the result stays on the device.
 * pairs: host int32 (i, j) x npairs.  Replaces one chunk of
 * computeGenDist_CUDA (genCUDA.cu:156-226). */
int crispy_nw_run(crispy_ctx *ctx, const int32_t *pairs, int64_t npairs, int band,
                  float threshold, int64_t *count);
/* Fused form: the pairs are the device-resident result of the last
 * crispy_kmer_run (no host round trip). */
int crispy_nw_run_on_kmer(crispy_ctx *ctx, int band, float threshold, int64_t *count);
int crispy_nw_fetch(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs,
                    float *dists);
/* The result as a CSR matrix over n_rows reads (upper triangle, i < j): built
 * on the device from the (i, j)-ordered pairs of crispy_nw_run_on_kmer (or of
 * crispy_nw_run on (i, j)-ordered input); row_ptr[n_rows + 1] on the host,
 * cols / vals (count entries, either may be NULL) the column ids and
 * distances.  CRISPY_ERR_STATE when the input pairs were not (i, j)-ordered,
 * CRISPY_ERR_ARG when some pair's i is not below n_rows. */
int crispy_nw_csr(crispy_ctx *ctx, int32_t n_rows, int64_t *row_ptr, int32_t *cols, float *vals);
/* Band cells per lane of the lane-sliced NW kernel (kB > 32): 0 = chosen per
 * kB bucket by the library's cost model (the default), 24 or 32 = forced.
 * The distances do not depend on it.  No counterpart in the reference: a
 * tuning and test switch. */
int crispy_nw_set_slice_width(crispy_ctx *ctx, int cells);
/* Launch form of the context's kernel chains: 1 (the default) = the k-mer
 * profile pass, the k-mer tile pass, the NW DP phase (plan kernel + family
 * kernels) and the NW filter each replay a captured HIP graph, captured on
 * first use and updated (or captured again) when a buffer, a size or a
 * setting changes; 0 = the same kernels launched one by one.  Results do not
 * depend on it.  With graphs on, the context's stream must not itself be
 * under capture by the caller.  No counterpart in the reference: a
 * launch-form switch. */
int crispy_set_graphs(crispy_ctx *ctx, int enable);
/* Autotuner for that choice: on the context's current k-mer result, times the
 * NW DP (bucketing, sort and fused launches; fastest of `repeats` runs after a
 * warm-up) under the cost model and under each forced width, keeps the
 * fastest setting for the context and returns it in *chosen (0, 24 or 32).
 * The NW result of the context is left as it was.  No counterpart in the
 * reference. */
int crispy_nw_autotune(crispy_ctx *ctx, int band, int repeats, int *chosen);
/* Unfiltered distances for host pairs (every pair gets a distance). */
int crispy_nw_dists(crispy_ctx *ctx, const int32_t *pairs, int64_t npairs, int band,
                    float *dists);
/* crispy_nw_dists without any device allocation, for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It synchronizes its stream (it reads back a
 * pair-range check), so it is not meant for graph capture.  d_pairs (int32 x2)
 * and d_dists are device arrays; the reads are those already uploaded to the
 * context's NW slot.  Bucketing, sort and DP plan use one fixed scratch array
 * per device (about 40 MB, held until the process ends), so: at most 2^20
 * pairs per call, band >= 2, ACGT-only reads and kB <= 767 (CRISPY_ERR_ARG
 * otherwise; such pairs need the generic kernel's per-lane scratch).  Every
 * call on a device shares that array: calls that overlap on one device race.
 * Blocks until the distances are written. */
int crispy_nw_dists_noalloc(crispy_ctx *ctx, const int32_t *d_pairs, int64_t npairs, int band,
                            float *d_dists);
/* Sorts host arrays (pairs int32 x2, dists) ascending by (dist, i, j) on the
 * device; the per-file order of writeVectorToFile_GPU (genCUDA.cu:39-45). */
int crispy_sort_by_dist(crispy_ctx *ctx, int32_t *pairs, float *dists, int64_t n);
/* The reference's output-file split of the current NW result: its input pairs
 * read in chunks of `chunk_pairs` (loadPairs, genMain.cpp:129-139; 589,824 in
 * the reference), a file closed after a chunk once it holds >= `flush_pairs`
 * passing pairs and a last file for the rest (genCPU.cpp:482-505: 147,456 on
 * the CPU path; genCUDA.cu:228-242: 589,824 on the GPU path).  Writes the
 * result index where each file ends into file_ends[0..*nfiles); read from the
 * filter's device scan, no pair is copied.  CRISPY_ERR_ARG when more than
 * max_files files would be needed (*nfiles says how many). */
int crispy_nw_file_split(crispy_ctx *ctx, int64_t chunk_pairs, int64_t flush_pairs, int64_t *file_ends,
                         int32_t max_files, int32_t *nfiles);
/* Result entries [first, first + count) sorted by (dist, i, j) on the device
 * and copied to the host: one output file's contents in the order
 * writeVectorToFile sorts them (thrust::sort_by_key, genCUDA.cu:39-45,
 * genCPU.cpp:36-40; ties by (i, j)).  The result itself is left as it was. */
int crispy_nw_fetch_sorted(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs, float *dists);

/* --------------------------------------------------------------- timing */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream.  Query returns the summed duration (ms), the number of
 * launches and the algorithmic work units since the last reset:
 *   "kmer_profile" reads, "kmer_tile" read pairs evaluated (i < j),
 *   "nw_band_reg" and "nw_generic" in-band DP cells, "kmer_sort",
 *   "nw_bucket", "nw_filter" elements. */
int crispy_timing_enable(crispy_ctx *ctx, int enable);
int crispy_timing_reset(crispy_ctx *ctx);
int crispy_timing_query(crispy_ctx *ctx, const char *kernel, double *total_ms,
                        int64_t *launches, double *work_units);

/* ---------------------------------------------------------- self checks */
/* Evaluates on the device the two float formulas the kernels use, for every
 * integer operand pair: out_kmer[t*(maxv+1)+m] = 1 - m/t for 1<=t<=maxv,
 * 0<=m<=maxv, and out_nw[a*(maxv+1)+m] = m/a likewise.  Lets a test compare
 * the device rounding with the host's bit for bit. */
int crispy_selftest_division(crispy_ctx *ctx, int32_t maxv, float *out_kmer, float *out_nw);

/* ------------------------------------------------------------ input side */
/* The reference's preprocess stage on the path crispy.sh runs
 * (`preprocess -i <in>`, preprocess.cpp:686-726): read the records, merge
 * exact duplicates, write <out> (default <in>_Clean) and <out>.frq in
 * (length, strcmp) order.
 *
 * Reader of preprocess.cpp:74-240 (FASTA, or FASTQ with fastq != 0): records
 * as concatenated sequence bytes + n+1 offsets, labels (every whitespace
 * character removed, READ::initialize :31-39) + n+1 offsets, and a frequency
 * per record (1; 0 for an empty FASTQ sequence).  Buffers are malloc'ed and
 * released with crispy_free.  Returns n, CRISPY_ERR_IO, or CRISPY_ERR_ARG when a
 * FASTA file does not start with '>'. */
int32_t crispy_preprocess_read(const char *path, int fastq, char **bases, int64_t **offsets,
                               char **labels, int64_t **label_offsets, int32_t **freq);
/* The -s length filter (processReadLength, preprocess.cpp:322-392): sets the
 * frequency of outliers and of reads with 'n'/'N' to 0; returns the count. */
int32_t crispy_preprocess_length_filter(const int64_t *offsets, const char *bases, int32_t n,
                                        double sd_allowed, int32_t *freq);
/* The -p primer filter (processPrimer + SWLocalAlignment,
 * preprocess.cpp:394-560) on the device, after the -s filter: every read with
 * frequency > 0 whose first min(2 primer_len, L) bytes hold no local alignment
 * of the primer within max_mismatches (mismatches, gaps and unaligned primer
 * ends) gets frequency 0; returns the count, or a negative error code. */
int32_t crispy_preprocess_primer_filter(crispy_ctx *ctx, const char *bases, const int64_t *offsets,
                                        int32_t n, const char *primer, int32_t primer_len,
                                        int32_t max_mismatches, int32_t *freq);
/* mergeReads (preprocess.cpp:244-280) on the device: the reads with frequency
 * > 0 in stable (length, strcmp) order (READ::operator< :42-49), each run of
 * identical sequences collapsed onto its last read with the run's summed
 * frequency.  order[k] / out_freq[k] (capacity n) receive the k-th output
 * read's index and frequency; *count the number of output reads. */
int crispy_preprocess_order(crispy_ctx *ctx, const char *bases, const int64_t *offsets,
                            const int32_t *freq, int32_t n, int32_t *order, int32_t *out_freq,
                            int32_t *count);
/* writeToFile (preprocess.cpp:562-601): "> " + label without its first
 * character (FASTQ: the whole label) and the sequence per output read, and
 * "label freq" lines in <out_path>.frq. */
int crispy_preprocess_write(const char *out_path, int fastq, const char *bases,
                            const int64_t *offsets, const char *labels,
                            const int64_t *label_offsets, const int32_t *order,
                            const int32_t *freq, int32_t count);

/* --------------------------------------------------------------- serving */
/* No counterpart in the reference, which runs one sample per process
 * (crispy.sh:19-20: kmerDist, then genDist on its pairs).  A server takes a
 * stream of such jobs: each job is one read set through the k-mer filter and
 * the banded NW on the passing pairs (the fused crispy_kmer_run +
 * crispy_nw_run_on_kmer), with the same results as running it alone.
 * `slots` worker threads each own a context; the head of the FIFO queue starts
 * as soon as a slot is free and its device-byte estimate
 * (crispy_server_job_bytes) fits `hbm_budget_bytes` together with what the
 * other slots hold, so admitted jobs overlap on the device. */
typedef struct crispy_server crispy_server;

int crispy_server_create(int device, int slots, int64_t hbm_budget_bytes, crispy_server **out);
/* Finishes the queued jobs, then stops the workers and frees the contexts. */
void crispy_server_destroy(crispy_server *srv);
/* Device bytes admission charges for a job (an estimate of what its context
 * will hold, from the read count, bases, longest read, K and band). */
int64_t crispy_server_job_bytes(const char *bases, const int64_t *offsets, int32_t n, int K,
                                int band);
/* Copies the reads (as for crispy_reads_upload; nw_bases NULL = the same reads
 * for both stages) and queues the job.  Returns its id (>= 0) or an error code;
 * CRISPY_ERR_NOMEM when the job alone exceeds the budget. */
int64_t crispy_server_submit(crispy_server *srv, const char *kmer_bases, const int64_t *kmer_offsets,
                             const char *nw_bases, const int64_t *nw_offsets, int32_t n, int K,
                             float kmer_threshold, int band, float nw_threshold);
/* Blocks until the job has finished; returns its status and its pair counts,
 * time queued and time running (ms). */
int crispy_server_wait(crispy_server *srv, int64_t job_id, int64_t *kmer_pairs, int64_t *nw_pairs,
                       double *queue_ms, double *run_ms);
/* Submission, start and end times (ms since crispy_server_create) of a
 * finished job, its admission estimate and the bytes its slot held after it. */
int crispy_server_job_times(crispy_server *srv, int64_t job_id, double *submit_ms, double *start_ms,
                            double *end_ms, int64_t *est_bytes, int64_t *held_bytes);
/* Error message of a finished job that failed ("" otherwise); the string
 * stays valid until crispy_server_destroy. */
const char *crispy_server_job_error(crispy_server *srv, int64_t job_id);
/* Waits for the job, copies its NW-passing pairs (int32 x2, ascending (i, j))
 * and distances out and forgets the job.  A job is fetched once: a second
 * fetch (or a wait that a fetch overtook) returns CRISPY_ERR_STATE, or
 * CRISPY_ERR_ARG once the job is forgotten. */
int crispy_server_fetch(crispy_server *srv, int64_t job_id, int32_t *pairs, float *dists);
/* Like crispy_server_fetch, as CSR: row_ptr[n_rows + 1] (built on the
 * device), cols and vals (nw_pairs entries, either may be NULL); forgets the
 * job.  n_rows must be the job's read count (CRISPY_ERR_ARG otherwise, the
 * job is kept). */
int crispy_server_fetch_csr(crispy_server *srv, int64_t job_id, int32_t n_rows, int64_t *row_ptr,
                            int32_t *cols, float *vals);
/* 1 if the job outgrew its admission charge while other jobs ran and was run
 * again alone (charged the whole budget), 0 if not; CRISPY_ERR_ARG for an
 * unknown or fetched job. */
int crispy_server_job_solo(crispy_server *srv, int64_t job_id);
/* Jobs running and queued, device bytes charged to the slots now and at most
 * so far (idle slots are charged what their contexts hold and give it back
 * when the queue head would not fit otherwise). */
int crispy_server_status(crispy_server *srv, int *running, int *queued, int64_t *reserved_bytes,
                         int64_t *peak_reserved_bytes);

#ifdef __cplusplus
}
#endif

#endif /* CRISPY_HIP_H */
