/*
 * libclast_hip — MI355X (gfx950) implementation of CLAST's device stages
 * behind a plain C ABI.  It replaces the reference's three device objects
 * and the result sort they feed:
 *
 *   CDeviceHashTable(setting, targetList, t_begin, t_end)   CDeviceHashTable.cuh:17-21
 *       -> clast_index_create / clast_index_destroy
 *   CDeviceSeqList_query(setting, &queryList, q_begin, q_end) CDeviceSeqList_query.cuh:14-18
 *       -> clast_query_create / clast_query_destroy
 *   CDeviceHitList(setting, hashTable, queryList, t_begin, 2*q_begin)
 *                                                           CDeviceHitList.cuh:23-28
 *       -> clast_search, then hitList.getResult(holder)     CDeviceHitList.cuh:29
 *       -> clast_hits_count / clast_hits_copy / clast_hits_destroy
 *   resultSorting(...) used by CHostResultHolder::fixResult utilResultSorting.cuh:28-37
 *       -> clast_result_sort
 *
 * Batched extension (no single counterpart in the reference): an index over
 * several target windows and a query set over several query windows are
 * searched in one launch chain by clast_search_batch; the output equals the
 * reference's per-window results appended in (target window, query window)
 * order and stable-sorted by (qID/2, score desc) as fixResult does
 * (CHostResultHolder.cu:75-128).
 *
 * Conventions: every call returns CLAST_OK (0) or an error code and never
 * throws; the message of the last failure on the calling thread is
 * available from clast_last_error().  Host buffers belong to the caller and
 * are not retained after a call returns; device memory belongs to the handle.
 *
 * Threads: each host thread has its own runtime on each device (a compute,
 * a side and a copy stream, and a search workspace that grows to the
 * largest search), so searches issued from different threads overlap on the
 * GPU.  Index and query handles may be searched from several threads at
 * once; searches of one query set that overlap in time must use the same
 * cut_off, total_db_size, K and lambda (a search stores the query's
 * per-strand score thresholds for those, under a lock of that query).
 * Create and destroy a handle on one thread, and not while it is searched.
 * A thread's runtimes go to a pool when the thread ends and are reused by
 * later threads; the pool lives until the process ends.
 */
#ifndef CLAST_HIP_H_
#define CLAST_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CLAST_OK = 0,
  CLAST_ERR_ARG = 1,    /* invalid argument or handle */
  CLAST_ERR_HIP = 2,    /* a HIP runtime call failed */
  CLAST_ERR_NOMEM = 3,  /* device allocation failed */
  CLAST_ERR_LIMIT = 4,  /* input exceeds a representable size (split it) */
  CLAST_ERR_IO = 5      /* an index file could not be written, read or is not one */
};

/* Stage-retention flag for clast_search_batch (stage-level parity checks). */
enum { CLAST_KEEP_STAGES = 1 };

/* Stages retained with CLAST_KEEP_STAGES, readable with clast_hits_stage_*. */
enum {
  CLAST_STAGE_SEEDS = 1,    /* seeds after the 3 boundary filters, generation order */
  CLAST_STAGE_SORTED = 2,   /* seeds after sortSeeds */
  CLAST_STAGE_HITS = 3,     /* seeds surviving isolate + duplicate (DP input) */
  CLAST_STAGE_ALIGNED = 4   /* DP output, one row per CLAST_STAGE_HITS entry */
};

/* CHostSetting's search parameters (CHostSetting.cuh:10-28).  total_db_size
 * is calcTotalDbSize(targets) (CHostSetting.cu:46-58); K and lambda are
 * GLOBAL_K / GLOBAL_LAMBDA (common.hpp:42-43). */
typedef struct clast_params {
  int32_t lmer;        /* -lMer   (<= 31) */
  int32_t stride;      /* -stride */
  int32_t cut_repeat;  /* -repeat (-1: unlimited) */
  int32_t width;       /* -width  */
  int32_t gap;         /* -gap    (2..16) */
  int32_t local;       /* -local  (0: global) */
  double cut_off;      /* -cutOff (-1: no cut off) */
  double total_db_size;
  double K;
  double lambda;
} clast_params;

/* One window of a host sequence list, exactly what
 * CHostSeqList::get{Base,Length,Gateway}Array(startID, endID) hands the
 * device (CHostSeqList.cu:15-70): the window's slot array (each sequence
 * followed by a copy of its first lMer-1 bases), per-sequence lengths and
 * per-sequence start slots rebased to 0.  id_begin is the list ID of the
 * window's first sequence (t_begin for targets, 2*q_begin for queries); the
 * per-slot ID and index arrays are implied and not passed. */
typedef struct clast_window {
  const char* bases;
  int64_t n_slots;
  const int32_t* lengths;
  const int32_t* gateway;
  int32_t n_seq;
  int32_t id_begin;
} clast_window;

typedef struct clast_index_s* clast_index;
typedef struct clast_query_s* clast_query;
typedef struct clast_hits_s* clast_hits;

/* cudaSetDevice(setting.getDeviceID()) (main.cu:30), and the calling
 * thread's runtime on that device.  Optional for every entry point but
 * clast_search_batch_noalloc: the first call on a thread without it uses
 * device 0.  Indexes and query sets may be shared between threads once
 * created (see Threads above). */
int clast_init(int32_t device);
const char* clast_last_error(void);
/* Library version string. */
const char* clast_version(void);

/* CDeviceHashTable ctor (CDeviceHashTable.cu:143-174): target k-mer index of
 * one target window. */
int clast_index_create(const clast_params* p, const clast_window* target, clast_index* out);
/* The same index over n_windows target windows at once (kept resident;
 * windows are numbered 0..n_windows-1 in the given order). */
int clast_index_create_multi(const clast_params* p, const clast_window* targets, int32_t n_windows,
                             clast_index* out);
/* Entries of the index (test introspection): per entry its k-mer code,
 * window, list ID and intra-sequence index, in key order. */
int64_t clast_index_entry_count(clast_index idx);
int clast_index_entries(clast_index idx, int64_t* key, int32_t* window, int32_t* tID, int32_t* tIdx);
int clast_index_destroy(clast_index idx);
/* Persisted index, no counterpart in the reference (which rebuilds the table
 * for every query chunk, main.cu:63-88, CHostSchedular.cu:70): for serving
 * start-up.  clast_index_save writes the target windows (class bytes,
 * sequence tables) and the index entries to a file; clast_index_load reads
 * one back on the current device (clast_init) and rebuilds the cell tables,
 * giving an index equal to the saved one.  A file that is not a complete
 * index of this format is rejected with CLAST_ERR_IO before any device
 * work. */
int clast_index_save(clast_index idx, const char* path);
int clast_index_load(const char* path, clast_index* out);

/* CDeviceSeqList_query ctor (CDeviceSeqList_query.cu:61-71): uploads one
 * query window ('+' and '-' strand of each query, IDs 2k and 2k+1). */
int clast_query_create(const clast_params* p, const clast_window* query, clast_query* out);
int clast_query_create_multi(const clast_params* p, const clast_window* queries, int32_t n_windows,
                             clast_query* out);
/* The reads as given, encoded on the device (replaces the host side of
 * CHostSeqList_query, CHostSeqList_query.cu:11-54 with addSequence,
 * utilAddSequence.cu:6-47, and the upload of CDeviceSeqList_query,
 * CDeviceSeqList_query.cuh:14-18): read k (bytes [offsets[k],
 * offsets[k+1]) of `bases`) becomes its '+' strand and its reverse
 * complement ('-': A<->T, C<->G, any other byte N), each followed by lMer-1
 * joint bases (the strand's first bases, N past a short strand's end).
 * Window w holds reads [win[2w], win[2w+1]); windows are consecutive and in
 * order; the list ID of window w's first strand is id_begin[w].  Equal to
 * clast_query_create_multi over the same strands built on the host. */
int clast_query_create_reads(const clast_params* p, const uint8_t* bases, const int64_t* offsets, const int32_t* win,
                             const int32_t* id_begin, int32_t n_win, clast_query* out);
int clast_query_destroy(clast_query q);

/* CDeviceHitList ctor (CDeviceHitList.cu:228-423): the whole seed -> filter
 * -> DP -> E-value -> sort -> dedup chain for one (target window, query
 * window).  idx and q must each hold one window whose id_begin equals
 * t_begin and q_begin. */
int clast_search(const clast_params* p, clast_index idx, clast_query q, int32_t t_begin, int32_t q_begin,
                 clast_hits* out);
/* Every (target window, query window) unit of idx x q in one launch chain,
 * output in fixResult order (see above).  flags: 0 or CLAST_KEEP_STAGES. */
int clast_search_batch(const clast_params* p, clast_index idx, clast_query q, int32_t flags,
                       clast_hits* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): clast_search_batch's launch chain and rows, with every
 * device buffer of the search carved from a fixed scratch array that the
 * library holds on each device it runs on (CLAST_NOALLOC_SCRATCH_BYTES, kept
 * until the process ends), and the rows written into the caller's host
 * arrays (`capacity` rows each; null arrays are skipped; twin / qwin as
 * clast_hits_copy_windows gives them) before the call returns.  It allocates
 * no device or pinned memory, so clast_init(device) must have run on the
 * calling thread.  Limits: a query set whose k-mer positions the scratch
 * cannot hold (16 B each) is refused before any device work, and a search
 * that needs more scratch fails, both with CLAST_ERR_LIMIT; more rows than
 * `capacity` fail with CLAST_ERR_LIMIT and *n_rows set to the rows needed.
 * Calls on one device share its scratch, so the library runs them one at a
 * time (a second call waits for the first); the scratch keeps the last
 * call's contents and every call clears what it needs cleared. */
#define CLAST_NOALLOC_SCRATCH_BYTES (256u << 20)
int clast_search_batch_noalloc(const clast_params* p, clast_index idx, clast_query q, int64_t capacity,
                               int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                               int32_t* qHL, int32_t* match, int32_t* score, double* evalue, int32_t* twin,
                               int32_t* qwin, int64_t* n_rows);

/* CDeviceHitList::getResult (CDeviceHitList.cu:425-451): the 9 result
 * arrays (tID and qID are list IDs).  Null pointers are skipped.  The rows
 * reach host memory asynchronously after a search returns (so the copy
 * overlaps the next search); clast_hits_copy and clast_hits_copy_windows
 * wait for them, clast_hits_count does not need to. */
int64_t clast_hits_count(clast_hits h);
/* No counterpart in the reference: blocks until the rows of h are in host
 * memory (what clast_hits_copy does before copying). */
int clast_hits_wait(clast_hits h);
int clast_hits_copy(clast_hits h, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                    int32_t* qHL, int32_t* match, int32_t* score, double* evalue);
/* Target / query window index of each result row (batched search). */
int clast_hits_copy_windows(clast_hits h, int32_t* twin, int32_t* qwin);
/* Retained stages (CLAST_KEEP_STAGES): seed stages fill tID..qIdx; the
 * ALIGNED stage fills all eight int arrays. */
int64_t clast_hits_stage_count(clast_hits h, int32_t stage);
int clast_hits_stage_copy(clast_hits h, int32_t stage, int32_t* tID, int32_t* tIdx, int32_t* qID,
                          int32_t* qIdx, int32_t* tHL, int32_t* qHL, int32_t* match, int32_t* score);
/* Device time of the search's stages in milliseconds, from hipEvents on the
 * library's compute stream: out[0] seed lookup kernel, [1] seed generation,
 * per-strand sort and isolate/duplicate filters (LDS passes, global-path
 * seed write, first-strand re-run), [2] global-path seed sort, [3]
 * global-path filters, [4] DP kernel, [5] E-value + filters + result order
 * + gather, [6] whole search (the copy to host memory runs on a second
 * stream and is not included); out[7] is the number of raw seeds (index
 * entries matched, before the corner filter) as a float. */
int clast_hits_stage_times(clast_hits h, float* out8);
/* No counterpart in the reference (work accounting for rooflines): out[0]
 * raw seeds (index entries matched, before the corner filter), [1] hits
 * entering the DP, [2] DP cells (band width x rows, both directions), [3]
 * result rows, [4] DP tasks (hit extensions, one per direction) run in packed
 * 16-bit lanes, [5] their cells (included in [2]), [6] DP tasks of at least
 * 1024 rows run one lane per band (none unless CLAST_DP_BAND=rest or all in
 * the environment), [7] their cells (included in [2]). */
int clast_hits_counts(clast_hits h, int64_t* out8);
/* No counterpart in the reference (rooflines): device times in milliseconds
 * of single kernels of the search, each from a hipEvent pair on the stream
 * it ran on: out[0] k_seed_count, [1] k_dp16 (DP in packed 16-bit lanes),
 * [2] k_dp (int32 DP; beside k_dp16 on a second stream in global mode),
 * [3] k_dp_band (one lane per band), [4] the DP stage, [5] seed generation
 * (as clast_hits_stage_times out[1]), [6] the result stage, [7] the whole
 * search.  0 for a kernel that did not run. */
int clast_hits_kernel_times(clast_hits h, float* out8);
/* No counterpart in the reference (rooflines): the DP cells (band width x
 * rows) and device times (ms, hipEvent pairs on each kernel's stream) of the
 * DP kernels: out[0] cells of k_dp16, [1] cells of the k_dp16 launch for
 * bands past a sequence end, [2] cells of k_dp, [3] cells of k_dp_band,
 * [4..8) the same four kernels' times; [8] tasks, [9] cells and [10] time of
 * the second pass of a two-pass DP stage, [11] 1 if the stage ran two passes
 * (the 16-bit kernels of the first then computed scores only).  The
 * past-the-end launch, k_dp, k_dp_band and the second pass run on a second
 * stream. */
int clast_hits_dp_kernels(clast_hits h, double* out12);
/* No counterpart in the reference (the DP autotuner): the DP launch
 * configuration the search ran with, out[0] schedule of the 16-bit kernel
 * (0 static, 1 persistent work-stealing wavefronts), [1] extensions run one
 * lane per band (0 none, 1 those the 16-bit kernel cannot take, 2 every one
 * over twice a SIMD's share), [2] radix digit bits of the task-key sort.
 * The configuration of each kind of search (gap, global / local, length
 * bucket of its longest extension: < 256, < 2048, < 8192 rows, longer; size
 * class: under 32768 DP tasks or not) is
 * picked by timed trials on the first such search in the process (rows are
 * the same under every configuration); CLAST_TUNE=0 turns trials off and
 * CLAST_DP_SCHED=static|persist, CLAST_DP_BAND=off|rest|all and
 * CLAST_RADIX_BITS=4..8 fix a knob.  clast_tuner_report writes the table
 * (one line per kind, with the trial times) into buf (cap bytes, NUL
 * terminated) and returns its full length. */
int clast_hits_dp_config(clast_hits h, int32_t* out3);
int64_t clast_tuner_report(char* buf, int64_t cap);
/* No counterpart in the reference (memory-budgeted admission): free and
 * total device memory of the calling thread's current device
 * (hipMemGetInfo), and the device bytes the calling thread's search
 * workspace holds there (it grows to the largest search and is kept). */
int clast_mem_info(int64_t* free_bytes, int64_t* total_bytes);
int clast_workspace_bytes(int64_t* bytes);
int clast_hits_destroy(clast_hits h);

/* resultSorting (utilResultSorting.cu:6-47): stable sort of n result rows by
 * (qID/2 ascending, score descending) on the device, in place. */
int clast_result_sort(int64_t n, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                      int32_t* qHL, int32_t* match, int32_t* score, double* evalue);

/* CHostResultHolder::fixResult + the row selection of printResult
 * (CHostResultHolder.cu:75-128, 131-243) on the device, for the n rows the
 * holder accumulated (its append order): order[k] receives the row index of
 * position k of fixResult's stable order (query pair ascending, score
 * descending), and printed[k] whether printResult writes that row: a row
 * equal to the last printed one in all ten printed fields is skipped, and
 * after the first row of a query label at most num_out rows of it print
 * (-1: no limit).  Labels are compared as integers: qlabel / tlabel give each
 * row's query / target label id (equal strings, equal ids; query id 0 must
 * be the empty label, which printResult's walk starts from), tstart each
 * row's target start offset ('@' line, added to tIdx as the row prints it).
 * *n_printed receives the number of printed rows. */
int clast_result_finalize(int64_t n, const int32_t* tIdx, const int32_t* qID, const int32_t* qIdx,
                          const int32_t* tHL, const int32_t* qHL, const int32_t* match, const int32_t* score,
                          const double* evalue, const int32_t* qlabel, const int32_t* tlabel, const int32_t* tstart,
                          int32_t num_out, int32_t* order, uint8_t* printed, int64_t* n_printed);

#ifdef __cplusplus
}
#endif

#endif /* CLAST_HIP_H_ */
