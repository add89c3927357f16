/*
 * pdx_amd.h -- C-ABI of the MI355X (gfx950) PDX search path.
 *
 * This is the drop-in boundary under the reference's C++ index API (include/pdx/index.hpp):
 * the host-side classes (IPDXIndex, PDXIndex<Q>, PDXTreeIndex<Q>) keep their interface and
 * hand their clusters to this library, which keeps them resident in HBM and runs the search
 * (probe, ADSampling-pruned PDX scan, top-k) as hand-written HIP kernels.
 *
 * Conventions: every call returns a pdxg_status (0 = success) and never throws; the message of
 * the last failure on the calling thread is available from pdxg_last_error(). Host buffers are
 * only read during the call. Device memory belongs to the index / workspace handles.
 */
#ifndef PDX_AMD_H
#define PDX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PDXG_OK = 0,
    PDXG_ERR_INVALID = 1,     /* bad argument or inconsistent index description */
    PDXG_ERR_HIP = 2,         /* a HIP runtime call failed */
    PDXG_ERR_UNSUPPORTED = 3, /* valid input outside what the GPU path implements */
    PDXG_ERR_NOMEM = 4        /* device allocation failed */
} pdxg_status;

/* PDXIndexType (include/pdx/common.hpp:78). */
enum { PDXG_F32 = 0, PDXG_U8 = 1, PDXG_TREE_F32 = 2, PDXG_TREE_U8 = 3 };

typedef struct pdxg_index pdxg_index;
typedef struct pdxg_workspace pdxg_workspace;

/* One cluster as the reference holds it in memory (Cluster<Q>, include/pdx/cluster.hpp:17-105):
 * `data` is the max_capacity-strided PDX buffer (layout.hpp:21-87, cluster.hpp:341-384),
 * `ids` the row id of each used slot, `tombstones` the deleted slot indices. */
typedef struct {
    uint32_t num_embeddings; /* live vectors (Cluster::num_embeddings) */
    uint32_t used_capacity;  /* slots in use incl. tombstones (Cluster::used_capacity) */
    uint32_t max_capacity;   /* stride of the PDX buffer (Cluster::max_capacity) */
    const void* data;        /* max_capacity * d elements (float for F32 levels, uint8 for U8) */
    const uint32_t* ids;     /* used_capacity row ids */
    const uint32_t* tombstones;
    uint32_t num_tombstones;
} pdxg_cluster_desc;

/* One IVF level (IVF<Q>, include/pdx/ivf_wrapper.hpp:15-179). */
typedef struct {
    uint32_t num_clusters;
    const pdxg_cluster_desc* clusters;
    const float* centroids; /* num_clusters * d row-major, or NULL (two-level L1 after restore) */
} pdxg_level_desc;

/* A whole index: replaces the state PDXIndex<Q>/PDXTreeIndex<Q> keep after Restore/BuildIndex
 * (index.hpp:241-263, 516-543) plus the ADSampling pruner built from the stored matrix
 * (pruners/adsampling.hpp:60-78). */
typedef struct {
    uint32_t index_type; /* PDXG_F32 .. PDXG_TREE_U8 */
    uint32_t num_dimensions;
    uint32_t num_vertical_dimensions;   /* GetPDXDimensionSplit, common.hpp:133-159 */
    uint32_t num_horizontal_dimensions;
    uint32_t is_normalized;             /* IVF::is_normalized */
    float quantization_base;            /* U8 levels (ivf_wrapper.hpp:34-37) */
    float quantization_scale;
    const float* rotation; /* d x d row-major (y = x . M) */
    pdxg_level_desc l1;    /* data clusters */
    pdxg_level_desc l0;    /* meso clusters over the L1 centroids (tree types only, F32) */
} pdxg_index_desc;

/* Upload an index (IVF<Q>::Load + ADSamplingPruner ctor). Replaces nothing in the reference
 * one-to-one; it is the residency step behind PDXIndex::Restore / BuildIndex. */
pdxg_status pdxg_index_create(int device, const pdxg_index_desc* desc, pdxg_index** out);
pdxg_status pdxg_index_destroy(pdxg_index* index);

/* Re-upload one cluster after Append/Delete/compaction on the host
 * (Cluster::AppendEmbedding / DeleteEmbedding / CompactCluster, cluster.hpp:121-154,260-294).
 * level: 1 = data clusters, 0 = meso clusters. max_capacity must be unchanged. The call waits for
 * searches still being planned on other threads (pdxg_search_device, servers) and for every kernel
 * in flight on the device before it writes; searches launched after it see the new cluster.
 * Captured graphs keep the sizes of their capture and must be re-created. */
pdxg_status pdxg_index_update_cluster(pdxg_index* index, uint32_t level, uint32_t cluster_id,
                                      const pdxg_cluster_desc* cluster);

/* Device bytes held by the index. */
size_t pdxg_index_device_bytes(const pdxg_index* index);

/* Search arguments. For pdxg_search every pointer is a HOST pointer; for pdxg_search_device every
 * pointer is a DEVICE pointer. k / nprobe may be NULL, then k_all / nprobe_all apply to every query.
 * nprobe == 0 means "all clusters" (searcher.hpp:614-616, index.hpp:547-550). */
typedef struct {
    uint32_t num_queries;
    const float* queries;          /* num_queries * d */
    uint32_t queries_transformed;  /* 1: already normalised + rotated (searcher.hpp:596-605) */
    const uint32_t* k;
    const uint32_t* nprobe;
    uint32_t k_all;
    uint32_t nprobe_all;
    uint32_t k_max;                /* upper bound of every k (sizes on-chip top-k buffers) */
    uint32_t nprobe_max;           /* upper bound of every nprobe (0 = may be "all") */
    uint32_t out_stride;           /* row stride of the outputs, >= k_max */
    uint32_t* out_ids;             /* num_queries * out_stride */
    float* out_distances;          /* num_queries * out_stride */
    uint32_t* out_int_distances;   /* optional (U8 levels): exact integer distances */
    uint32_t* out_counts;          /* num_queries: results per query (<= k) */
    uint64_t* out_bytes;           /* optional: algorithmic bytes read per query */
} pdxg_search_args;

/* Batched search with host buffers; synchronous. Replaces PDXearch::Search driven by
 * PDXIndex::Search (index.hpp:269-271) and PDXTreeIndex::Search (index.hpp:545-566), one row per
 * query. Not reentrant on one index handle (nor is the reference's Search). */
pdxg_status pdxg_search(pdxg_index* index, const pdxg_search_args* args);

/* Scratch for pdxg_search_device, sized for at most max_queries queries with k <= max_k and
 * nprobe <= max_nprobe (0 = all clusters). */
pdxg_status pdxg_workspace_create(pdxg_index* index, uint32_t max_queries, uint32_t max_k,
                                  uint32_t max_nprobe, pdxg_workspace** out);
pdxg_status pdxg_workspace_destroy(pdxg_workspace* ws);
/* Device bytes a workspace of that size takes (for admission control). */
size_t pdxg_workspace_bytes(const pdxg_index* index, uint32_t max_queries, uint32_t max_k,
                            uint32_t max_nprobe);

/* Per-stage device timing of searches run with this workspace: when enabled, every
 * pdxg_search_device call records HIP events on its stream around its stages;
 * pdxg_workspace_stage_times waits for the last call and writes the elapsed milliseconds of
 * [0] preprocessing (normalise + rotate + quantise), [1] probe (centroid distances + ordering),
 * [2] meso-level scan (two-level indexes), [3] data-level scan. Not in the reference. */
pdxg_status pdxg_workspace_enable_timing(pdxg_workspace* ws, int enable);
pdxg_status pdxg_workspace_stage_times(pdxg_workspace* ws, float* out_ms4);
/* Algorithmic bytes read by [0] the meso-level scan and [1] the data-level scan of the last timed
 * search (summed over its queries; centroid probe bytes excluded). */
pdxg_status pdxg_workspace_stage_bytes(pdxg_workspace* ws, uint64_t* out2);
/* Split of the meso-level stage of the last timed search: [0] the two-phase scan's first phase
 * (visit-list inversion and partial-distance summaries), [1] the scan proper (the replay, or the
 * pruned scan when the two-phase scan did not run, with [0] = 0). Not in the reference. */
pdxg_status pdxg_workspace_meso_times(pdxg_workspace* ws, float* out2);
/* (query, meso member) pairs summarised by the first phase of the last timed search (0 when the
 * two-phase scan did not run): each costs d subtractions, d fused multiply-adds and, per 64-dim
 * horizontal block, a 15-add reduction. Not in the reference. */
pdxg_status pdxg_workspace_meso_pairs(pdxg_workspace* ws, uint64_t* out1);

/* Scan workgroup size per level ("waves per cluster": 64 .. 1024 threads per query; 0 = the
 * built-in heuristic). pdxg_workspace_autotune runs the given search (device buffers, as for
 * pdxg_search_device) with every size, level by level, and keeps the fastest; it writes the
 * outputs of the search as a side effect. Results never depend on the sizes. Not in the
 * reference. */
pdxg_status pdxg_workspace_set_plan(pdxg_workspace* ws, uint32_t block_meso, uint32_t block_data);
pdxg_status pdxg_workspace_get_plan(pdxg_workspace* ws, uint32_t* out2);
/* Two-phase meso-level scan of two-level indexes (on by default where the workspace holds its
 * summaries): phase 1 computes every partial distance the meso scan can compare, for all
 * (query, visited meso member) pairs, sharing each meso cluster's reads across the queries that
 * visit it; phase 2 replays the reference's Warmup / Prune / Merge on those values. Results are
 * identical either way; the autotuner also picks between the two. pdxg_workspace_meso_summary
 * returns 1 when the next search will use it (tombstoned meso members or meso clusters grown past
 * the workspace's plan fall back to the pruned scan). Not in the reference. */
pdxg_status pdxg_workspace_set_meso_summary(pdxg_workspace* ws, int enable);
int pdxg_workspace_meso_summary(const pdxg_workspace* ws);
pdxg_status pdxg_workspace_autotune(pdxg_index* index, pdxg_workspace* ws, const pdxg_search_args* args,
                                    void* stream);
/* Order of the per-query scans (a query's clusters always run in order on one workgroup, as the
 * reference's threshold chain requires): 0 = one workgroup per query in query order;
 * 1 (default) = longest first: the queries sorted on device by descending data-level visit count,
 * so the longest cluster walks start first and short ones fill in behind them; 2 = longest first
 * through a work queue: a persistent grid of as many workgroups as fit on the device at once takes
 * the sorted queries with an atomic counter until none is left. Results do not depend on it; the
 * autotuner keeps the fastest. pdxg_workspace_order returns the current mode. Not in the
 * reference. */
pdxg_status pdxg_workspace_set_order(pdxg_workspace* ws, int mode);
int pdxg_workspace_order(const pdxg_workspace* ws);

/* Asynchronous batched search on `stream` (a hipStream_t, NULL = default stream) with device
 * buffers. The workspace must not be shared by calls that overlap in time. Calls with workspaces
 * and outputs of their own may overlap on different streams over one index (the index is only
 * read): throughput callers keep several batches in flight that way, so that one batch's
 * memory-bound data-level scan runs beside another's compute-bound meso phase. */
pdxg_status pdxg_search_device(pdxg_index* index, pdxg_workspace* ws, const pdxg_search_args* args,
                               void* stream);

/* pdxg_search_device without allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments (device buffers) and results as pdxg_search_device; its per-call
 * scratch is carved out of one fixed 64 MiB array that the library keeps on each device it runs on
 * until the process ends, so no workspace is needed and nothing is allocated or freed. The meso
 * level of two-level indexes always runs the pruned scan (the two-phase scan's summaries need far
 * more room). A call whose scratch would exceed 64 MiB (about 2 x num_queries x d x 4 bytes plus
 * num_queries x (centroids probed + nprobe) x 8) returns PDXG_ERR_UNSUPPORTED. All calls on one GPU
 * share its array: calls that overlap in time on one GPU (other threads, other streams) race and
 * must be serialised by the caller. Asynchronous on `stream`. */
pdxg_status pdxg_search_device_noalloc(pdxg_index* index, const pdxg_search_args* args, void* stream);

/* Query preprocessing only (normalise if the index is normalised, rotate; searcher.hpp:601-613),
 * host buffers: out_rotated num_queries*d floats, out_quantized (U8 indexes, may be NULL)
 * num_queries*d bytes (ScalarQuantizer::QuantizeEmbedding, quantizers/scalar.hpp:76-94). */
pdxg_status pdxg_preprocess(pdxg_index* index, const float* queries, uint32_t num_queries,
                            uint32_t queries_transformed, float* out_rotated, uint8_t* out_quantized);

/* Filtered search of one query (PDXearch::FilteredSearch, searcher.hpp:726-856, as driven by
 * PDXIndex::FilteredSearch, index.hpp:273-280). selection: one byte per slot of the data level
 * in cluster-offset order (PredicateEvaluator::selection_vector, db_mock/predicate_evaluator.hpp;
 * offsets = running sum of max_capacity, ivf_wrapper.hpp:76-87), host buffer. */
pdxg_status pdxg_filtered_search(pdxg_index* index, const float* query, uint32_t query_transformed,
                                 uint32_t k, uint32_t nprobe, const uint8_t* selection,
                                 uint32_t* out_ids, float* out_distances, uint32_t* out_int_distances,
                                 uint32_t* out_count);

/* Search of the meso level only (two-level indexes): the top-level PDXearch<F32> over the L1
 * centroids that PDXTreeIndex::Append runs to pick a cluster (index.hpp:596-607). Queries are
 * already normalised + rotated (host buffers); nprobe applies to the meso clusters. Output ids
 * are L1 cluster ids, ascending by distance. */
pdxg_status pdxg_search_meso(pdxg_index* index, const float* queries_transformed, uint32_t num_queries,
                             uint32_t k, uint32_t nprobe, uint32_t* out_ids, float* out_distances,
                             uint32_t* out_counts);

/* ---------------------------------------------------------------------------------------------
 * Serving: continuous batching over a resident index (not in the reference, which serves one
 * query per call on one thread). Requests (one query each, own k and nprobe) are queued by
 * pdxg_server_submit; a dispatcher thread packs the waiting requests into a batch whenever one of
 * the server's slots is free and launches it on that slot's stream, so batches overlap on the
 * device; a completion thread polls the batches' events in dispatch order and hands a batch's
 * results back once it and every batch dispatched before it are done. Admission is by scratch
 * memory: each slot holds a fixed part sized for max_batch queries and, for two-level indexes, a
 * buffer for the two-phase meso scan's summaries; a request's footprint there is the summarised
 * vectors its nprobe may need (the padded sizes of the largest meso clusters it can visit), and a
 * batch takes waiting requests in arrival order while their footprints fit. With a scratch budget,
 * the budget is shared evenly by max_slots slots (fewer when a share would not hold the fixed part
 * and one request of the largest footprint) and each slot's summary buffer takes what its share
 * leaves. Per-query results do not depend on what a query is batched with.
 * ------------------------------------------------------------------------------------------- */
typedef struct pdxg_server pdxg_server;

typedef struct {
    uint32_t max_batch;          /* queries per launch */
    uint32_t max_k;              /* largest accepted k */
    uint32_t max_nprobe;         /* largest accepted nprobe; 0 = any, including "all" */
    uint32_t max_slots;          /* upper bound on batches in flight (streams) */
    size_t scratch_budget_bytes; /* device scratch for all slots; 0 = no limit beyond max_slots */
    uint32_t scan_block[2];      /* scan workgroup sizes (meso, data), as pdxg_workspace_set_plan */
    uint32_t pipeline_min;       /* a batch starts while another is in flight only when at least this
                                    many requests wait (0 = whenever a slot is free) */
} pdxg_server_config;

typedef struct {
    uint64_t submitted, completed, batches;
    uint32_t slots;              /* slots admitted by the scratch budget */
    size_t scratch_bytes;        /* device scratch held by the slots */
    /* sums over completed batches (ms): device time of the search (HIP events around it), host
     * time from taking the requests to the last launch, and taking the requests to results ready */
    double gpu_ms, host_ms, turnaround_ms;
    /* admission by footprint (two-level indexes with the two-phase meso scan): each slot's
     * meso-summary buffer, the sum over dispatched batches of their requests' footprints, and the
     * batches that the footprint closed before max_batch */
    size_t summary_capacity_bytes;
    double admitted_summary_bytes;
    uint64_t footprint_limited_batches;
} pdxg_server_stats;

pdxg_status pdxg_server_create(pdxg_index* index, const pdxg_server_config* config, pdxg_server** out);
/* Stops the threads after the queued requests complete. */
pdxg_status pdxg_server_destroy(pdxg_server* server);
/* Enqueue one query (host buffer, d floats, copied). tag identifies the result. */
pdxg_status pdxg_server_submit(pdxg_server* server, const float* query, uint32_t k, uint32_t nprobe,
                               uint64_t tag);
/* Collect up to max_results completed requests, waiting up to timeout_ms for the first one.
 * Row r of out_ids / out_distances (stride out_stride >= max_k) belongs to out_tags[r];
 * out_latency_ms (optional) is submit-to-completion time. */
pdxg_status pdxg_server_poll(pdxg_server* server, uint32_t max_results, int timeout_ms, uint64_t* out_tags,
                             uint32_t out_stride, uint32_t* out_ids, float* out_distances, uint32_t* out_counts,
                             float* out_latency_ms, uint32_t* num_results);
pdxg_status pdxg_server_get_stats(pdxg_server* server, pdxg_server_stats* out);
/* Scan plans per batch size: runs pdxg_workspace_autotune on one slot for batches of 1, 4, 16, ...
 * queries and max_batch, built from the given sample (host buffers, n queries with their k and
 * nprobe, repeated when a batch is larger than n); afterwards every batch takes the plan tuned for
 * the smallest size >= its query count. The server must be idle (nothing submitted and not yet
 * completed). Not in the reference. */
pdxg_status pdxg_server_autotune(pdxg_server* server, const float* queries, const uint32_t* k, const uint32_t* nprobe,
                                 uint32_t n);
/* The tuned plans, one row of 5 values each: [largest batch, meso block, data block, two-phase
 * meso scan 0/1, scan order mode 0/1/2 (pdxg_workspace_set_order)]; *num_rows = the number of rows
 * (0 before pdxg_server_autotune). */
pdxg_status pdxg_server_get_plans(pdxg_server* server, uint32_t* out, uint32_t max_rows, uint32_t* num_rows);
/* Load driver: submit query i (row i of queries, k[i], nprobe[i]) at arrival_s[i] seconds after the
 * call starts (ascending), collect every result, and write its submit-to-completion latency in ms.
 * Returns the wall time from first arrival to last completion in *elapsed_s. */
pdxg_status pdxg_server_replay(pdxg_server* server, const float* queries, const uint32_t* k, const uint32_t* nprobe,
                               const double* arrival_s, uint32_t n, float* out_latency_ms, double* elapsed_s);

/* ---------------------------------------------------------------------------------------------
 * Dense first pass (not in the reference): a copy of an F32 index's data level (every used slot,
 * normalised and rotated as the index stores it) searched exhaustively by a low-precision MFMA
 * GEMM (BF16, or FP8 e4m3 with per-tensor scales; FP32 accumulation) over the expansion
 * ||q||^2 + ||x||^2 - 2 q.x (squared norms accumulated in FP64), keeping k * oversample
 * candidates per query, which are reranked with the exact FP32 PDX distance -- the one
 * PDXearch's Start computes for a vector it scans whole (searcher.hpp:218-281): the vertical chain
 * over the first v dims, then the 64-dim horizontal blocks. Results are (distance, id) ascending,
 * ties by id. Rows of the copy are addressed by id; the row filter is the bitmap form of the
 * passing row ids of PDXIndex::FilteredSearch (index.hpp:273-280, 833-845).
 * ------------------------------------------------------------------------------------------- */
typedef struct pdxg_dense pdxg_dense;
/* Precision of the copy; PDXG_DENSE_NO_RESIDUAL (or-ed in) turns residual coding off. By default,
 * when the level has centroids, the low-precision rows hold x - c (c: the centroid of the row's
 * cluster; appended rows take their nearest centroid) and q.x = q.c + q.(x - c), with q.c from an
 * FP32 GEMM per search. */
enum { PDXG_DENSE_BF16 = 0, PDXG_DENSE_FP8 = 1, PDXG_DENSE_NO_RESIDUAL = 0x100 };

/* Device buffers; the call is asynchronous on `stream` except when a selection is given (the
 * passing-row count is read back to size the GEMM). */
typedef struct {
    uint32_t num_queries;          /* <= max_queries of the dense copy */
    const float* queries;          /* num_queries * d */
    uint32_t queries_transformed;  /* 1: already normalised + rotated */
    uint32_t k;                    /* results per query */
    uint32_t oversample;           /* k * oversample (<= 256) candidates are reranked in FP32 */
    const uint8_t* selection;      /* NULL, or one byte per id (id capacity): 1 = row passes */
    uint32_t out_stride;           /* >= k */
    uint32_t* out_ids;             /* num_queries * out_stride */
    float* out_distances;          /* num_queries * out_stride, may be NULL */
    uint32_t* out_counts;          /* num_queries */
} pdxg_dense_search_args;

/* Builds the copy from the index's data level (PDXG_F32 / PDXG_TREE_F32). extra_rows reserves
 * room for appended rows, whose ids may go up to (largest id + extra_rows). precision: see above. */
pdxg_status pdxg_dense_create(pdxg_index* index, uint32_t precision, uint64_t extra_rows, uint32_t max_queries,
                              pdxg_dense** out);
pdxg_status pdxg_dense_destroy(pdxg_dense* dense);
pdxg_status pdxg_dense_search(pdxg_dense* dense, const pdxg_dense_search_args* args, void* stream);
/* Appends n rows (host buffers: n * d floats, n ids); raw rows are normalised and rotated with the
 * index's rotation first. An id that is already live is replaced. Synchronous; waits first for every
 * kernel in flight on the device (a search of this copy on any stream). */
pdxg_status pdxg_dense_append(pdxg_dense* dense, const float* rows, const uint32_t* ids, uint32_t n,
                              uint32_t rows_transformed);
/* Deletes rows by id (host buffer); ids without a live row are ignored. Synchronous; waits first
 * for every kernel in flight on the device. */
pdxg_status pdxg_dense_delete(pdxg_dense* dense, const uint32_t* ids, uint32_t n);
/* Device timing of the searches of this copy (HIP events on their stream): when enabled,
 * pdxg_dense_stage_times waits for the last search and writes the milliseconds of [0] query
 * preparation and row filtering, [1] the first pass (the low-precision GEMM with the top-R of every
 * row segment kept in its epilogue), [2] the merge of the segments' lists, [3] the FP32 rerank. */
pdxg_status pdxg_dense_enable_timing(pdxg_dense* dense, int enable);
pdxg_status pdxg_dense_stage_times(pdxg_dense* dense, float* out4);
/* [0] rows (including deleted ones), [1] row capacity, [2] id capacity, [3] device bytes,
 * [4] rows that entered the last search's GEMM, [5] precision | 0x100 when residual coding is on. */
pdxg_status pdxg_dense_info(const pdxg_dense* dense, uint64_t* out6);

/* ---------------------------------------------------------------------------------------------
 * Captured searches (not in the reference): one pdxg_search_device call recorded into a HIP graph
 * and replayed with a single launch, for callers that repeat a search of fixed shape (a serving
 * loop's slot, a fixed batch size, single queries). The graph keeps the call's device pointers,
 * counts and the workspace's plan as they were at capture; each replay searches what those
 * buffers hold when it runs. Re-create the graph after pdxg_index_update_cluster or a plan change.
 * While the graph exists the workspace belongs to it: replays must not overlap each other or
 * other searches on that workspace.
 * ------------------------------------------------------------------------------------------- */
typedef struct pdxg_graph pdxg_graph;
pdxg_status pdxg_graph_create(pdxg_index* index, pdxg_workspace* ws, const pdxg_search_args* args,
                              pdxg_graph** out);
pdxg_status pdxg_graph_launch(pdxg_graph* graph, void* stream);
/* Nodes (kernel launches) of the captured search. */
uint32_t pdxg_graph_nodes(const pdxg_graph* graph);
pdxg_status pdxg_graph_destroy(pdxg_graph* graph);

/* ---------------------------------------------------------------------------------------------
 * Index build on the device: the steps of BuildIndex (index.hpp:640-757) other than k-means, for
 * a builder that keeps the rows in HBM. Every array argument is a device pointer, except the
 * per-cluster metadata of pdxg_build_layout (host arrays); `stream` is the HIP stream (null: the
 * default stream). Each call returns when its work is done.
 * ------------------------------------------------------------------------------------------- */
/* Random rotation (pruners/adsampling.hpp:44-57): Q of the FP64 Householder QR G = QR of the
 * d x d row-major matrix g, with the sign of each diag(R) entry folded into its column of Q (so R's
 * diagonal is positive: the unique QR of a full-rank G). Writes q32 (row-major float, the index's
 * rotation matrix) and, when q64 is not null, the same in double. */
pdxg_status pdxg_build_rotation(const double* g, uint32_t d, float* q32, double* q64, void* stream);
/* y = x . m for n rows of d floats (x normalised first when `normalize`, as the search normalises
 * a query: quantizers/scalar.hpp:23-40), with the search's own rotation kernels, so each row equals
 * what a search computes for the same vector. minmax[0..1] = the smallest and largest value of y
 * (the range of ScalarQuantizer::ComputeQuantizationParams, scalar.hpp:60-74). */
pdxg_status pdxg_build_transform(const float* x, uint64_t n, uint32_t d, const float* m, int normalize, float* y,
                                 float* minmax, void* stream);
/* PDX layout of every cluster (layout.hpp:21-87): position p of `order` (n positions, cluster by
 * cluster: cluster c holds positions cluster_first[c] .. cluster_first[c + 1]) puts row order[p] of
 * y (n_rows x d) in slot p - cluster_first[c] of cluster c, whose buffer of cluster_cap[c] * d
 * elements starts at element cluster_offset[c] of `out` (out_elems elements: bytes when u8, floats
 * otherwise). v = vertical dims (the reference's dimension split; d - v a multiple of 64). U8
 * stores the SQ8 codes clamp(round((y - base) * scale), 0, 255) (scalar.hpp:76-94); slots past a
 * cluster's count are zero. cluster_first (nclusters + 1), cluster_cap and cluster_offset
 * (nclusters) are host arrays. */
pdxg_status pdxg_build_layout(const float* y, uint32_t d, uint32_t v, int u8, float base, float scale,
                              const uint32_t* order, uint64_t n, uint32_t nclusters, const uint64_t* cluster_first,
                              const uint32_t* cluster_cap, const uint64_t* cluster_offset, void* out,
                              uint64_t out_elems, void* stream);

const char* pdxg_last_error(void);

/* Library / device description: "gfx950 <name> <CUs> CUs <LDS> B LDS". */
pdxg_status pdxg_device_info(int device, char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PDX_AMD_H */
