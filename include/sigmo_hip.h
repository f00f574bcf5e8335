/*
 * sigmo_hip.h -- C ABI of the MI355X (gfx950) implementation of SIGMo's batched
 * subgraph-isomorphism hot path: signatures -> candidate filter/refine -> GMCR -> join.
 *
 * Every entry point returns an int status (SIGMO_OK == 0, negative on failure) and
 * records a message retrievable with sigmo_last_error().  All device buffers belong to
 * opaque handles created here; host buffers are caller-owned.  A context binds one
 * device and one HIP stream; it is externally synchronised (one host thread per
 * context).  Calls enqueue on the context stream and return; calls that deliver data
 * to host memory synchronise the stream first.
 *
 * Reference interfaces replaced (paths relative to the SIGMo repository):
 *   createDeviceCSRGraph / destroyDeviceCSRGraph   library/include/graph.hpp:255-322
 *   getDeviceGraphAllocSize                          library/include/graph.hpp:324-332,423-432
 *   signature::Signature<PowerGraph,4>               library/include/signature.hpp:20-212,336-390
 *   candidates::Candidates                           library/include/candidates.hpp:14-204
 *   isomorphism::filter::filterCandidates            library/include/isomorphism.hpp:22-62
 *   isomorphism::filter::refineCandidates            library/include/isomorphism.hpp:64-127
 *   isomorphism::mapping::GMCR::generateGMCR         library/include/gmcr.hpp:34-128
 *   isomorphism::join::joinCandidates                library/include/isomorphism.hpp:270-381
 *   io::loadCSRGraphsFromFile / FromLines            library/include/io.hpp:37-55, graph.hpp:151-247
 *   utils::BatchedEvent::getProfilingInfo            library/include/utils.hpp:127-161
 */
#ifndef SIGMO_HIP_H
#define SIGMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
#define SIGMO_OK 0
#define SIGMO_ERR_INVALID (-1)     /* bad argument / malformed input                    */
#define SIGMO_ERR_NOMEM (-2)       /* device or host allocation failed                  */
#define SIGMO_ERR_HIP (-3)         /* HIP runtime error                                 */
#define SIGMO_ERR_UNSUPPORTED (-4) /* shape outside the supported envelope              */
#define SIGMO_ERR_OVERFLOW (-5)    /* a counter exceeded its documented width           */
#define SIGMO_ERR_IO (-6)          /* file read/write failure                           */

/* Limits of this implementation (the reference assumes k <= 29 and n <= 64 silently). */
#define SIGMO_MAX_QUERY_NODES 32     /* nodes per query graph  (reference: stack[30])    */
#define SIGMO_MAX_DATA_GRAPH_NODES 128 /* nodes per data graph (reference: 64, ZINC filter) */
#define SIGMO_WILDCARD_LABEL 255     /* types.hpp:22 WILDCARD_NODE                        */

/* Signature scope, signature.hpp:18 SignatureScope{Data, Query}. */
#define SIGMO_SCOPE_DATA 0
#define SIGMO_SCOPE_QUERY 1

/* Phase ids for sigmo_phase_ms (per-phase accumulated kernel time, like BatchedEvent). */
#define SIGMO_PHASE_SIG_DATA 0
#define SIGMO_PHASE_SIG_QUERY 1
#define SIGMO_PHASE_FILTER 2 /* filter + every refine iteration (sigmo.cpp filter_times) */
#define SIGMO_PHASE_GMCR 3
#define SIGMO_PHASE_JOIN 4
#define SIGMO_PHASE_COUNT 5

typedef struct sigmo_ctx sigmo_ctx;
typedef struct sigmo_host_graphs sigmo_host_graphs;
typedef struct sigmo_graph sigmo_graph;
typedef struct sigmo_am_graph sigmo_am_graph;
typedef struct sigmo_signatures sigmo_signatures;
typedef struct sigmo_candidates sigmo_candidates;
typedef struct sigmo_gmcr sigmo_gmcr;

/* ---- errors ------------------------------------------------------------------------ */
/* Message of the last failing call on this host thread ("" if none). */
const char* sigmo_last_error(void);

/* ---- context (device + stream); replaces the sycl::queue every reference class holds -- */
int sigmo_ctx_create(int device, sigmo_ctx** out);
int sigmo_ctx_destroy(sigmo_ctx* ctx);
int sigmo_ctx_sync(sigmo_ctx* ctx);
/* Accumulated kernel milliseconds of one phase since the last reset (hipEvent timing,
 * the analogue of utils.hpp:149-157 BatchedEvent::getProfilingInfo). */
int sigmo_phase_ms(sigmo_ctx* ctx, int phase, double* ms);
int sigmo_phase_reset(sigmo_ctx* ctx);
/* Launch marks for per-call timing (utils.hpp:127-161 BatchedEvent): every timed kernel
 * launch on the context gets a sequence number; sigmo_event_mark returns the next one and
 * sigmo_event_elapsed sums the kernel milliseconds of launches [from, to) (synchronises). */
int sigmo_event_mark(sigmo_ctx* ctx, uint64_t* mark);
int sigmo_event_elapsed(sigmo_ctx* ctx, uint64_t from, uint64_t to, double* ms);
/* Device name and global memory size (sigmo.cpp:20-21). */
int sigmo_device_info(sigmo_ctx* ctx, char* name, size_t name_len, uint64_t* global_mem);

/* ---- host-side batched CSR (the arrays createDeviceCSRGraph concatenates) ------------ */
/* A host batch holds graph_offsets[G+1], row_offsets[N+1] (global), column_indices[E]
 * (global node ids), node_labels[N], edge_labels[E] exactly as graph.hpp:264-294 builds. */
int sigmo_host_graphs_from_arrays(uint32_t num_graphs, const uint32_t* graph_offsets, const uint32_t* row_offsets,
                                  const uint32_t* column_indices, const uint8_t* node_labels, const uint8_t* edge_labels,
                                  sigmo_host_graphs** out);
/* .dat loader, one graph per line "n#N l#L (id label)*N e#E (u v [label])*E"
 * (io.hpp:46-55 loadCSRGraphsFromFile via graph.hpp:154-247). */
int sigmo_host_graphs_load_dat(const char* path, sigmo_host_graphs** out);
/* Same parser over an in-memory buffer of '\n'-separated lines (io.hpp:37-44). */
int sigmo_host_graphs_parse_dat(const char* text, size_t len, sigmo_host_graphs** out);
/* The .dat lines that rank `rank` of `nprocs` reads in the reference's multi-node driver
 * (library/src/sigmo_mpi.cpp:18-80, loadFileLinesMPI), with plain file reads in place of MPI-IO:
 * the bytes [rank * (size / nprocs), + 2048 * max_lines + 1024) (past the end of the file: NUL),
 * the partial first line dropped on ranks > 0 and the partial last line on ranks < nprocs - 1, at
 * most max_lines lines, then only lines with exactly one 'n', 'e' and 'l' kept (:72-77).  The
 * kept lines, each ended by '\n', go to text[0..*len) when cap >= *len (call with text = NULL to
 * size it); *num_lines = lines kept.  A NUL run past the end of the file is returned as one NUL. */
int sigmo_dat_shard_lines(const char* path, uint32_t rank, uint32_t nprocs, uint32_t max_lines, char* text, size_t cap, size_t* len,
                          uint32_t* num_lines);
/* Those lines parsed as io::loadCSRGraphsFromLines does (io.hpp:37-44; sigmo_mpi.cpp:110-114). */
int sigmo_host_graphs_load_dat_shard(const char* path, uint32_t rank, uint32_t nprocs, uint32_t max_lines, sigmo_host_graphs** out);
int sigmo_host_graphs_save_dat(const sigmo_host_graphs* g, const char* path);
int sigmo_host_graphs_info(const sigmo_host_graphs* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* total_edges);
int sigmo_host_graphs_arrays(const sigmo_host_graphs* g, const uint32_t** graph_offsets, const uint32_t** row_offsets,
                             const uint32_t** column_indices, const uint8_t** node_labels, const uint8_t** edge_labels);
/* Graphs [first, first+count) as a new batch (ids rebased to 0). */
int sigmo_host_graphs_slice(const sigmo_host_graphs* g, uint32_t first, uint32_t count, sigmo_host_graphs** out);
/* Concatenate b after a. */
int sigmo_host_graphs_concat(const sigmo_host_graphs* a, const sigmo_host_graphs* b, sigmo_host_graphs** out);
int sigmo_host_graphs_free(sigmo_host_graphs* g);

/* ---- deterministic synthetic workload generator (no reference counterpart) ---------- */
/* size_dist: 0 = clamp(round(Normal(24,8)),6,64); 1 = clamp(round(LogNormal(ln 24,0.45)),6,128).
 * Molecules [first, first+count) of the sequence seeded by `seed`: any slice equals the
 * same slice of a longer run (per-molecule PRNG streams). */
#define SIGMO_SIZE_NORMAL 0
#define SIGMO_SIZE_LOGNORMAL 1
int sigmo_gen_molecules(uint64_t seed, uint64_t first, uint64_t count, int size_dist, int threads,
                        sigmo_host_graphs** out);
/* BFS-grown induced query graphs with k ~ U[kmin,kmax].  perturb_frac: exact fraction of
 * queries with one relabelled node; ring_frac < 0: no ring constraint, else the exact
 * fraction of queries containing a cycle (the rest acyclic). */
int sigmo_gen_queries(uint64_t seed, uint32_t count, uint32_t kmin, uint32_t kmax, int size_dist, double perturb_frac,
                      double ring_frac, sigmo_host_graphs** out);
/* Poisson arrivals (exponential gaps at `qps`) and per-request query counts U[qmin,qmax]. */
int sigmo_gen_arrivals(uint64_t seed, uint32_t count, double qps, uint32_t qmin, uint32_t qmax, double* arrival_s,
                       uint32_t* queries_per_request);

/* ---- device batched CSR graph (createDeviceCSRGraph, graph.hpp:255-314) --------------- */
int sigmo_graph_upload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_graph** out);
int sigmo_graph_free(sigmo_graph* g); /* destroyDeviceCSRGraph, graph.hpp:316-322 */
int sigmo_graph_info(const sigmo_graph* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* total_edges);
/* Bytes of the reference CSR arrays (graph.hpp:324-328) and of the whole device image. */
int sigmo_graph_alloc_size(const sigmo_graph* g, uint64_t* csr_bytes, uint64_t* device_bytes);

/* ---- device batched adjacency-matrix graph (DeviceBatchedAMGraph, graph.hpp:114-148) -- */
/* createDeviceAMGraph (graph.hpp:345-392) of a host batch, each graph's matrix built as
 * toAMGraph does (graph.hpp:208-214): bit u*W + v of graph g's ceil(n*n/64) words (from word
 * graph_offsets[g]) is set for every edge u-v, W = (uint16)sqrt(float(64 * words))
 * (utils.hpp:17-28); num_nodes is cumulative.  SIGMO_ERR_INVALID for a graph of more than 64
 * nodes (this form is for query graphs). */
int sigmo_am_graph_upload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_am_graph** out);
int sigmo_am_graph_free(sigmo_am_graph* g); /* destroyDeviceAMGraph, graph.hpp:394-399 */
/* adjacency_words: the sum of the graphs' word counts (the adjacency_t count of
 * getDeviceAMGraphAllocSize(vector<AMGraph>), graph.hpp:410-421). */
int sigmo_am_graph_info(const sigmo_am_graph* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* adjacency_words);
/* Copy the device arrays back (any pointer may be null): graph_offsets[G], num_nodes[G],
 * adjacency[adjacency_words], node_labels[total_nodes]. */
int sigmo_am_graph_download(sigmo_ctx* ctx, const sigmo_am_graph* g, uint32_t* graph_offsets, uint32_t* num_nodes, uint64_t* adjacency,
                            uint8_t* node_labels);

/* ---- signatures (signature::Signature<PowerGraph,4>, signature.hpp:157-212) ---------- */
int sigmo_signatures_create(sigmo_ctx* ctx, uint64_t data_nodes, uint64_t query_nodes, sigmo_signatures** out);
int sigmo_signatures_free(sigmo_signatures* s);
/* generateCSRSignatures (signature.hpp:125-152). */
int sigmo_signatures_generate(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope);
/* refineCSRSignatures<PowerGraph>(view_size) (signature.hpp:336-390). */
int sigmo_signatures_refine(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size);
/* Copy the packed 16x4-bit signatures (SignatureDevice::signature) of a scope. */
/* refineCSRSignatures<Algorithm::ViewBased>(view_size) (signature.hpp:298-334): every node adds,
 * per CSR entry to a neighbour u and label l, (uint8)(before[u][l] - (l == own label ?
 * view_size : 0)) through incrementLabelCount's uint8 arithmetic.  Accumulates onto the current
 * signatures, as the reference does. */
int sigmo_signatures_refine_view_based(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size);
/* generateAMSignatures (signature.hpp:100-120): each node's first four neighbours (ascending;
 * MAX_NEIGHBORS = 4, types.hpp:28) add their labels onto its signature. */
int sigmo_signatures_generate_am(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope);
/* refineAMSignatures<PowerGraph>(view_size) (signature.hpp:250-296): the signature is cleared
 * and rebuilt as the label histogram of a view_size-layer BFS over 32-bit position sets
 * (Bitset<uint32_t>: positions >= 32 alias mod 32), the node's own position excluded. */
int sigmo_signatures_refine_am(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope, uint32_t view_size);
/* refineAMSignatures<ViewBased>(view_size) (signature.hpp:214-248): the ViewBased rule above over
 * each node's first four AM neighbours. */
int sigmo_signatures_refine_am_view_based(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope, uint32_t view_size);
int sigmo_signatures_download(sigmo_ctx* ctx, const sigmo_signatures* s, int scope, uint64_t* out, uint64_t count);
/* getDataSignatureAllocationSize / getQuerySignatureAllocationSize (signature.hpp:181-198). */
int sigmo_signatures_alloc_size(const sigmo_signatures* s, int scope, uint64_t* bytes);

/* ---- candidates (candidates::Candidates, candidates.hpp:165-204) ---------------------- */
/* Zero-initialised source_nodes x target_nodes bitset (candidates.hpp:165-179). */
int sigmo_candidates_create(sigmo_ctx* ctx, uint64_t source_nodes, uint64_t target_nodes, sigmo_candidates** out);
int sigmo_candidates_free(sigmo_candidates* c);
/* filterCandidates<Query> (isomorphism.hpp:22-62). */
int sigmo_candidates_filter(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data,
                            const sigmo_signatures* s);
/* refineCandidates<Query> (isomorphism.hpp:64-127) with the current-level signatures. */
/* filterCandidates<CandidatesDomain::Data> (isomorphism.hpp:22-62, :49-52): the same label
 * test stored transposed -- create the candidates with source_nodes = data nodes and
 * target_nodes = query nodes; row d lists the query nodes data node d may match. */
int sigmo_candidates_filter_data(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data);
int sigmo_candidates_refine(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data,
                            const sigmo_signatures* s);
/* Download in the reference layout: uint32 words, row stride ceil(target/32)
 * (getHostCandidates, candidates.hpp:185-190).  `words` holds rows*stride words. */
int sigmo_candidates_download(sigmo_ctx* ctx, const sigmo_candidates* c, uint32_t* words, uint64_t n_words);
/* getCandidatesCount(source) for every source row (candidates.hpp:77-81). */
int sigmo_candidates_counts(sigmo_ctx* ctx, const sigmo_candidates* c, uint32_t* counts, uint64_t n);
/* getAllocationSize (candidates.hpp:196) in the reference's bytes, and device bytes. */
int sigmo_candidates_alloc_size(const sigmo_candidates* c, uint64_t* ref_bytes, uint64_t* device_bytes);

/* ---- GMCR (isomorphism::mapping::GMCR, gmcr.hpp:18-135) ------------------------------- */
int sigmo_gmcr_create(sigmo_ctx* ctx, sigmo_gmcr** out);
int sigmo_gmcr_free(sigmo_gmcr* m);
/* generateGMCR (gmcr.hpp:34-128).  Query graphs with fewer than 2 nodes are rejected. */
int sigmo_gmcr_generate(sigmo_ctx* ctx, sigmo_gmcr* m, const sigmo_graph* query, const sigmo_graph* data,
                        const sigmo_candidates* c);
int sigmo_gmcr_total(const sigmo_gmcr* m, uint64_t* total_query_indices);
/* data_graph_offsets[Gd+1] and query_graph_indices[P], ascending query id per segment. */
int sigmo_gmcr_download(sigmo_ctx* ctx, const sigmo_gmcr* m, uint32_t* data_graph_offsets, uint64_t n_offsets,
                        uint32_t* query_graph_indices, uint64_t n_indices);

/* ---- join (isomorphism::join::joinCandidates, isomorphism.hpp:270-381) ---------------- */
/* Adds the match count to *num_matches (the reference accumulates, sigmo.cpp:145-146).
 * per_pair_counts (nullable, host, GMCR order, length total_query_indices) receives each
 * pair's count: 0/1 in find-first mode, leaf hits in find-all mode. */
int sigmo_join(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c,
               const sigmo_gmcr* m, int find_first, uint64_t* num_matches, int32_t* per_pair_counts);

/* sigmo_join without allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, accumulation and per-pair output as sigmo_join; the join's
 * counters, per-pair counts and list of wide-molecule pairs live in a static device array of the
 * library (16 MiB per device, held until the process ends) instead of allocated buffers, so the
 * GMCR may hold at most SIGMO_JOIN_NOALLOC_MAX_PAIRS pairs (more: SIGMO_ERR_UNSUPPORTED).  All
 * calls on one device share that array: calls that overlap on one device (several contexts or
 * threads at once) race; calls in sequence are independent. */
#define SIGMO_JOIN_NOALLOC_MAX_PAIRS (1u << 21)
int sigmo_join_noalloc(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c,
                       const sigmo_gmcr* m, int find_first, uint64_t* num_matches, int32_t* per_pair_counts);

/* ---- wildcard / edge-label join (isomorphism::join::joinWildcardCandidates,
 * isomorphism.hpp:428-556, with defineMatchingOrder :383-408 and isValidMapping :410-426) ----
 * Per GMCR pair: the query node with most candidates in the molecule goes first (ties: larger
 * index; all k <= 32 nodes are considered), then the others in ascending order; each candidate
 * of the first node starts its own DFS; adjacent query/data pairs must carry equal edge labels.
 * In find-first mode a pair counts the first-node candidates that extend to a match.  Same
 * accumulation and per-pair output as sigmo_join. */
int sigmo_join_wildcard(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c,
                        const sigmo_gmcr* m, int find_first, uint64_t* num_matches, int32_t* per_pair_counts);

/* ---- whole plain pipeline (sigmo.cpp:102-159 phase order) ---------------------------- */
typedef struct sigmo_pipeline_result {
  uint64_t num_matches;    /* "# Matches" (sigmo.cpp:215)                               */
  uint64_t gmcr_pairs;     /* total_query_indices                                         */
  uint64_t matched_pairs;  /* pairs with count > 0 (== num_matches in find-first mode)    */
  double phase_ms[SIGMO_PHASE_COUNT];
  double total_ms;         /* device time from first signature kernel to join completion */
  uint64_t dfs_steps;      /* join DFS frame steps (accepts + backtracks + leaves)         */
  uint64_t join_wave_steps; /* join loop iterations summed over wavefronts (lane use = dfs_steps / (64 x this)) */
} sigmo_pipeline_result;
/* Runs gen(D), gen(Q), filter, R x {refsig(D,r), refsig(Q,r), refine}, GMCR, join on
 * device-resident graphs, with one launch per phase; allocates and frees its buffers.  With
 * R >= 1 the filter runs after the level-1 signature passes, fused with the first refine into one
 * launch (the same bits: the filter reads no signatures, the signature passes no candidates);
 * sigmo_kernel_stats then counts that launch under SIGMO_KERNEL_FILTER and R - 1 refines. */
int sigmo_pipeline_run(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, uint32_t refinement_steps,
                       int find_first, sigmo_pipeline_result* result);

/* The same plain pipeline with its signature/candidate/GMCR buffers allocated once
 * (the reference's setup phase, sigmo.cpp:75-91) and reused by every execute. */
typedef struct sigmo_pipeline sigmo_pipeline;
int sigmo_pipeline_create(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, sigmo_pipeline** out);
int sigmo_pipeline_execute(sigmo_pipeline* p, uint32_t refinement_steps, int find_first, sigmo_pipeline_result* result);
/* One hipGraph for the whole pipeline: sigmo_pipeline_capture executes once to size every
 * buffer, then records the same launches (signature resets, every phase kernel, GMCR with the
 * pair count of that execute, the join) as a graph; sigmo_pipeline_replay launches it.  Work per
 * replay is that of an execute; only host round trips and launch gaps go away.  The graphs must
 * stay allocated and unchanged; a replay whose pair counts differ from the capture's returns
 * SIGMO_ERR_INVALID.  phase_ms is not filled by a replay (total_ms is). */
int sigmo_pipeline_capture(sigmo_pipeline* p, uint32_t refinement_steps, int find_first);
int sigmo_pipeline_replay(sigmo_pipeline* p, sigmo_pipeline_result* result);
/* Per-pair counts of the last execute or replay (GMCR order) and the GMCR arrays. */
int sigmo_pipeline_pairs(sigmo_pipeline* p, uint32_t* data_graph_offsets, uint64_t n_offsets, uint32_t* query_graph_indices,
                         int32_t* counts, uint64_t n_pairs);
int sigmo_pipeline_free(sigmo_pipeline* p);

/* ---- per-kernel timing (hipEvents on the context stream) ----------------------------- */
#define SIGMO_KERNEL_SIG_GENERATE 0
#define SIGMO_KERNEL_SIG_REFINE 1
#define SIGMO_KERNEL_FILTER 2
#define SIGMO_KERNEL_REFINE 3
#define SIGMO_KERNEL_GMCR 4
#define SIGMO_KERNEL_JOIN 5
#define SIGMO_KERNEL_PAIR_FILTER 6 /* batched form: fused filter/refine/GMCR predicate */
#define SIGMO_KERNEL_DB_PREPARE 7  /* batched form: per-molecule level planes            */
#define SIGMO_KERNEL_COUNT 8
/* Launch count, summed milliseconds and summed algorithmic HBM bytes of one kernel family
 * since the last sigmo_phase_reset. */
int sigmo_kernel_stats(sigmo_ctx* ctx, int kernel, uint64_t* launches, double* ms, double* bytes);

/* ---- batched form (no reference counterpart: same results as the plain pipeline) ------- */
/* The data side is prepared once per device graph and stays resident: every molecule's
 * PowerGraph level signatures for levels 1..R as bit-planes over its node positions.  A
 * batch run then evaluates each (query graph, molecule) pair directly -- the label filter
 * and the R refine levels per query node inside the molecule, stopping at the first query
 * node without candidates (the GMCR predicate) -- keeps the surviving pairs' candidate masks,
 * orders them as GMCR does and joins them.  No Q x D bitset is materialised.  GMCR arrays and
 * per-pair counts equal those of sigmo_pipeline_run with the same R. */
typedef struct sigmo_db sigmo_db;
typedef struct sigmo_batch sigmo_batch;
typedef struct sigmo_batch_result {
  uint64_t num_matches;   /* sum of per-pair counts                      */
  uint64_t matched_pairs; /* pairs with count > 0                         */
  uint64_t gmcr_pairs;    /* surviving (query graph, molecule) pairs      */
  uint64_t dfs_steps;     /* join DFS frame steps                         */
  uint64_t filter_wave_steps; /* pair evaluation: wavefront steps        */
  uint64_t filter_lane_steps; /* pair evaluation: active lanes summed    */
  uint64_t filter_scans;      /* pair evaluation: pairs that passed the pre-test */
  uint64_t join_wave_steps;   /* join loop iterations summed over wavefronts */
  uint64_t one_sync;          /* 1 if the run first took the one-synchronisation path */
  uint64_t arena_repeats;     /* pair evaluations repeated after the candidate arena overflowed */
} sigmo_batch_result;
/* Launch shape of the batched form's persistent kernels; 0 fields keep the built-in choice. */
typedef struct sigmo_batch_tuning {
  uint32_t filter_blocks_per_cu; /* pair evaluation: 4-wavefront blocks per CU (capped by residency); 0: all that stay resident (<= 8) */
  uint32_t filter_claim;         /* molecule mode: molecules a wavefront claims at a time (0: ~count / (4 x wavefronts), 1..64);
                                    list mode: 64-pair groups a wavefront claims at a time (0: 2) */
  uint32_t join_blocks_per_cu;   /* lane join: 2-wavefront blocks per CU (1..4); 0: 4 */
  uint32_t join_chunk;           /* GMCR pairs per claim of the lane join; 0: ~P / (16 x wavefronts), 64..1024 */
  uint32_t filter_mode;          /* pair evaluation: 1 molecule by molecule from LDS-staged planes, 2 pair by pair from the
                                    pre-tested pair list; 0: list mode below 256 query graphs */
} sigmo_batch_tuning;
int sigmo_db_prepare(sigmo_ctx* ctx, const sigmo_graph* data, uint32_t refinement_steps, sigmo_db** out);
/* Recompute the level planes from the data graph's current device contents (same result as a
 * fresh sigmo_db_prepare; no allocation). */
int sigmo_db_refresh(sigmo_db* db);
int sigmo_db_free(sigmo_db* db);
int sigmo_db_alloc_size(const sigmo_db* db, uint64_t* bytes);
int sigmo_batch_create(sigmo_ctx* ctx, const sigmo_db* db, sigmo_batch** out);
int sigmo_batch_run(sigmo_batch* b, const sigmo_graph* query, int find_first, sigmo_batch_result* result);
/* GMCR arrays (data_graph_offsets[Gd+1], query_graph_indices[P]) and per-pair counts of the
 * last run; any pointer may be null. */
int sigmo_batch_pairs(sigmo_batch* b, uint32_t* data_graph_offsets, uint64_t n_offsets, uint32_t* query_graph_indices,
                      int32_t* counts, uint64_t n_pairs);
int sigmo_batch_free(sigmo_batch* b);
int sigmo_batch_set_tuning(sigmo_batch* b, const sigmo_batch_tuning* t);
int sigmo_batch_get_tuning(const sigmo_batch* b, sigmo_batch_tuning* t);
/* Autotuner: times sigmo_batch_run over a small grid of launch shapes (pair evaluation mode and shape first,
 * then the join), each the best of `repeats` runs, and keeps the fastest in the batch.  Results
 * do not depend on the tuning.  *best_ms (nullable) receives the winner's time. */
int sigmo_batch_autotune(sigmo_batch* b, const sigmo_graph* query, int find_first, uint32_t repeats, sigmo_batch_tuning* chosen,
                         double* best_ms);
/* One hipGraph for a whole run: sigmo_batch_capture runs the batch once to size every buffer,
 * then records [the database refresh when refresh_db,] query preparation, pair evaluation,
 * GMCR and join, with their sizes fixed, into a graph; sigmo_batch_replay launches it and reads
 * the results (sigmo_batch_pairs works after it).  The query graph and database must stay
 * allocated; if their contents changed so that the pair counts differ from the captured run, the
 * replay returns SIGMO_ERR_INVALID (capture again). */
int sigmo_batch_capture(sigmo_batch* b, const sigmo_graph* query, int find_first, int refresh_db);
int sigmo_batch_replay(sigmo_batch* b, sigmo_batch_result* res);
/* Per-request results of the last run, for a batch whose query graphs are the concatenation of
 * requests: request r owns graphs [request_graph_offsets[r], request_graph_offsets[r+1]).  The
 * hash is order-independent over the request's (molecule, request-local query graph, count)
 * triples, so equal hashes mean equal results whatever else shared the batch.  At most 256
 * requests per batch. */
typedef struct sigmo_request_stats {
  uint64_t matched_pairs;
  uint64_t embeddings; /* sum of per-pair counts */
  uint64_t hash;
} sigmo_request_stats;
int sigmo_batch_request_stats(sigmo_batch* b, const uint32_t* request_graph_offsets, uint32_t num_requests, sigmo_request_stats* out);
/* Device bytes a run over `num_query_graphs` query graphs of `query_nodes` nodes in total is
 * expected to hold beyond the database: the per-pair offsets (4 B per query graph x molecule),
 * the survival bits and the first sizing of the candidate arena.  Runs whose survivors outgrow
 * the arena's first sizing grow it (and keep it for later runs). */
int sigmo_batch_workspace_bytes(const sigmo_batch* b, uint32_t num_query_graphs, uint64_t query_nodes, uint64_t* bytes);

/* ---- serving loop (continuous batching over a resident database) ------------------------ */
/* No counterpart in the reference, whose driver runs one query file per process
 * (library/src/sigmo.cpp:102-159).
 *
 * Request r arrives at arrival_s[r] seconds after the call starts (real time; arrival_s must not
 * decrease) and carries the query graphs [request_graph_offsets[r], request_graph_offsets[r+1]).
 * Each worker, whenever it is free, admits in arrival order the requests that have arrived,
 * while its batch stays within max_batch_requests and max_batch_graphs and the bytes of every
 * batch in flight stay within memory_budget (a batch costs bytes_per_graph per query graph plus
 * bytes_per_batch).  Requests that arrive while batches run wait for the next free worker; with
 * several workers their batches overlap on the device.  A request whose own cost exceeds the
 * budget is rejected; one that waited longer than deadline_ms before admission expires.  Both
 * are reported and never run. */
enum { SIGMO_SERVE_DONE = 0, SIGMO_SERVE_REJECTED = 1, SIGMO_SERVE_EXPIRED = 2, SIGMO_SERVE_PENDING = 3 };
typedef struct sigmo_serve_options {
  uint32_t max_batch_requests; /* 0 or > 256: 256 (the per-request statistics limit) */
  uint32_t max_batch_graphs;   /* 0: no limit; a single larger request still runs alone */
  uint64_t memory_budget;      /* bytes over all batches in flight; 0: no limit */
  uint64_t bytes_per_graph;    /* cost model; sigmo_serve fills 0 from sigmo_batch_workspace_bytes */
  uint64_t bytes_per_batch;
  double deadline_ms;          /* 0: none */
  int find_first;
} sigmo_serve_options;
typedef struct sigmo_serve_result {
  int32_t status;    /* SIGMO_SERVE_* */
  uint32_t batch;    /* batch that served the request (dispatch order) */
  uint32_t batch_requests;
  uint32_t worker;
  double admit_s;    /* from the call's start */
  double done_s;
  double latency_ms; /* completion (or rejection/expiry) minus arrival */
  uint64_t matched_pairs;
  uint64_t embeddings;
  uint64_t hash;
} sigmo_serve_result;
typedef struct sigmo_serve_summary {
  uint32_t batches;
  uint32_t done, rejected, expired;
  double wall_s;          /* call start to last completion */
  double busy_s;          /* time with at least one batch in flight */
  uint64_t max_inflight_bytes;
  uint64_t matched_pairs, embeddings, gmcr_pairs;
} sigmo_serve_summary;
/* Runs one batch for the scheduler: fills out[i] for requests[0..count) and *gmcr_pairs.
 * Returns SIGMO_OK or an error status, which stops the loop. */
typedef int (*sigmo_serve_executor)(void* user, uint32_t worker, const uint32_t* requests, uint32_t count, sigmo_request_stats* out,
                                    uint64_t* gmcr_pairs);
/* The scheduler alone, with a caller-supplied executor (host-testable). */
int sigmo_serve_schedule(uint32_t num_workers, const uint32_t* request_graph_offsets, const double* arrival_s, uint32_t num_requests,
                         const sigmo_serve_options* options, sigmo_serve_executor executor, void* user, sigmo_serve_result* results,
                         sigmo_serve_summary* summary);
/* The serving loop on the device: worker w runs its batches with workers[w] (batch objects over
 * one database, each created on its own context so that their streams overlap), uploading each
 * batch's query graphs from `queries` into a buffer it reuses. */
int sigmo_serve(sigmo_batch* const* workers, uint32_t num_workers, const sigmo_host_graphs* queries, const uint32_t* request_graph_offsets,
                const double* arrival_s, uint32_t num_requests, const sigmo_serve_options* options, sigmo_serve_result* results,
                sigmo_serve_summary* summary);
/* The same loop for requests that arrive while it runs (a service's request handler): open
 * starts one thread per worker batch; submit queues a request (its query graphs are copied,
 * arrival = now, ids count from 0) from any host thread; poll copies a request's result,
 * waiting up to timeout_ms for it to settle (0: no wait, < 0: no limit; status
 * SIGMO_SERVE_PENDING until then); close stops accepting, serves what is queued, joins the
 * threads and frees the server (no other call on the server may overlap close).  Times in the
 * results count from open.  With bytes_per_graph = 0 each request is priced at
 * sigmo_batch_workspace_bytes of a batch holding just it.  After a batch fails, submit and the
 * polls of unserved requests return its error. */
typedef struct sigmo_server sigmo_server;
int sigmo_server_open(sigmo_batch* const* workers, uint32_t num_workers, const sigmo_serve_options* options, sigmo_server** out);
int sigmo_server_submit(sigmo_server* server, const sigmo_host_graphs* query_graphs, uint64_t* request_id);
int sigmo_server_poll(sigmo_server* server, uint64_t request_id, double timeout_ms, sigmo_serve_result* result);
int sigmo_server_close(sigmo_server* server, sigmo_serve_summary* summary);
/* The online scheduler alone, with a caller-supplied executor (host-testable): requests are
 * submitted as query-graph counts. */
int sigmo_server_open_schedule(uint32_t num_workers, const sigmo_serve_options* options, sigmo_serve_executor executor, void* user,
                               sigmo_server** out);
int sigmo_server_submit_count(sigmo_server* server, uint32_t num_query_graphs, uint64_t* request_id);

/* ---- streaming (a molecule shard larger than one device pass) ---------------------------- */
/* Replaces one rank's run of the reference's multi-node driver (library/src/sigmo_mpi.cpp:
 * 110-159) when its shard does not fit the candidate memory: the shard stays in host memory
 * and runs in chunks of chunk_molecules molecules.  With `overlap`, a loader thread slices and
 * uploads chunk i+1 on its own stream while chunk i runs on the context's stream.  Each chunk's
 * GMCR (chunk-local molecule ids) and per-pair counts are handed to on_chunk (when given); their
 * concatenation equals one run over the whole shard. */
enum { SIGMO_FORM_PLAIN = 0, SIGMO_FORM_BATCHED = 1 };
typedef struct sigmo_stream_options {
  uint32_t chunk_molecules; /* 0: the whole shard in one chunk */
  uint32_t refinement_steps;
  int find_first;
  int form;                 /* SIGMO_FORM_PLAIN: per-phase kernels over the chunk's bitset; SIGMO_FORM_BATCHED */
  int overlap;              /* nonzero: load the next chunk while the current one runs */
  uint32_t workers;         /* batched form with overlap: chunks run on this many contexts at once (each
                             * its own stream, database and batch; at most 4; 0 or 1: one).  Kernel
                             * timings of the extra contexts are added to the caller's after a pass. */
} sigmo_stream_options;
typedef struct sigmo_stream_result {
  uint64_t num_matches, matched_pairs, gmcr_pairs, dfs_steps;
  uint32_t chunks;
  double wall_ms;      /* first load to last chunk's results */
  double compute_ms;   /* chunks' device work, host-timed */
  double load_ms;      /* chunk slicing and upload */
  double load_wait_ms; /* time the compute side waited for a chunk */
} sigmo_stream_result;
/* Receives chunk `chunk` (molecules [first_molecule, first_molecule + num_molecules)): GMCR
 * offsets[num_molecules + 1], query graph indices and counts of its num_pairs pairs.  Returns 0
 * to go on. */
typedef int (*sigmo_stream_chunk_fn)(void* user, uint32_t chunk, uint32_t first_molecule, uint32_t num_molecules, const uint32_t* offsets,
                                     const uint32_t* indices, const int32_t* counts, uint64_t num_pairs);
/* One pass over the shard: sigmo_stream_open, sigmo_stream_pass, sigmo_stream_close. */
int sigmo_stream_run(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_host_graphs* data, const sigmo_stream_options* options,
                     sigmo_stream_chunk_fn on_chunk, void* user, sigmo_stream_result* result);
/* A streaming session: the chunk plan and the loader are set up once, and the device buffers
 * (upload slots, the plain pipeline's signatures, bitset and GMCR arrays, or the batched form's
 * database and batch) are allocated by the first pass and kept for later passes, as one rank of
 * the reference's multi-node run allocates once (library/src/sigmo_mpi.cpp:153-195) before its
 * timed phases.  `data` and `query` must outlive the session; the session uses `ctx` (its stream)
 * and a loader context of its own. */
typedef struct sigmo_stream sigmo_stream;
int sigmo_stream_open(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_host_graphs* data, const sigmo_stream_options* options,
                      sigmo_stream** out);
int sigmo_stream_pass(sigmo_stream* stream, sigmo_stream_chunk_fn on_chunk, void* user, sigmo_stream_result* result);
int sigmo_stream_close(sigmo_stream* stream);

#ifdef __cplusplus
}
#endif
#endif /* SIGMO_HIP_H */
