/*
 * b40c_mi355x — C ABI of the MI355X (gfx950) single-GPU BFS library.
 *
 * This is the drop-in boundary for b40c's single-GPU BFS path.  The reference
 * is a header-only C++ template library (no extern "C", no shared library);
 * each entry point below names the reference interface it replaces
 * (file:line in the b40c tree).  All pointers are plain host pointers unless
 * the name says d_ (device pointer).  No entry point throws; every one returns
 * a b40c_status and b40c_last_error_string() describes the last failure on
 * the calling thread.
 *
 * Conventions (identical to the reference): vertex ids, offsets and labels
 * are int32; depth label -1 = unreached, source depth 0; predecessor -1 =
 * unreached, source predecessor -2.
 */
#ifndef B40C_MI355X_H
#define B40C_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  B40C_ERR_QUEUE_OVERFLOW is the reference's
 * cudaErrorInvalidConfiguration "Frontier queue overflow. Please increase
 * queue-sizing factor" (b40c/graph/bfs/enactor_contract_expand.cuh:319-324);
 * B40C_ERR_UNSUPPORTED its cudaErrorInvalidDeviceFunction for an untuned
 * architecture (enactor_contract_expand.cuh:594-595). */
typedef int32_t b40c_status;
#define B40C_OK                 0
#define B40C_ERR_QUEUE_OVERFLOW 1
#define B40C_ERR_UNSUPPORTED    2
#define B40C_ERR_OOM            3
#define B40C_ERR_INVALID_ARG    4
#define B40C_ERR_HIP            5
#define B40C_ERR_ADMISSION      6

/* Strategy numbering mirrors test/bfs/test_bfs.cu:72-79 (HOST = -1 is the
 * CPU path and is not served by this library).  PERSISTENT and
 * DIRECTION_OPTIMIZING have no counterpart in the reference enum. */
#define B40C_EXPAND_CONTRACT      0
#define B40C_CONTRACT_EXPAND      1
#define B40C_TWO_PHASE            2
#define B40C_HYBRID               3
#define B40C_PERSISTENT           4
#define B40C_DIRECTION_OPTIMIZING 5

/* FrontierType (b40c/graph/bfs/csr_problem.cuh:54-61). */
#define B40C_VERTEX_FRONTIERS 0
#define B40C_EDGE_FRONTIERS   1
#define B40C_MIXED_FRONTIERS  2

/* Enactor launch modes: the strategy's own default (the form test_bfs.cu:
 * 690-718 uses), the single persistent launch (EnactFusedSearch) or one
 * launch per level (EnactIterativeSearch). */
#define B40C_MODE_DEFAULT   0
#define B40C_MODE_FUSED     1
#define B40C_MODE_ITERATIVE 2

typedef struct b40c_graph b40c_graph;
typedef struct b40c_problem b40c_problem;
typedef struct b40c_enactor b40c_enactor;

/* Runtime tuning knobs.  They replace the compile-time KernelPolicy blocks
 * the reference hand-sets per architecture (e.g. enactor_hybrid.cuh:569-655,
 * contract_expand_atomic/kernel_policy.cuh:61-279).  Zero means "library
 * default / autotuned". */
typedef struct b40c_tuning {
    int32_t grid_size;             /* workgroups per level launch (0: 4 x CUs)              */
    int32_t cta_gather_threshold;  /* degree >= this: whole-workgroup gather                */
    int32_t wave_gather_threshold; /* degree >= this: whole-wavefront gather                */
    int32_t levels_per_sync;       /* iterative enactors: levels launched between host checks */
    int32_t saturation_quit;       /* hybrid: fused kernel quits when frontier > grid * this */
    int32_t do_alpha;              /* direction-optimizing: switch to bottom-up when m_f > m_u / alpha */
    int32_t do_beta;               /* direction-optimizing: back to top-down when n_f < n / beta   */
    int32_t persistent_wgs_per_cu; /* persistent kernels: resident workgroups per CU        */
    int32_t grid_gather_threshold; /* degree >= this: list split over all workgroups (grid tier) */
    int32_t bottom_up_grid_size;   /* direction-optimizing: workgroups per level launch, for its   */
                                   /* bottom-up steps; top-down steps use grid_size of them       */
} b40c_tuning;

/* ---- errors / device --------------------------------------------------- */
const char* b40c_last_error_string(void);
/* Reports device name, CU count and gfx architecture string. */
b40c_status b40c_device_info(int32_t* num_cus, char* arch, int32_t arch_len);

/* ---- device-resident CSR graphs ---------------------------------------- */
/* Upload a host CSR (CsrProblem::FromHostProblem's H2D copy,
 * csr_problem.cuh:328-403).  Host arrays stay caller-owned. */
b40c_status b40c_graph_from_host_csr(int32_t nodes, int64_t edges, const int32_t* row_offsets,
                                     const int32_t* column_indices, b40c_graph** out);
/* GPU COO -> CSR with the reference's semantics: stable sort by row only,
 * columns in input order (CsrGraph::FromCoo, b40c/graph/csr_graph.cuh:108-152),
 * done with the GPU LSD radix sort instead of std::stable_sort.  Host COO in. */
b40c_status b40c_graph_from_coo(int32_t nodes, int64_t edges, const int32_t* rows, const int32_t* cols,
                                b40c_graph** out);
/* GPU COO -> CSR with sorted, duplicate-free columns per row: packed
 * (row, col) keys, GPU radix sort, consecutive removal and row offsets (the
 * steps the reference builds from consecutive_removal::Enactor::Trim,
 * consecutive_removal/enactor.cuh:119-126, and scan::Enactor::Scan,
 * scan/enactor.cuh:122-136).  flags: B40C_COO_DROP_SELF_LOOPS drops (v, v),
 * B40C_COO_SYMMETRIZE adds every edge's reverse (the graph is then marked
 * symmetric).  Host COO in. */
#define B40C_COO_DROP_SELF_LOOPS 1
#define B40C_COO_SYMMETRIZE      2
b40c_status b40c_graph_from_coo_unique(int32_t nodes, int64_t edges, const int32_t* rows, const int32_t* cols,
                                       int32_t flags, b40c_graph** out);
/* Synthetic RMAT graph built on the GPU: 2^scale vertices, edge_factor*2^scale
 * sampled edges, quadrant probabilities (a,b,c,1-a-b-c), symmetrized,
 * self-loops and duplicates removed, int32 CSR from a radix sort of packed
 * 64-bit (src,dst) keys (replaces BuildRmatGraph, graph/builder/rmat.cuh:127-206,
 * and the host FromCoo sort). */
b40c_status b40c_graph_build_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                                  uint64_t seed, b40c_graph** out);
/* Synthetic w^3 6-neighbour lattice with seeded random vertex relabelling and
 * no self loops (replaces BuildGrid3dGraph, graph/builder/grid3d.cuh:53-127).
 * perm (nullable, w^3 entries) receives original id -> new id. */
b40c_status b40c_graph_build_lattice3d(int32_t width, uint64_t seed, int32_t* perm, b40c_graph** out);
b40c_status b40c_graph_info(const b40c_graph* g, int32_t* nodes, int64_t* edges);
b40c_status b40c_graph_copy_to_host(const b40c_graph* g, int32_t* row_offsets, int32_t* column_indices);
/* Graphs are reference counted: problems hold a reference. */
void b40c_graph_destroy(b40c_graph* g);

/* ---- graph files (test_bfs.cu:976-1037: graph types dimacs, metis, market) */
#define B40C_FILE_MARKET 0 /* MatrixMarket coordinate (b40c/graph/builder/market.cuh) */
#define B40C_FILE_METIS  1 /* METIS adjacency lists (b40c/graph/builder/metis.cuh)    */
#define B40C_FILE_DIMACS 2 /* DIMACS "p"/"a" lines (b40c/graph/builder/dimacs.cuh)   */
typedef struct b40c_edge_list b40c_edge_list;
/* Parses a graph file on the host (path NULL: stdin), in the reference
 * readers' edge order: ReadMarketStream (market.cuh:50-158), ReadMetisStream
 * (metis.cuh:52-158), ReadDimacsStream (dimacs.cuh:54-220).  undirected adds
 * each edge's reverse (MatrixMarket, DIMACS); splice > 0 reads DIMACS files
 * path.0 .. path.<splice-1> merged by row (BuildDimacsGraph, dimacs.cuh:230-270).
 * Ids outside [1, n] are an error (B40C_ERR_INVALID_ARG); weighted METIS
 * formats are B40C_ERR_UNSUPPORTED.  Needs no GPU. */
b40c_status b40c_edge_list_read(const char* path, int32_t format, int32_t undirected, int32_t splice,
                                b40c_edge_list** out);
/* Vertex count, directed edge count, and 1 when the list is a CSR (METIS)
 * rather than a COO list. */
b40c_status b40c_edge_list_info(const b40c_edge_list* el, int32_t* nodes, int64_t* edges, int32_t* is_csr);
/* Copies the COO rows (edges entries, file order) or CSR offsets (nodes + 1)
 * and the columns; either pointer may be NULL. */
b40c_status b40c_edge_list_copy(const b40c_edge_list* el, int32_t* rows_or_offsets, int32_t* cols);
/* Device CSR of a parsed file: COO through the GPU stable sort by row
 * (CsrGraph::FromCoo, csr_graph.cuh:108-152), METIS uploaded as read. */
b40c_status b40c_graph_from_edge_list(const b40c_edge_list* el, b40c_graph** out);
void b40c_edge_list_destroy(b40c_edge_list* el);
/* Read + build in one call (BuildMarketGraph market.cuh:172, BuildMetisGraph
 * metis.cuh:184, BuildDimacsGraph dimacs.cuh:230, then the device CSR). */
b40c_status b40c_graph_read(const char* path, int32_t format, int32_t undirected, int32_t splice,
                            b40c_graph** out);
/* Writes a graph as METIS text, "n edges" then one line of 1-based
 * neighbours per vertex (WriteMetisStream, metis.cuh:220-245); path NULL:
 * stdout. */
b40c_status b40c_graph_write_metis(const b40c_graph* g, const char* path);

/* ---- consecutive removal and scan ------------------------------------- */
/* consecutive_removal::Enactor::Trim (b40c/consecutive_removal/enactor.cuh:
 * 119-126) on device arrays: keeps the first key (and value) of every run of
 * equal consecutive keys, in order; *num_compacted receives the count.  Values
 * are optional (both NULL or both set); outputs hold n entries and must not
 * alias the inputs. */
b40c_status b40c_consecutive_removal_u64(const uint64_t* d_keys_in, const uint32_t* d_values_in,
                                         uint64_t* d_keys_out, uint32_t* d_values_out, int64_t n,
                                         int64_t* num_compacted);
/* scan::Enactor::Scan (b40c/scan/enactor.cuh:122-136): prefix sum of n
 * unsigned 32-bit values (mod 2^32), exclusive or inclusive; d_out may equal
 * d_in; *total (nullable) receives the exact 64-bit sum. */
b40c_status b40c_scan_u32(const uint32_t* d_in, uint32_t* d_out, int64_t n, int32_t exclusive, uint64_t* total);

/* ---- radix sort (b40c/radix_sort/enactor.cuh:153-266) ------------------- */
/* Stable LSD sort of device keys over bits [begin_bit, end_bit) with 8-bit
 * digits (upsweep -> spine -> downsweep per digit place).  d_keys[0] holds the
 * input; d_keys[1] must have room for n keys.  On return *selector says which
 * buffer holds the result (DoubleBuffer::selector, util/multiple_buffering.cuh:
 * 65-80).  Digit places whose keys all share one digit are skipped. */
b40c_status b40c_radix_sort_u64(uint64_t* d_keys0, uint64_t* d_keys1, int64_t n, int32_t begin_bit,
                                int32_t end_bit, int32_t* selector);
b40c_status b40c_radix_sort_pairs_u32(uint32_t* d_keys0, uint32_t* d_keys1, uint32_t* d_vals0,
                                      uint32_t* d_vals1, int64_t n, int32_t begin_bit, int32_t end_bit,
                                      int32_t* selector);
/* b40c_radix_sort_u64 without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code that must
 * not call hipMalloc, tight serving loops, harnesses that forbid it).  The
 * scratch is a fixed 1 MiB device array per GPU, held until the process
 * ends, which covers any n <= 2^32 - 1 on a device of up to 256 CUs (larger
 * requests fail with B40C_ERR_UNSUPPORTED).  All calls on one GPU share that
 * array, so calls that overlap in time on one GPU (other host threads) race
 * and must be serialized by the caller.  Synchronous, on the default stream. */
b40c_status b40c_radix_sort_u64_noalloc(uint64_t* d_keys0, uint64_t* d_keys1, int64_t n, int32_t begin_bit,
                                        int32_t end_bit, int32_t* selector);
/* Key types of the generic sort: the reference's KeyTraits
 * (b40c/radix_sort/sort_utils.cuh:87-183).  Signed keys are ordered by
 * flipping the sign bit, IEEE floats by flipping the sign bit of positive
 * values and every bit of negative ones (so -0.0 sorts before +0.0 and NaNs
 * by their bits); begin_bit/end_bit index those ordered bits.  The keys in
 * memory are never rewritten by the conversion. */
#define B40C_KEY_U32 0
#define B40C_KEY_I32 1
#define B40C_KEY_F32 2
#define B40C_KEY_U64 3
#define B40C_KEY_I64 4
#define B40C_KEY_F64 5
/* radix_sort::Enactor::Sort over DoubleBuffer<K, V> for any key type K above,
 * keys only (d_vals0 = d_vals1 = NULL) or with uint32 values
 * (radix_sort/enactor.cuh:153-266).  Scratch is allocated for the call. */
b40c_status b40c_radix_sort(int32_t key_type, void* d_keys0, void* d_keys1, uint32_t* d_vals0, uint32_t* d_vals1,
                            int64_t n, int32_t begin_bit, int32_t end_bit, int32_t* selector);
/* A sort enactor owns its scratch (the spine) across calls, like the
 * reference Enactor's spine storage (radix_sort/enactor.cuh:96-140): a sort
 * no larger than an earlier one on the same enactor allocates nothing.  One
 * HIP stream per enactor; calls are synchronous (the selector is read back). */
typedef struct b40c_sort_enactor b40c_sort_enactor;
b40c_status b40c_sort_enactor_create(b40c_sort_enactor** out);
b40c_status b40c_sort_enactor_sort(b40c_sort_enactor* e, int32_t key_type, void* d_keys0, void* d_keys1,
                                   uint32_t* d_vals0, uint32_t* d_vals1, int64_t n, int32_t begin_bit,
                                   int32_t end_bit, int32_t* selector);
/* The enactor's launch policy, the run-time form of the reference's
 * per-architecture sort policies (radix_sort/autotuned_policy.cuh:190-230):
 * keys per thread of a downsweep tile (8 or 16; 0 = default) and
 * workgroups per CU (1-8; 0 = default).  The defaults are the tuned values
 * the stateless sorts and the CSR builds use. */
b40c_status b40c_sort_enactor_set_policy(b40c_sort_enactor* e, int32_t items_per_thread, int32_t wgs_per_cu);
void b40c_sort_enactor_destroy(b40c_sort_enactor* e);
/* Device allocation helpers so FFI callers can stage radix-sort buffers. */
b40c_status b40c_device_alloc(void** d_ptr, int64_t bytes);
b40c_status b40c_device_free(void* d_ptr);
b40c_status b40c_memcpy_h2d(void* d_dst, const void* src, int64_t bytes);
b40c_status b40c_memcpy_d2h(void* dst, const void* d_src, int64_t bytes);

/* ---- problem (CsrProblem, b40c/graph/bfs/csr_problem.cuh) --------------- */
/* FromHostProblem (csr_problem.cuh:328-334): copies the host CSR. */
b40c_status b40c_problem_create(int32_t nodes, int64_t edges, const int32_t* column_indices,
                                const int32_t* row_offsets, int32_t flags, b40c_problem** out);
/* Same, over an already device-resident graph (no copy; takes a reference). */
b40c_status b40c_problem_create_from_graph(b40c_graph* g, int32_t flags, b40c_problem** out);
/* Reset (csr_problem.cuh:547-750): (re)allocate queues for frontier_type x
 * queue_sizing and clear labels/visited state.  Must precede every search.
 * queue_sizing <= 0 selects the reference default 1.3 (test_bfs.cu:802). */
b40c_status b40c_problem_reset(b40c_problem* p, int32_t frontier_type, double queue_sizing);
/* ExtractResults (csr_problem.cuh:255-322): depth and predecessor labels
 * (either pointer may be NULL).  Synchronizes the problem's stream. */
b40c_status b40c_problem_extract(b40c_problem* p, int32_t* depth, int32_t* pred);
/* The HIP stream all of the problem's work runs on (hipStream_t). */
void* b40c_problem_stream(b40c_problem* p);
/* Device pointers of the labels (for callers that keep results resident). */
b40c_status b40c_problem_device_labels(b40c_problem* p, int32_t** d_depth, int32_t** d_pred);
void b40c_problem_destroy(b40c_problem* p);

/* ---- enactors (EnactorContractExpand / ExpandContract / TwoPhase / Hybrid) */
/* Enactor construction (enactor_*.cuh constructors; INSTRUMENT template flag). */
b40c_status b40c_enactor_create(int32_t strategy, int32_t instrument, b40c_enactor** out);
b40c_status b40c_enactor_set_tuning(b40c_enactor* e, const b40c_tuning* t);
b40c_status b40c_enactor_get_tuning(const b40c_enactor* e, b40c_tuning* t);
/* The frontier type a problem must be Reset with for this enactor
 * (GetFrontierType, e.g. enactor_contract_expand.cuh:205). */
int32_t b40c_enactor_frontier_type(const b40c_enactor* e);
/* EnactSearch / EnactFusedSearch / EnactIterativeSearch from `src`
 * (enactor_hybrid.cuh:559-563; enactor_contract_expand.cuh:338-342, 525-529).
 * Asynchronous on the problem's stream except for the level-termination
 * checks; returns B40C_ERR_QUEUE_OVERFLOW if any frontier overflowed. */
b40c_status b40c_enact_search(b40c_enactor* e, b40c_problem* p, int32_t src, int32_t max_grid_size,
                              int32_t mode);
/* GetStatistics (enactor_contract_expand.cuh:236-250): total queued frontier
 * elements, search depth (max label + 1) and average workgroup duty
 * (instrumented enactors only, else 0). */
b40c_status b40c_get_statistics(b40c_enactor* e, int64_t* total_queued, int32_t* search_depth,
                                double* avg_duty);
void b40c_enactor_destroy(b40c_enactor* e);

/* ---- batched and served traversal (no counterpart in the reference) ------
 * Many single-source queries over several resident graphs share one
 * bit-parallel multi-source traversal: up to 64 queries per graph run in the
 * same level steps, each with its own int32 depth/predecessor label slot
 * (local vertex ids).  Queries join at level boundaries as a graph slot and
 * label memory (8 bytes x graph vertices, from the budget given at creation)
 * become free, and leave when their traversal ends — continuous batching.
 * Graphs must be symmetric (every edge present in both directions). */
typedef struct b40c_multigraph b40c_multigraph;
typedef struct b40c_msbfs b40c_msbfs;
/* Packs device graphs into one CSR (graph g owns vertices
 * [vbase[g], vbase[g+1])).  The source graphs may be destroyed afterwards. */
b40c_status b40c_multigraph_pack(b40c_graph* const* graphs, int32_t ngraphs, b40c_multigraph** out);
/* vbase: ngraphs + 1 entries, ebase: ngraphs + 1 entries (either nullable). */
b40c_status b40c_multigraph_info(const b40c_multigraph* mg, int32_t* ngraphs, int32_t* vbase, int64_t* ebase);
/* Drops the caller's reference; engines created on the pack keep their own,
 * so the pack may be destroyed before them. */
void b40c_multigraph_destroy(b40c_multigraph* mg);
/* label_budget_bytes: device memory for per-query depth + pred slots (8 B per
 * vertex of the query's graph).  max_queries: queries submitted and not yet
 * collected by b40c_msbfs_result (ids are reused after collection). */
b40c_status b40c_msbfs_create(b40c_multigraph* mg, int64_t label_budget_bytes, int32_t max_queries,
                              b40c_msbfs** out);
/* Direction switch parameters and workgroups per level launch (0 keeps). */
b40c_status b40c_msbfs_set_tuning(b40c_msbfs* m, int32_t alpha, int32_t beta, int32_t grid);
/* Admission policy: while queries are running, waiting queries join only at
 * level boundaries at least `cohort_gap` levels after the last admission, so
 * queries arrive in cohorts that share their expensive first levels (1 =
 * admit at every boundary; default 6).  An idle engine admits at once. */
b40c_status b40c_msbfs_set_admission(b40c_msbfs* m, int32_t cohort_gap);
/* Launch form of a step (north-star batched form; no counterpart in the
 * reference).  levels >= 1: each b40c_msbfs_step replays one hipGraph that
 * holds `levels` levels (six kernels each, the phase read from device
 * memory) and one phase-advance node, then synchronizes once; queries that
 * finish inside the replay are retired on the device and reported at its
 * end.  0 (default): each step launches one level's kernels directly. */
b40c_status b40c_msbfs_set_levels_per_step(b40c_msbfs* m, int32_t levels);
/* Persistent form (north-star batched form; no counterpart in the
 * reference): on = 1 runs each step's levels (levels_per_step, at least 1)
 * in one launch of a co-resident grid, the six steps of every level
 * separated by software grid barriers; 0 (default) uses launches or the
 * captured graph.  The grid is the occupancy API's co-resident maximum
 * capped by the tuning's grid; a barrier wait that times out (workgroups not
 * co-resident) fails the step with B40C_ERR_HIP, after which the queries in
 * flight hold partial labels (destroy the engine). */
b40c_status b40c_msbfs_set_persistent(b40c_msbfs* m, int32_t on);
/* Queues a query (source `src` local to graph `graph`); returns its id.
 * B40C_ERR_ADMISSION if its labels can never fit the budget. */
b40c_status b40c_msbfs_submit(b40c_msbfs* m, int32_t graph, int32_t src, int32_t* query_id);
/* One level boundary: admit waiting queries, run one level of every query in
 * flight, retire finished ones.  Reports queries running and waiting. */
b40c_status b40c_msbfs_step(b40c_msbfs* m, int32_t* running, int32_t* waiting);
/* Steps until every submitted query has finished. */
b40c_status b40c_msbfs_run(b40c_msbfs* m);
/* Finished queries not reported before (up to max_ids). */
b40c_status b40c_msbfs_poll(b40c_msbfs* m, int32_t* ids, int32_t max_ids, int32_t* count);
/* Labels of a finished query (nullable outputs) and its search depth; frees
 * its label slot, and its id may be returned by a later submission. */
b40c_status b40c_msbfs_result(b40c_msbfs* m, int32_t query_id, int32_t* depth, int32_t* pred,
                              int32_t* search_depth);
void* b40c_msbfs_stream(b40c_msbfs* m);
void b40c_msbfs_destroy(b40c_msbfs* m);

/* ---- diagnostics ---------------------------------------------------------- */
/* Times `iterations` back-to-back grid barriers of one launch of a
 * co-resident grid of `grid_size` workgroups (test/gbarrier/test_gbarrier.cu:43-113).
 * light = 0: agent-scope release/acquire at every workgroup (any data);
 * light = 1: the persistent kernels' mode (data passed with agent-scope
 * accesses, no cache maintenance).  Returns milliseconds per barrier in
 * *ms_per_sync. */
b40c_status b40c_gbarrier_benchmark(int32_t grid_size, int32_t iterations, int32_t light, double* ms_per_sync);
/* Correctness check of the same barrier (the reference has none): every
 * workgroup writes a slab, passes the barrier and compares three other
 * workgroups' slabs (through agent-scope accesses when light = 1);
 * *mismatches counts stale or wrong words. */
b40c_status b40c_gbarrier_check(int32_t grid_size, int32_t iterations, int32_t light, int64_t* mismatches);

/* ---- per-stage BFS microbenchmarks ----------------------------------------
 * A level-synchronous BFS split into its status lookup (edge frontier ->
 * culled, labelled vertices as row ranges) and neighbour gather (row ranges ->
 * unculled neighbour ids, load-balanced over equal tiles of the output) stages; each stage runs once as a bench pass (culls
 * and reads only, on private copies of the visited state) and once for real,
 * every launch timed with HIP events (EnactorGatherLookup::EnactSearch,
 * test/bfs/microbench/enactor_gather_lookup.cuh:152-382, and
 * EnactorCull::EnactSearch, test/bfs/microbench/enactor_cull.cuh:152-385). */
#define B40C_MB_GATHER_LOOKUP 0 /* bench lookup: bitmask cull + label read (microbench_bfs.cu) */
#define B40C_MB_CULL 1          /* bench lookup: bitmask, label, wave and history culls (microbench_cull.cu) */

typedef struct b40c_mb_level {
    int32_t iteration;        /* depth labelled by this level's lookup (0: the source) */
    int32_t reserved;
    int64_t lookup_in;        /* ids entering the lookup (the previous gather's output) */
    int64_t bitmask_pass;     /* bench lookup: ids left after the bitmask cull (-1 at iteration 0) */
    int64_t label_pass;       /* bench lookup: ... and whose label was unset */
    int64_t local_pass;       /* B40C_MB_CULL: ... after the wave and history culls; else -1 */
    int64_t discovered;       /* lookup output: vertices labelled `iteration` */
    int64_t gathered;         /* gather output: neighbours of those vertices, with repeats */
    uint64_t gather_checksum; /* bench gather: sum of the gathered neighbour ids (mod 2^64) */
    float offsets_ms;         /* gather set-up: scan of the row lengths + tile partition */
    float bench_lookup_ms, lookup_ms, bench_gather_ms, gather_ms;
} b40c_mb_level;

/* Runs the microbenchmark search from `src`.  Records one b40c_mb_level per
 * iteration into levels[0 .. max_levels); *num_levels receives the number of
 * iterations the search took (more than max_levels: the rest were not
 * recorded).  depth (nullable, host int32[V]) receives the labels. */
b40c_status b40c_microbench_search(const b40c_graph* g, int32_t src, int32_t mode, int32_t* depth,
                                   b40c_mb_level* levels, int32_t max_levels, int32_t* num_levels);

#ifdef __cplusplus
}
#endif

#endif /* B40C_MI355X_H */
