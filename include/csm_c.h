/* C-ABI of the MI355X incremental match engine.
 *
 * One engine owns a device-resident dynamic labelled graph and one or more
 * registered query graphs; every edge update is answered with the exact
 * number of query embeddings it creates (insertion) or destroys (deletion).
 * Plain pointers and sizes only; no exceptions cross this boundary and the
 * library never exits the process.  Calls on one engine must be serialised by
 * the caller (one host thread per engine); csm_classify is read-only.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository, ParaCOSM/CSM/...).
 */
#ifndef CSM_C_H
#define CSM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes.  0 = success, negative = error (details: csm_last_error). */
#define CSM_OK 0
#define CSM_E_ARG -1            /* bad argument / out-of-range id          */
#define CSM_E_UNKNOWN_ALGO -2   /* main.cpp:760-764 "Unknown algorithm"    */
#define CSM_E_DELETE_MISSING -3 /* graph.cpp:98-102 "deletion error"       */
#define CSM_E_DEVICE -4         /* HIP runtime error or no device          */
#define CSM_E_CAPACITY -5       /* batch/label/id exceeds a fixed capacity */
#define CSM_E_UNSUPPORTED -6    /* not implemented on the device yet       */
#define CSM_E_TIME_LIMIT -7     /* watchdog fired; counts are partial      */
#define CSM_E_ADMISSION -8      /* work refused: device memory budget      */

typedef struct csm_engine csm_engine;

/* A labelled undirected graph given as arrays (vertex ids are 0..n-1).
 * Mirrors what Graph::LoadFromFile builds (graph.cpp:217-249). */
typedef struct csm_graph {
  uint32_t num_vertices;
  const uint32_t* vlabels;   /* [num_vertices]                      */
  size_t num_edges;
  const uint32_t* src;       /* [num_edges]                         */
  const uint32_t* dst;       /* [num_edges]                         */
  const uint32_t* elabels;   /* [num_edges]                         */
} csm_graph;

/* One stream record; mirrors InsertUnit (utils/types.h:34-42). */
typedef struct csm_update {
  uint8_t type;    /* 'e' or 'v'                                   */
  uint8_t is_add;  /* 1 insertion, 0 deletion                      */
  uint16_t pad;
  uint32_t id1;    /* vertex id, or edge endpoint                  */
  uint32_t id2;    /* edge endpoint (edges only)                   */
  uint32_t label;  /* vertex label, or edge label (insertions)     */
} csm_update;

/* Options; mirror the reference CLI flags (main.cpp:681-703) plus device
 * knobs.  Zero-initialise and set what you need. */
typedef struct csm_opts {
  uint64_t max_results;     /* --max-results: per update and query, the
                               search stops once the count reaches this and
                               the count reported is min(exact, max_results)
                               (CSM status bit CSM_USTAT_CAPPED); 0 = none.
                               The reference stops each update's search at
                               max_num_results_ (parallel_graphflow.cpp:386,
                               2101) and keeps the count it had reached.    */
  uint32_t print_prep;      /* --print-prep                                */
  uint32_t print_enum;      /* --print-enum: keep one row per embedding
                               each call finds (csm_embeddings); searches
                               then take the hashing path, and batches are
                               not replayed as HIP graphs                  */
  uint32_t homo;            /* --homo: drop the injectivity test           */
  uint32_t threads;         /* -t (accepted; device sizes itself)          */
  uint32_t auto_tuning;     /* --auto-tuning                               */
  uint32_t max_batch;       /* max updates per csm_apply_batch; 0 = 65534  */
  uint32_t want_hash;       /* compute per-update embedding-set hashes     */
  double time_limit_s;      /* --time-limit: wall-clock seconds per engine
                               call that searches (apply_batch, add/remove
                               edge, initial); 0 = none.  When it expires the
                               search stops, every update is still applied to
                               the graph and index, the call returns
                               CSM_E_TIME_LIMIT and csm_update_status flags
                               the updates whose counts are partial (lower
                               bounds).  Replaces execute_with_time_limit
                               (utils/utils.h:19-36) and the reach_time_limit
                               polls (parallel_graphflow.cpp:288,386-387).  */
  uint64_t device_budget;   /* bytes of HBM the engine may use; 0 = 90%    */
  uint32_t split_chunk;     /* candidates per depth-2 task; 0 = autotuned  */
  uint32_t device;          /* HIP device ordinal                          */
  uint64_t match_budget;    /* per query and batch: stop searching a query
                               once its matches in the batch exceed this
                               (0 = unbounded); see csm_query_status       */
  uint32_t spill_budget;    /* work units (loop iterations, 64-entry
                               walks) before a search task hands the rest
                               of its subtree to the next round (0 =
                               autotuned).  A budget of 256 or more is
                               quartered for batches of up to 2,048
                               updates and halved up to 8,192 (not below
                               64): shorter rounds for small batches     */
  uint32_t reference_orders; /* 1: search in the reference plugin's own
                               matching orders: parallel_graphflow the
                               GraphFlow rule (graphflow.cpp:36-160),
                               parallel_turboflux TurboFlux's path-count rule
                               (parallel_turboflux.cpp:533-700), and
                               parallel_symbi SymBi's search-time rule (the
                               next vertex picked at every step by least
                               n2 count, non-isolated first,
                               parrallel_symbi.cpp:4231-4262; its seeds skip
                               the neighbourhood certificate and its batches
                               are not replayed as HIP graphs).  0 (default):
                               the same seeds with positions 2..k-1 chosen by
                               estimated candidate counts from the data
                               graph's label statistics.  Counts are the same
                               either way; only the search effort differs.  */
  uint32_t want_work;       /* 1: count the candidate entries each update's
                               search tests (csm_update_work), the size of
                               its search tree                             */
  uint32_t no_graphs;       /* 1: never replay steady-state batches as HIP
                               graphs (every launch issued one by one)     */
  uint32_t enum_cap;        /* print_enum: rows kept per call (0 = 2^20);
                               csm_embeddings reports how many were found  */
} csm_opts;

/* Creates an engine and uploads the data graph.  `algo` is one of the
 * reference plugin names: "parallel_graphflow", "parallel_symbi",
 * "parallel_turboflux", "graphflow", "symbi", "turboflux"
 * (main.cpp:740-764).  Replaces `new Parallel_Graphflow(...)` etc. and the
 * matching ctor (matching.h:38-42).  On failure returns NULL and stores a
 * CSM_E_* code in *err (if err is not NULL). */
csm_engine* csm_create(const char* algo, const csm_graph* query, const csm_graph* data,
                       const csm_opts* opts, int* err);

/* csm_create with the data graph given as graph-file records already in HBM
 * (d_records: n csm_update records of type 'v' / 'e', e.g. from
 * csm_parse_text_device or csm_parse_text_to_device in CSM_PARSE_GRAPH
 * mode).  The CSR is built from them on the device, as Graph::LoadFromFile
 * and AddVertex / AddEdge build it (graph.cpp:23-93, 217-249): vertex ids run
 * to the largest id of any record, a vertex keeps the label of its first 'v'
 * record (one without any has no label and matches nothing), edges keep file
 * order and the first copy of a repeated edge.  Only the per-vertex labels
 * and the set of edge labels come back to the host. */
csm_engine* csm_create_from_records(const char* algo, const csm_graph* query,
                                    const csm_update* d_records, size_t n, const csm_opts* opts,
                                    int* err);

/* Host text of a graph file parsed on `device` into records that stay in HBM
 * (allocated by the library; free with csm_free_records): the input of
 * csm_create_from_records without a host copy of the records.  Same grammar,
 * errors and *err_line as csm_parse_text.  No reference counterpart. */
int csm_parse_text_to_device(int device, int mode, const char* text, size_t len,
                             csm_update** d_records, size_t* n, uint64_t* err_line);
void csm_free_records(int device, csm_update* d_records);

/* Registers one more standing query on the same data graph (serving); the
 * first query is id 0.  No reference counterpart (the reference runs one
 * query per process). */
int csm_register_query(csm_engine* e, const csm_graph* query, uint32_t* qid);
uint32_t csm_num_queries(const csm_engine* e);

/* matching::Preprocessing (matching.h:45): matching orders / SymBi DAG and
 * index build (graphflow.cpp:31-34, symbi.cpp:108-111). */
int csm_preprocess(csm_engine* e);

/* matching::InitialMatching (matching.h:46): embeddings of query 0 in the
 * initial graph. */
int csm_initial(csm_engine* e, uint64_t* num_results);

/* matching::Classify (matching.h:56; parallel_graphflow.cpp:1990-2016):
 * *safe = 1 iff no registered query edge matches (L(v1), L(v2), label) in
 * either orientation. */
int csm_classify(csm_engine* e, uint32_t v1, uint32_t v2, uint32_t label, int* safe);

/* matching::AddEdge (matching.h:48; parallel_graphflow.cpp:2058-2138):
 * inserts (v1, v2, label) and returns the new embeddings of query 0 in *pos
 * (summed over queries when several are registered).  One update per launch
 * sequence: the plain per-update GPU path. */
int csm_add_edge(csm_engine* e, uint32_t v1, uint32_t v2, uint32_t label, uint64_t* pos);

/* matching::RemoveEdge (matching.h:49; parallel_graphflow.cpp:2171-2250):
 * counts the embeddings containing (v1, v2) before removing it.  Deleting an
 * absent edge returns CSM_E_DELETE_MISSING (graph.cpp:98-102). */
int csm_remove_edge(csm_engine* e, uint32_t v1, uint32_t v2, uint64_t* neg);

/* matching::AddVertex / RemoveVertex (matching.h:50-51; graph.cpp:24-56). */
int csm_add_vertex(csm_engine* e, uint32_t id, uint32_t label);
int csm_remove_vertex(csm_engine* e, uint32_t id);

/* Batched form of AddEdge/RemoveEdge over a stream slice (no reference
 * counterpart: the reference applies updates one at a time in
 * main.cpp:253-397).  Results are exact per update and per query, as if the
 * updates were applied in order: row-major [update][query] arrays of
 * n * num_queries entries.  pos/neg receive the count of insertions/deletions
 * respectively (the other is 0); safe (n entries, nullable) is the Classify
 * result; emb_hash (nullable, needs want_hash) the embedding-set hash.
 * Host pointers. */
int csm_apply_batch(csm_engine* e, const csm_update* updates, size_t n, uint64_t* pos,
                    uint64_t* neg, uint8_t* safe, uint64_t* emb_hash);

/* Same with every pointer in device memory (HBM) and no host round trip;
 * `counts` is [n * num_queries] u64 (insertions and deletions together, the
 * update's type says which).  Asynchronous on the engine's stream; call
 * csm_sync before reading results. */
int csm_apply_batch_device(csm_engine* e, const csm_update* d_updates, size_t n,
                           uint64_t* d_counts, uint8_t* d_safe, uint64_t* d_hash);
int csm_sync(csm_engine* e);

/* Per update of the last csm_apply_batch / csm_apply_batch_device call (or
 * entry 0 after csm_initial): bit CSM_USTAT_TIME_LIMIT = the time limit
 * stopped part of its search (its counts are lower bounds); bit
 * CSM_USTAT_CAPPED = a count of it reached max_results; bit
 * CSM_USTAT_CERTIFIED = it passed the label filter (Classify) but was
 * certified before its search to change no count: for every seed, some other
 * query neighbour of a seed endpoint has no candidate with its labels among
 * that endpoint's neighbours (the safe-update classifier's second stage).
 * No reference counterpart (the reference reports one flag for the whole
 * run). */
#define CSM_USTAT_TIME_LIMIT 1
#define CSM_USTAT_CAPPED 2
#define CSM_USTAT_CERTIFIED 4
int csm_update_status(const csm_engine* e, uint8_t* status, size_t n);

/* Per update of the last csm_apply_batch_device call (needs want_work):
 * the candidate entries its search tested over all queries, a measure of the
 * size of its search tree.  csm_apply_batch keeps those of its last slice.
 * No reference counterpart (the reference's intermediate counters,
 * matching.h:28-33, are not printed). */
int csm_update_work(const csm_engine* e, uint64_t* work, size_t n);

/* The embeddings found by the last call that searched (csm_initial,
 * csm_add_edge / csm_remove_edge, csm_apply_batch[_device]), with
 * csm_opts.print_enum set: the output of --print-enum, which the reference
 * prints as each search reaches it (SymBi/symbi.cpp:564-569,
 * Parallel_GraphFlow/parallel_graphflow.cpp:367).  Row r of `rows` holds
 * *width words: the update's index in the call (0 for csm_initial), the
 * query id, then the data vertex matched to each query vertex 0..k-1, and
 * 0xFFFFFFFF past the query's k.  Rows come in no particular order.  Up to
 * cap_rows rows are copied (*n_rows); *n_found counts every embedding found,
 * more than the rows kept once csm_opts.enum_cap is reached.  rows = NULL:
 * only the counts and *width. */
int csm_embeddings(const csm_engine* e, uint32_t* rows, size_t cap_rows, size_t* n_rows,
                   uint64_t* n_found, uint32_t* width);

/* Sets csm_opts.time_limit_s for the following calls (the CLI hands each
 * call what is left of the run's --time-limit).  No reference counterpart. */
int csm_set_time_limit(csm_engine* e, double seconds);

/* Per query, for the last batch: bit 1 if its search stopped because its
 * match budget was exceeded, bit 2 if the time limit dropped part of its
 * search (its counts for that batch are then partial).  No reference
 * counterpart (the reference's --max-results caps one update). */
int csm_query_status(const csm_engine* e, uint8_t* aborted, size_t n);

/* matching::GetNumInitialResults/Positive/Negative (matching.h:60-62). */
int csm_totals(const csm_engine* e, uint64_t* initial, uint64_t* positive, uint64_t* negative);

/* matching::GetMemoryCost (matching.h:53): edges and vertices held. */
int csm_memory(const csm_engine* e, size_t* num_edges, size_t* num_vertices);

/* Device bytes currently held by the engine (admission control).  Reads one
 * word, so a serving thread may call it while a batch call runs. */
int csm_device_bytes(const csm_engine* e, uint64_t* bytes);

/* Engine statistics (cumulative): [0] search tasks run, [1] candidate
 * entries tested, [2] algorithmic bytes of the search (4 B per candidate
 * entry scanned plus 4 B per entry of every partner segment, merge model),
 * [3] batches, [4] pool repacks, [5] search-kernel time in ns (HIP events on
 * the engine stream), [6] search launches timed, [7] most search rounds one
 * batch needed, [8] split_chunk and [9] spill budget in use for the first
 * registered query's size class, [10] 1 while the autotuner (auto_tuning
 * != 0 with neither knob set) is still trying candidates on large batches,
 * [11] batches replayed as HIP graphs, [12 + 2c] split_chunk and [13 + 2c]
 * spill budget of size class c: queries of up to 8, 12, 16 and 32 vertices
 * (c = 0..3); the autotuner picks each class's pair separately. */
int csm_stats(const csm_engine* e, uint64_t* out, size_t n_out);

/* Diagnostic read-back of the candidate index of one query (parallel_symbi
 * and parallel_turboflux engines), the state behind Parrllel_SymBi's counters
 * (Parallel_SymBi/parallel.h:67-76; symbi.cpp:246-377) and Parallel_TurboFlux's
 * DCG over its spanning tree (parallel_turboflux.h:34-55,
 * parallel_turboflux.cpp:326-469): what = 0 DAG/tree root
 * (out[0]); 1 d1, 2 d2 (0/1 per query vertex and data vertex, [k][nv]);
 * 3 np1, 4 nc2 ([k][nv]); 5 n1 (per query vertex, per parent, [nv] each);
 * 6 n2 (per query vertex, per neighbour in id order, [nv] each). */
int csm_index_dump(csm_engine* e, uint32_t qid, uint32_t what, uint32_t* out, size_t n);

/* The matching orders of query qid, one per query edge (the query vertex at
 * each position; positions 0 and 1 are the edge's endpoints): out receives
 * *n_orders x k entries (out == NULL: only *n_orders).  For
 * parallel_turboflux with reference_orders these follow TurboFlux's rule
 * (Parallel_TurboFlux::GenerateMatchingOrder, parallel_turboflux.cpp:533-700,
 * which the reference prints with --print-prep), and paths (k entries,
 * nullable) receives its explicit-path counts per query vertex
 * (CountDownwards, :469-491); otherwise paths are 0.  For parallel_symbi
 * with reference_orders only positions 0 and 1 (the seed edge) are used:
 * the search picks every later vertex as it goes (SymBi's rule,
 * parrallel_symbi.cpp:4231-4262). */
int csm_query_orders(csm_engine* e, uint32_t qid, uint32_t* out, size_t n, uint32_t* paths,
                     size_t* n_orders);

/* Device and pinned-host buffers on the engine's device and runtime, for
 * callers that keep inputs resident in HBM without a HIP runtime of their
 * own (one HIP runtime per process).  Device bytes count towards
 * csm_device_bytes.  Copies are synchronous on the engine stream.  No
 * reference counterpart. */
int csm_device_alloc(csm_engine* e, size_t bytes, void** ptr);
int csm_device_free(csm_engine* e, void* ptr);
int csm_host_alloc(csm_engine* e, size_t bytes, void** ptr);
int csm_host_free(csm_engine* e, void* ptr);
int csm_copy_to_device(csm_engine* e, void* dst, const void* src, size_t bytes);
int csm_copy_to_host(csm_engine* e, void* dst, const void* src, size_t bytes);
int csm_copy_device(csm_engine* e, void* dst, const void* src, size_t bytes);

/* The HIP stream the engine launches on (hipStream_t). */
void* csm_stream(csm_engine* e);

/* Overlapping copies with the engine's work (serving).  A second HIP stream,
 * the transfer stream, carries asynchronous copies; markers (slots
 * 0..CSM_MARKS-1) order the two streams: csm_mark records a marker on a
 * stream, csm_wait_mark makes a stream wait for it on the device, and
 * csm_sync_mark waits for it on the host.  Host memory passed to
 * csm_copy_async must be pinned (csm_host_alloc).  These calls touch only the
 * transfer stream and the markers, so a second host thread may make them
 * while a batch call runs, provided the two threads use different slots.
 * No reference counterpart (the reference's driver is a serial loop,
 * main.cpp:253-397). */
#define CSM_STREAM_ENGINE 0
#define CSM_STREAM_TRANSFER 1
#define CSM_MARKS 16
int csm_copy_async(csm_engine* e, void* dst, const void* src, size_t bytes);
int csm_mark(csm_engine* e, int stream, uint32_t slot);
int csm_wait_mark(csm_engine* e, int stream, uint32_t slot);
int csm_sync_mark(csm_engine* e, uint32_t slot);

/* Free and total device memory (hipMemGetInfo) on the engine's device: the
 * serving loop's HBM admission.  No reference counterpart. */
int csm_mem_info(csm_engine* e, uint64_t* free_bytes, uint64_t* total_bytes);

/* Text ingestion on the device.  Parses the reference's input formats, one
 * record per line, into csm_update records:
 *   CSM_PARSE_GRAPH  - query / data graph files (Graph::LoadFromFile,
 *                      graph.cpp:217-249): the type is the line's first
 *                      char; "t ..." headers skipped, "v id label" -> type
 *                      'v', any other char "x a b label" -> type 'e'
 *                      (is_add = 1);
 *   CSM_PARSE_STREAM - update streams (Graph::LoadUpdateStream,
 *                      graph.cpp:251-285): the type is the first token;
 *                      "v"/"-v id label" -> type 'v' (is_add = token "v"),
 *                      any other token "tok a b label" -> type 'e' with
 *                      is_add = (tok == "e").
 * Blank lines are skipped; ids and labels are decimal uint32.  Returns
 * CSM_E_ARG for a malformed line (its 1-based number in *err_line) and
 * CSM_E_CAPACITY when more than `cap` records result (*n holds the count,
 * so cap = 0 sizes the output).  Replaces the host iostream parsing of
 * Graph::LoadFromFile / LoadUpdateStream; no reference counterpart as an
 * entry point. */
#define CSM_PARSE_GRAPH 0
#define CSM_PARSE_STREAM 1

/* Host text in, host records out, on device `device` (no engine needed, so
 * a data graph can be parsed before csm_create). */
int csm_parse_text(int device, int mode, const char* text, size_t len, csm_update* out,
                   size_t cap, size_t* n, uint64_t* err_line);

/* Text already in HBM (16-byte aligned) to device records, on the engine's
 * stream; scratch counts towards csm_device_bytes.  Stream records with edge
 * updates only can go straight to csm_apply_batch_device. */
int csm_parse_text_device(csm_engine* e, int mode, const char* d_text, size_t len,
                          csm_update* d_out, size_t cap, size_t* n, uint64_t* err_line);

/* csm_parse_text_device with no device allocation.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its working space is a fixed scratch array of 12 MiB per GPU
 * inside the library, so a call takes at most 8 MiB of text and 262,144
 * lines (larger inputs: CSM_E_CAPACITY, nothing parsed).  Calls on one GPU
 * share that array: calls that overlap on one GPU race and must be
 * serialised by the caller.  Results equal csm_parse_text_device's. */
int csm_parse_text_device_static(csm_engine* e, int mode, const char* d_text, size_t len,
                                 csm_update* d_out, size_t cap, size_t* n, uint64_t* err_line);

const char* csm_last_error(const csm_engine* e);
void csm_destroy(csm_engine* e);

#ifdef __cplusplus
}
#endif
#endif /* CSM_C_H */
