/*
 * updec.h -- C ABI of the MI355X-native decimal expression path.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory
 * owned by the caller; every call is ordered on the given HIP stream
 * (NULL = the null stream).  Functions return UPDEC_OK (0) or an error code;
 * updec_last_error() gives the message of the calling thread's last error.
 * No C++ exception crosses this boundary.
 *
 * Each entry point names the reference interface it replaces.  Paths are
 * relative to the reference repository.
 */
#ifndef UPDEC_H
#define UPDEC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UPDEC_OK 0
#define UPDEC_EINVAL 1       /* bad argument */
#define UPDEC_ECOMPILE 2     /* parse, rewrite or hipRTC compile failure */
#define UPDEC_EHIP 3         /* HIP runtime error */
#define UPDEC_EUNSUPPORTED 4 /* construct outside what the HIP path implements */
#define UPDEC_EFULL 5        /* batch has no room for this query: launch it first */

#define UPDEC_MODE_AUTO 0     /* faithful inside the reference's envelope, else extended */
#define UPDEC_MODE_FAITHFUL 1 /* 32-limb wrap and DivOrMod path quirks */
#define UPDEC_MODE_EXTENDED 2 /* same rules, exact, no wrap */

typedef void* updec_stream_t; /* hipStream_t */
typedef struct updec_expr_s* updec_expr_t;
typedef struct updec_agg_s* updec_agg_t;
typedef struct updec_module_s* updec_module_t;

/* Column of the compact decimal layout: DECIMAL(precision, scale), row
 * stride updec_compact_bytes(precision).  `bits` is an optional statistic
 * (every stored magnitude < 2^bits, 0 = unknown) from updec_column_maxbits. */
typedef struct updec_column_desc {
    const char* name;
    int32_t precision;
    int32_t scale;
    int32_t bits;
} updec_column_desc;

const char* updec_last_error(void);
int updec_version(void);

/* Select the device (hipSetDevice) and warm the runtime. */
int updec_init(int device);

/* Device-memory and stream plumbing for hosts without their own HIP
 * runtime binding (the Python layer uses these; C++ hosts can use HIP
 * directly).  No counterpart in the reference, whose buffers are CUDA
 * managed memory (src/CudaAcc/AriesEngineDef.h:72-122). */
int updec_malloc(void** d_ptr, size_t bytes);
int updec_free(void* d_ptr);
/* stream-ordered allocation (hipMallocAsync / hipFreeAsync): no device-wide
 * synchronization, for buffers whose lifetime follows one stream's work */
int updec_malloc_async(void** d_ptr, size_t bytes, updec_stream_t stream);
int updec_free_async(void* d_ptr, updec_stream_t stream);
int updec_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, updec_stream_t stream);
int updec_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, updec_stream_t stream);
int updec_memset(void* d_ptr, int value, size_t bytes, updec_stream_t stream);
int updec_stream_create(updec_stream_t* stream);
int updec_stream_destroy(updec_stream_t stream);
int updec_stream_sync(updec_stream_t stream);
int updec_device_sync(void);
/* device properties: compute units, clock kHz, HBM bytes */
int updec_device_info(int32_t* cus, int32_t* clock_khz, int64_t* mem_bytes, char* name, size_t name_cap);
/* free and total device memory of the current device (hipMemGetInfo), for
 * admission control */
int updec_mem_info(int64_t* free_bytes, int64_t* total_bytes);
/* event timing helpers: record on a stream, elapsed ms between two events */
int updec_event_create(void** ev);
int updec_event_record(void* ev, updec_stream_t stream);
int updec_event_elapsed(void* ev0, void* ev1, float* ms);
/* *done = 1 when the work before the event has finished (non-blocking) */
int updec_event_query(void* ev, int32_t* done);
int updec_event_sync(void* ev);
int updec_event_destroy(void* ev);

/* Plain streaming copy kernel (16-byte vector loads/stores), used to
 * calibrate HBM counters and bandwidth; bytes must be a multiple of 16 and
 * both pointers 16-byte aligned.  No counterpart in the reference. */
int updec_copy_kernel(void* d_dst, const void* d_src, size_t bytes, updec_stream_t stream);

/* Issue-rate probe of the limb multiply (32 x 32 -> 64 multiply-add,
 * v_mad_u64_u32): blocks x 256 lanes, 8 independent chains of `iters` steps
 * per lane; returns the kernel time and the multiply-adds it did
 * (synchronous).  The measured peak the multiply-bound kernels are compared
 * with.  No counterpart in the reference. */
int updec_mad_peak(int32_t blocks, int32_t iters, double* ms_out, double* macs_out);

/* Compact bytes per row of DECIMAL(P, *): GetDecimalRealBytes
 * (src/datatypes/decimal.cu:464-467, table at :92). */
int updec_compact_bytes(int precision);

/* Largest magnitude bit length over a compact column (synchronous). */
int updec_column_maxbits(const uint8_t* d_col, int precision, int64_t rows, int32_t* bits_out, updec_stream_t stream);

/* ---------------- expression path ----------------
 * Replaces AriesEngineShell::AlignOptimize (src/AriesEngineWrapper/AriesEngineShell.cpp:646-685),
 * AEExprCalcNode::GenerateCudaFunction (src/AriesEngine/AriesExprCalcNode.cpp:1285-1318) and
 * AriesDynamicKernelManager::CompileKernels (src/CudaAcc/DynamicKernel.cu:60-158):
 * parse `text` over the named columns, rewrite with the reference's
 * alignment scheduling and constant pre-calculation, type every node,
 * generate a gfx950 kernel and compile it with hipRTC (module cache keyed
 * by source text). */
int updec_expr_compile(const char* text, const updec_column_desc* cols, int ncols, int mode, updec_expr_t* out);
/* result type of the materialized column, its compact bytes, and whether
 * the faithful contract applies */
int updec_expr_info(updec_expr_t e, int32_t* precision, int32_t* scale, int32_t* out_bytes, int32_t* faithful);
/* algorithmic work of one row: 32 x 32 -> 64 limb products at the plan's
 * static widths (products, the division's dividend scaling and Knuth D,
 * alignment multiplies; a word division counted as one) */
int updec_expr_work(updec_expr_t e, int64_t* limb_products_per_row);
/* canonical text of the rewritten, typed plan (NUL-terminated; returns the
 * needed size in *needed) */
int updec_expr_plan(updec_expr_t e, char* buf, size_t cap, size_t* needed);
int updec_expr_source(updec_expr_t e, char* buf, size_t cap, size_t* needed);
/* source of the column-iterator variant (updec_expr_calc_iter) */
int updec_expr_source_iter(updec_expr_t e, char* buf, size_t cap, size_t* needed);
/* Evaluate every row into compact output rows of out_bytes each:
 * AEExprDynKernelNode::Process / RunKernelFunction
 * (src/AriesEngine/AriesExprCalcNode.cpp:543-575) and
 * AriesDynamicKernelManager::CallKernel (src/CudaAcc/DynamicKernel.cu:281-334).
 * d_cols[i] is column i of the compile-time descriptor list; d_out must be
 * 4-byte aligned. */
int updec_expr_calc(updec_expr_t e, const uint8_t* const* d_cols, int64_t rows, uint8_t* d_out, updec_stream_t stream);
/* Lanes per row of updec_expr_calc: the reference's multi-lane (XMP) form of
 * calc_expr_<id>, one value spread over a group of lanes
 * (src/AriesEngine/AriesExprCalcNode.cpp:1320-1400, the TPI choice of
 * AriesCalcTreeGenerator.cpp:40-46).  An extended-mode expression whose
 * exact intermediates need W >= 64 limbs has it, with W / 4 or W / 8 lanes
 * per row (at most 64); 1 selects one row per lane (the default).
 * *width_out (optional) receives W, or 0 when there is no multi-lane form.
 * lanes_per_row = 0 only queries.  Results are identical in every form. */
int updec_expr_lanes(updec_expr_t e, int32_t lanes_per_row, int32_t* width_out);
/* source of the multi-lane calc kernel at lanes_per_row lanes per row */
int updec_expr_source_ml(updec_expr_t e, int32_t lanes_per_row, char* buf, size_t cap, size_t* needed);
void updec_expr_free(updec_expr_t e);

/* Block-column inputs: AriesColumnDataIterator (src/AriesColumnDataIterator.hxx:8-41,
 * AriesColumnDataIterator.cu:1-117).  A column is nblocks device blocks of
 * compact rows; block k holds rows [d_starts[k], d_starts[k+1]) (the last one
 * up to `rows`), d_starts[0] = 0, ascending.  Row i of a launch reads
 * position d_indices[i] of every column when d_indices is given (int32; -1 =
 * NULL, as a join's outer rows), else position i.  Both pointer arrays are
 * device memory. */
typedef struct updec_column_blocks {
    const uint8_t* const* d_blocks;
    const int64_t* d_starts;
    int32_t nblocks;
    int64_t rows;
} updec_column_blocks;
/* updec_expr_calc over block columns and indices: output row i is the
 * expression at position d_indices[i] (or i).  A NULL position, or one at or
 * past a column's rows, gives a zero row and d_valid[i] = 0 (d_valid: one
 * byte per row, 1 = value; may be NULL).  d_out 4-byte aligned. */
int updec_expr_calc_blocks(updec_expr_t e, const updec_column_blocks* cols, const int32_t* d_indices, int64_t rows,
                           uint8_t* d_out, uint8_t* d_valid, updec_stream_t stream);

/* Column iterators: AriesColumnDataIterator (src/AriesColumnDataIterator.hxx:8-41)
 * with its operator[] (src/AriesColumnDataIterator.cu:1-117), one per column
 * of a launch.  The column's data is nblocks device blocks of rows; block k
 * holds positions [d_starts[k], d_starts[k+1]) (the last one up to
 * `positions`), d_starts[0] = 0, ascending.  A row is the compact decimal
 * (B(P) bytes), or with `nullable` the reference's nullable_type<CompactDecimal>:
 * a flag byte (0 = NULL) then the B(P) compact bytes.  Without an index,
 * launch row i reads position i.  With an index, launch row i reads the
 * position the index holds at i: UPDEC_INDEX_INT32 entries are int32 (-1 =
 * NULL), UPDEC_INDEX_NULLABLE_INT32 entries are nullable_type<int32_t> (5
 * packed bytes: flag, then the int32; flag 0 = NULL); the index is one array
 * (nindex_blocks = 1, d_index, 4-byte aligned for int32) or nindex_blocks
 * blocks (d_index_blocks, with block k starting at launch row
 * d_index_starts[k]; device arrays).  A NULL index entry, a position outside
 * [0, positions) or a NULL row gives a NULL value.  Every pointer is
 * device memory. */
#define UPDEC_INDEX_NONE 0
#define UPDEC_INDEX_INT32 1
#define UPDEC_INDEX_NULLABLE_INT32 2
typedef struct updec_column_iter {
    const uint8_t* const* d_blocks;
    const int64_t* d_starts;
    int32_t nblocks;
    int32_t nullable;
    int64_t positions;
    int32_t index_kind;
    int32_t nindex_blocks;
    const void* d_index;
    const void* const* d_index_blocks;
    const int64_t* d_index_starts;
} updec_column_iter;
/* updec_expr_calc over column iterators, the reference's calc_expr_<id>
 * with its input iterators (AriesExprCalcNode.cpp:1292-1296, load and store
 * code :76-108, :1402-1425).  A result is NULL when a column it reads is NULL,
 * or when a division or MOD with an operand that can be NULL (a column with
 * nullable rows or an index, or an expression over one) has a zero divisor
 * (nullable_type operator/ and %, src/datatypes/aries_types.hxx:626-663).
 * Output row i: with out_nullable = 0, B(P) compact bytes (zero for NULL);
 * with out_nullable = 1, the reference's nullable output row, a flag byte
 * then B(P) compact bytes (zero for NULL).  d_valid (optional): one byte per
 * row, 1 = value, 0 = NULL.  d_out 4-byte aligned. */
int updec_expr_calc_iter(updec_expr_t e, const updec_column_iter* cols, int64_t rows, uint8_t* d_out,
                         int32_t out_nullable, uint8_t* d_valid, updec_stream_t stream);

/* ---------------- fused aggregation ----------------
 * Replaces AEExprAggFunctionNode::RunKernelFunction
 * (src/AriesEngine/AriesExprCalcNode.cpp:1623-1700) with
 * AggregateColumnData(SUM/COUNT) (src/CudaAcc/AriesSqlOperator_group.cu:360-424)
 * for the case the planner can fuse: SUM(expr_k) for k < nexprs over rows
 * passing an optional int32 filter, grouped by up to two 1-byte key columns.
 * Groups are numbered in ascending key order (the order SortColumns gives,
 * AriesSqlOperator_group.cu:14-54): one pass with per-block key tables
 * merged on the device, or, when a block meets more than 16 keys, a key
 * discovery pass and a host-numbered key table first.
 * Output: images[g * nexprs + k] is the 133-byte nullable_type<Decimal>
 * image of SUM(expr_k) in group g (AVG images follow, see avg_mask);
 * counts[g] = COUNT(*) of group g; keys_out[g] = key0 | key1 << 8 of group
 * g (host memory).  An aggregation object keeps device state between runs
 * (the one-pass form's key tables are cleared by the previous run), so it
 * serves one run at a time: callers that run one aggregation from several
 * threads at once compile one object per thread. */
typedef struct updec_agg_desc {
    const updec_column_desc* cols;
    int32_t ncols;
    const char* const* exprs;
    int32_t nexprs;
    int32_t has_filter; /* filter on the int32 column passed to updec_agg_run */
    int32_t filter_op;  /* 0 <=, 1 <, 2 >=, 3 >, 4 ==, 5 != */
    int64_t filter_value;
    int32_t nkeys; /* 0, 1 or 2 one-byte key columns */
    int32_t mode;
    /* bit k set: also AVG(expr_k) = SUM(expr_k) / COUNT(*) per group by the
     * DivisionInt64 rule (as updec_div_int64 with this mode), written after
     * the SUM images: d_images then holds max_groups * (nexprs + navg)
     * images, AVG j of group g at index max_groups * nexprs + g * navg + j
     * (navg = bits set, j counting them from bit 0) */
    uint32_t avg_mask;
} updec_agg_desc;

int updec_agg_compile(const updec_agg_desc* desc, updec_agg_t* out);
int updec_agg_run(updec_agg_t a, const uint8_t* const* d_cols, const int32_t* d_filter, const uint8_t* d_key0,
                  const uint8_t* d_key1, int64_t rows, uint8_t* d_images, int64_t* d_counts, int32_t max_groups,
                  int32_t* ngroups_out, uint16_t* keys_out, updec_stream_t stream);
/* device time of the last run's main aggregation kernel, in milliseconds
 * (HIP events on the run's stream; synchronizes) */
int updec_agg_last_kernel_ms(updec_agg_t a, float* ms);
/* generated source for a given group count of the two-pass form, or of the
 * one-pass form with ngroups = 0 (no device needed) */
int updec_agg_source(updec_agg_t a, int32_t ngroups, char* buf, size_t cap, size_t* needed);
/* source of the column-iterator variant (updec_agg_run_iter) for ngroups >= 1 */
int updec_agg_source_iter(updec_agg_t a, int32_t ngroups, char* buf, size_t cap, size_t* needed);
void updec_agg_free(updec_agg_t a);
/* updec_agg_run over block columns and indices, for aggregations without
 * filter and keys (one group): SUM of each expression over the positions
 * d_indices[0..rows) (or 0..rows); NULL positions (-1) and positions past a
 * column add to no SUM and to no COUNT.  d_images: nexprs 133-byte images,
 * then one AVG image per avg_mask bit; d_counts: one int64. */
int updec_agg_run_blocks(updec_agg_t a, const updec_column_blocks* cols, const int32_t* d_indices, int64_t rows,
                         uint8_t* d_images, int64_t* d_counts, updec_stream_t stream);
/* updec_agg_run over column iterators, with the descriptor's filter and
 * keys: d_filter, d_key0 and d_key1 hold one value per launch row (already
 * gathered, as the filter and group production of the reference hand
 * them).  SUM(expr_k) adds the non-NULL values of expr_k (agg_sum_t of
 * nullable_type, src/CudaAcc/algorithm/operators.hxx:1175-1195); counts[g]
 * is COUNT(*) of the rows of group g that pass the filter.  Groups come
 * from the two-pass form (key discovery, then one pass).  AVG through
 * avg_mask needs non-NULL counts, so it is refused (UPDEC_EUNSUPPORTED) when
 * a column can be NULL. */
int updec_agg_run_iter(updec_agg_t a, const updec_column_iter* cols, int64_t rows, const int32_t* d_filter,
                       const uint8_t* d_key0, const uint8_t* d_key1, uint8_t* d_images, int64_t* d_counts,
                       int32_t max_groups, int32_t* ngroups_out, uint16_t* keys_out, updec_stream_t stream);

/* AVG = SUM / COUNT per group: DivisionInt64 (src/CudaAcc/AriesSqlOperator_helpers.cu:717-764),
 * division_data_int64 (src/CudaAcc/AriesEngineAlgorithm.cu:9476-9535), do_div
 * (src/CudaAcc/algorithm/kernel_util.hxx:427-438): Decimal(sum) /= Decimal((int64)count).
 * d_sums: n images of 133 bytes at a stride of stride_bytes. */
int updec_div_int64(const uint8_t* d_sums, int64_t stride_bytes, const int64_t* d_counts, uint8_t* d_out, int64_t n,
                    int mode, updec_stream_t stream);

/* Segmented SUM of a compact column in `associated` order with segment
 * starts `groups`: aggregate_column<CompactDecimal> SUM
 * (src/CudaAcc/AriesEngineAlgorithm.h:832-857) through segreduce
 * (src/CudaAcc/algorithm/kernel_segreduce.hxx:535-612).  Position i reads
 * row d_associated[i] (d_associated may be NULL: row i); segment s sums
 * positions [groups[s], groups[s+1]) (the last up to `rows`); positions
 * before groups[0] are in no segment.  Starts must ascend (as the
 * reference's group production gives them); they are read on the device
 * only, so descending starts give meaningless sums but no access outside
 * the buffers.  d_out: ngroups 133-byte images (prec P, frac S, +0 for an
 * empty segment).  Rows of at most 128 bytes (the 32-limb Decimal) and at
 * most 2^31 positions.  Stream-ordered; allocates its accumulators with
 * hipMallocAsync on the stream. */
int updec_segsum_compact(const uint8_t* d_col, int precision, int scale, int64_t rows, const int32_t* d_associated,
                         const int32_t* d_groups, int32_t ngroups, uint8_t* d_out, int mode, updec_stream_t stream);
/* updec_segsum_compact for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and kernels;
 * the accumulators live in a fixed 8 MiB device array of the library (one
 * per GPU, held until the process ends, its address looked up by
 * updec_init or by the first call on a device), so the call allocates
 * nothing.  At most updec_segsum_noalloc_max_segments(precision) segments
 * (ngroups * (W + 2) 64-bit words, W = 2, 4, 8, 16 or 32 limbs by B(P)):
 * beyond that UPDEC_EINVAL, nothing launched.  Every call on a GPU uses the
 * same array: calls that overlap on one GPU (different streams or threads)
 * race and give wrong sums; the array is cleared by each call before use. */
int updec_segsum_compact_noalloc(const uint8_t* d_col, int precision, int scale, int64_t rows,
                                 const int32_t* d_associated, const int32_t* d_groups, int32_t ngroups, uint8_t* d_out,
                                 int mode, updec_stream_t stream);
int updec_segsum_noalloc_max_segments(int precision, int32_t* max_segments);

/* ---------------- codec: ingest and output of compact columns ----------------
 * Bulk ingest of one text column into DECIMAL(P, S) compact rows: the
 * per-value step of the CSV import (src/server/mysql/importCsvFile.cpp:640-660
 * and :1883-1897), i.e. Decimal(P, S, text) (src/datatypes/decimal.cu:154-165:
 * StringToDecimal :2980-3025, then cast :471-533) and ToCompactDecimal
 * (:337-341).  Row r is d_chars[d_offsets[r] .. d_offsets[r+1]) (rows + 1
 * offsets), cut at its first NUL.  The reference's character rules are kept
 * for every byte (a non-digit c still contributes c - '0'); its reading past
 * the terminator after a trailing '.' is not: a trailing '.' ends the row.
 * The reference reports no error on this path (its overflow check is
 * disabled, decimal.cu:400-425), so neither does this one.  B(P) must be
 * at most 128 bytes (the reference's 32-limb Decimal). */
int updec_parse_compact(const char* d_chars, const int64_t* d_offsets, int64_t rows, int32_t precision,
                        int32_t scale, uint8_t* d_out, updec_stream_t stream);

/* Bulk output of a compact column into fixed stride-byte slots, each
 * NUL-terminated and zero-filled.  UPDEC_FORMAT_HEX is GetDecimal
 * (src/datatypes/decimal.cu:362-397) of Decimal(compact, P, S) (:167-176),
 * the string AriesDataBuffer::GetCompactDecimalAsString returns
 * (src/CudaAcc/AriesEngineDef.cu:817-829): "+0000000C(002)".
 * UPDEC_FORMAT_TEXT is the decimal text with exactly S fractional digits
 * ("-123.4500"), the form the reference's decimal tests print
 * (test/testcases/datatypes/test_decimal.cu).  stride: a multiple of 4 of
 * at least updec_format_stride(); d_out 4-byte aligned. */
#define UPDEC_FORMAT_HEX 0
#define UPDEC_FORMAT_TEXT 1
int updec_format_stride(int32_t precision, int32_t scale, int32_t form, int64_t* stride);
int updec_format_compact(const uint8_t* d_col, int64_t rows, int32_t precision, int32_t scale, int32_t form,
                         char* d_out, int64_t stride, updec_stream_t stream);

/* ---------------- dynamic-code interface ----------------
 * AriesDynamicKernelManager::CompileKernels / CallKernel
 * (src/CudaAcc/DynamicKernel.h:50-99): compile caller-supplied kernel
 * source (which may #include "dec_device.h") for gfx950 with hipRTC and
 * launch an extern "C" kernel by name with one by-value argument block. */
int updec_compile(const char* src, size_t len, updec_module_t* out);
int updec_launch(updec_module_t m, const char* fn, uint32_t grid, uint32_t block, const void* args, size_t args_size,
                 updec_stream_t stream);
void updec_module_free(updec_module_t m);
/* hipRTC-compile a source for gfx950 without loading it (no device needed);
 * returns UPDEC_ECOMPILE with the compiler log on failure */
int updec_check_compile(const char* src, size_t len);
/* Module cache counters since the process started: hipRTC compiles, hits in
 * the in-memory LRU (100 modules, keyed by source text as the reference's
 * DynamicKernel.cu:10-58) and loads from the on-disk store of compiled code
 * objects ($UPDEC_KERNEL_CACHE, "off" disables; else
 * $XDG_CACHE_HOME/updec or ~/.cache/updec).  Generated sources carry no
 * constant values (constants are kernel data, the reference's constValues),
 * so expressions that differ only in literals of the same limb widths share
 * one compile. */
int updec_module_cache_stats(int64_t* rtc_compiles, int64_t* memory_hits, int64_t* disk_hits);

/* ---------------- interpreter and batches ----------------
 * The same plans as the expression path, run by an interpreter compiled once
 * with the library (no per-query hipRTC compile): the alternative to
 * AriesDynamicKernelManager::CompileKernels / CallKernel
 * (src/CudaAcc/DynamicKernel.h:50-99) for a stream of distinct queries, with
 * the SUM / COUNT of AEExprAggFunctionNode::RunKernelFunction
 * (src/AriesEngine/AriesExprCalcNode.cpp:1623-1700) over all rows.
 *
 * updec_vm_compile: parse, rewrite and type `text` exactly as
 * updec_expr_compile does, then build a register program.  Every register
 * holds W limbs, W in {2, 4, 8, 16, 32, 64, 128, 256, 512} the smallest
 * that holds every intermediate exactly; UPDEC_EUNSUPPORTED when none does.
 * Up to 32 limbs one lane holds a value; the 64- to 512-limb classes
 * (extended-mode programs only) spread a value over a group of lanes (the
 * reference's multi-lane XMP form, AriesExprCalcNode.cpp:1320-1400).
 *
 * A batch collects queries (program, device columns, rows, output image and
 * count) and runs them with one set of launches: persistent kernels per
 * register width take row ranges from a work counter; with use_graph the
 * launches are one HIP graph (captured at the first launch).  Results are
 * written to each query's 133-byte SUM image and int64 COUNT.  A batch is
 * not thread-safe across launches on different streams: use one stream per
 * batch object.  No counterpart in the reference (its queries launch one
 * kernel each, synchronously). */
typedef struct updec_vm_s* updec_vm_t;
typedef struct updec_batch_s* updec_batch_t;
int updec_vm_compile(const char* text, const updec_column_desc* cols, int ncols, int mode, updec_vm_t* out);
int updec_vm_info(updec_vm_t v, int32_t* width, int32_t* nins, int32_t* nregs, int32_t* precision, int32_t* scale,
                  int32_t* faithful);
int updec_vm_listing(updec_vm_t v, char* buf, size_t cap, size_t* needed);
void updec_vm_free(updec_vm_t v);
int updec_batch_create(int32_t max_queries, int32_t max_items, updec_batch_t* out);
/* per width class (0..4 for W = 2..32): rows per work item and resident
 * blocks per compute unit of the persistent kernel */
int updec_batch_tuning(updec_batch_t b, int32_t width_class, int64_t rows_per_item, int32_t blocks_per_cu);
/* lanes per value of a multi-lane class (width_class 5..8, W = 64 .. 512
 * limbs): W / 4 (4 limbs per lane, the default) or W / 8 (8 limbs per lane),
 * at most 64 (the 512-limb class: 64 lanes of 8 limbs) */
int updec_batch_lanes(updec_batch_t b, int32_t width_class, int32_t lanes);
/* *index_out (optional): the query's position in the pending batch.
 * UPDEC_EFULL (nothing staged) when the batch's query, work-item or program
 * space would overflow: launch the batch and add the query to the next. */
int updec_batch_add(updec_batch_t b, updec_vm_t v, const uint8_t* const* d_cols, int64_t rows, uint8_t* d_image,
                    int64_t* d_count, int32_t* index_out);
/* number of queries added since the last launch */
int updec_batch_pending(updec_batch_t b, int32_t* n);
int updec_batch_launch(updec_batch_t b, int use_graph, updec_stream_t stream);
void updec_batch_free(updec_batch_t b);

/* ---------------- seeded synthetic columns ----------------
 * Counter-based generator (splitmix64 keyed by seed, column id and row), so
 * any row range can be produced independently and identically on host and
 * device.  kind:
 *   UPDEC_GEN_RANGE:  magnitude = (lo + r mod (hi - lo + 1)) * step   (64-bit)
 *   UPDEC_GEN_DIGITS: magnitude uniform with exactly `digits` decimal digits
 *                     (leading digit nonzero: uniform in [10^(digits-1), 10^digits))
 *   UPDEC_GEN_BELOW:  magnitude uniform in [0, bound), bound given as
 *                     bound_words little-endian 32-bit words (<= 16)
 * neg_permille: probability of a negative sign in 1/1000. */
#define UPDEC_GEN_RANGE 0
#define UPDEC_GEN_DIGITS 1
#define UPDEC_GEN_BELOW 2
typedef struct updec_gen_desc {
    int32_t kind;
    int32_t precision;
    int32_t digits;
    int32_t neg_permille;
    uint64_t lo, hi, step;
    uint64_t seed;
    uint32_t column_id;
    int32_t bound_words;
    uint32_t bound[16];
} updec_gen_desc;
int updec_gen_compact(const updec_gen_desc* g, int64_t row0, int64_t rows, uint8_t* out, int on_device,
                      updec_stream_t stream);
/* int32 uniform in [lo, hi] and 1-byte values drawn from a set */
int updec_gen_int32(uint64_t seed, uint32_t column_id, int32_t lo, int32_t hi, int64_t row0, int64_t rows, int32_t* out,
                    int on_device, updec_stream_t stream);
int updec_gen_byte(uint64_t seed, uint32_t column_id, const uint8_t* choices, int32_t nchoices, int64_t row0,
                   int64_t rows, uint8_t* out, int on_device, updec_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
