/*
 * omni_gpu.h -- C ABI of the MI355X (gfx950) hash-aggregation and hash-join hot path.
 *
 * The reference library (OmniOperator) runs group-by aggregation and hash join
 * as single-threaded CPU operators behind
 *     Operator::AddInput(VectorBatch*) / Operator::GetOutput(VectorBatch**)
 * (core/src/operator/operator.h:22-264).  This header is the thin, plain-C
 * boundary under GPU-backed operators with the same surface
 * (omnioperator_amd/host/gpu_operators.h).  Every entry point takes plain
 * pointers and sizes; no C++ type and no exception crosses it.
 *
 * Conventions
 *  - Every function returns OMNI_GPU_OK (0) or a negative OMNI_GPU_ERR_* code;
 *    omni_gpu_last_error() returns the message of the last failure on the
 *    calling thread.
 *  - Column pointers passed to *_add / *_probe / *_submit are DEVICE pointers
 *    (memory from omni_gpu_malloc or any hipMalloc'd buffer of the current
 *    device).  Null bitmaps use the OmniVec layout: LSB-first, bit = 1 means
 *    NULL (core/src/vector/nulls_buffer.h:50-84, core/src/util/bit_util.h:62-71);
 *    a NULL `nulls` pointer means "no NULLs".
 *  - Type ids are the reference's DataTypeId values (core/src/type/data_type.h:40-65).
 *  - Aggregate function ids are the reference's FunctionType values
 *    (core/src/operator/util/function_type.h:11-27).
 *  - A handle is used by one host thread at a time (the reference's operators
 *    are single-threaded by contract); distinct handles may be used
 *    concurrently.  Each handle owns one HIP stream.
 */
#ifndef OMNI_GPU_H
#define OMNI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define OMNI_GPU_OK 0
#define OMNI_GPU_ERR_INVALID (-1)      /* bad argument or unsupported type/function combination */
#define OMNI_GPU_ERR_HIP (-2)          /* a HIP runtime call failed */
#define OMNI_GPU_ERR_NOMEM (-3)        /* device allocation failed */
#define OMNI_GPU_ERR_OVERFLOW (-4)     /* decimal overflow with overflowAsNull == false */
#define OMNI_GPU_ERR_OUTPUT_FULL (-5)  /* caller's output capacity too small; *needed reports the size */
#define OMNI_GPU_ERR_STATE (-6)        /* call out of order (e.g. probe before build finished) */

/* ---- type / function ids (values of the reference enums) -------------------------------------- */
#define OMNI_TYPE_INT 1
#define OMNI_TYPE_LONG 2
#define OMNI_TYPE_DOUBLE 3
#define OMNI_TYPE_BOOLEAN 4
#define OMNI_TYPE_SHORT 5
#define OMNI_TYPE_DECIMAL64 6
#define OMNI_TYPE_DECIMAL128 7
#define OMNI_TYPE_DATE32 8
#define OMNI_TYPE_TIMESTAMP 12
#define OMNI_TYPE_BYTE 18       /* int8; a group-by or join key column only */

#define OMNI_AGG_SUM 0
#define OMNI_AGG_COUNT_COLUMN 1
#define OMNI_AGG_COUNT_ALL 2
#define OMNI_AGG_AVG 3
#define OMNI_AGG_MAX 5
#define OMNI_AGG_MIN 6

/* Aggregation strategies (omni_gpu_agg_spec.strategy). */
#define OMNI_GPU_STRATEGY_AUTO 0
#define OMNI_GPU_STRATEGY_LDS 1      /* workgroup-local LDS table, spill to the global table */
#define OMNI_GPU_STRATEGY_GLOBAL 2   /* global open-addressing table, 64-bit CAS insert */
#define OMNI_GPU_STRATEGY_RADIX 3    /* radix-partition rows by key hash, then per-partition tables */
/* OMNI_GPU_STRATEGY_RADIX (many groups): the batches are recorded, and read when the groups are first asked for
 * (omni_gpu_agg_num_groups / omni_gpu_agg_get_output / omni_gpu_agg_wait), so their columns must stay valid until
 * one of those returns.  The rows are partitioned twice by bits of their key hash (2^8, then up to 2^11 children
 * each, chosen from a HyperLogLog estimate of the groups) and each final piece is aggregated in one workgroup's LDS
 * table, its groups written once -- no device-wide atomic per row.  It reorders rows, so it is refused for a
 * checked decimal SUM (an input or output that can leave its limits) and for double MIN / MAX, whose results
 * depend on the row order.  AUTO takes it for a capacity hint of 2^21 groups or more when it applies.  Batches
 * added after the groups were read go to the hash table (every existing group is inserted into it first). */

/* Join types and build sides: the values of OmniOperator's JoinType and BuildSide
 * (core/src/plannode/planNode.h:406-414 and :28-32). */
#define OMNI_GPU_JOIN_INNER 0
#define OMNI_GPU_JOIN_LEFT 1
#define OMNI_GPU_JOIN_RIGHT 2
#define OMNI_GPU_JOIN_FULL 3
#define OMNI_GPU_JOIN_LEFT_SEMI 4
#define OMNI_GPU_JOIN_LEFT_ANTI 5
#define OMNI_GPU_JOIN_EXISTENCE 6
#define OMNI_GPU_BUILD_UNKNOWN 0
#define OMNI_GPU_BUILD_LEFT 1
#define OMNI_GPU_BUILD_RIGHT 2

#define OMNI_GPU_HASH_TABLE_NORMAL 0
#define OMNI_GPU_HASH_TABLE_ARRAY 1

#define OMNI_GPU_MAX_KEYS 4
#define OMNI_GPU_MAX_AGGS 16

/* One input column: mirrors Vector<T> values + NullsBuffer of a flat OmniVec vector
 * (core/src/vector/vector.h:202-426).  Element i is values[i]. */
typedef struct omni_gpu_column {
    int32_t type_id;        /* DataTypeId */
    int32_t reserved;
    const void *values;     /* device pointer */
    const uint8_t *nulls;   /* device pointer to LSB-first null bitmap (bit=1 NULL) or NULL */
} omni_gpu_column;

/* One output column: the library writes values[0..n) and the null bitmap bits [0..n). */
typedef struct omni_gpu_out_column {
    int32_t type_id;
    int32_t reserved;
    void *values;           /* device pointer, n * sizeof(type) bytes */
    uint8_t *nulls;         /* device pointer, (n + 7) / 8 bytes; written in full */
} omni_gpu_out_column;

/* ---- runtime ----------------------------------------------------------------------------------- */
int omni_gpu_init(int device);
const char *omni_gpu_last_error(void);
int omni_gpu_malloc(void **ptr, size_t bytes);
int omni_gpu_free(void *ptr);
/* Copies and memset on device memory; each returns when the device has finished it (so that work queued on any
 * handle afterwards sees the result). */
int omni_gpu_memcpy_h2d(void *dst, const void *src, size_t bytes);
int omni_gpu_memcpy_d2h(void *dst, const void *src, size_t bytes);
int omni_gpu_memcpy_d2d(void *dst, const void *src, size_t bytes);
int omni_gpu_memset(void *dst, int value, size_t bytes);
/* Page-locked host memory (hipHostMalloc), for staging many small VectorBatches into one large copy; no reference
 * counterpart (the reference's operators never leave host memory). */
int omni_gpu_host_alloc(void **ptr, size_t bytes);
int omni_gpu_host_free(void *ptr);
int omni_gpu_device_synchronize(void);

/* Device memory the library holds: *in_use_bytes in live buffers of open handles, *cached_bytes in freed blocks it
 * keeps for reuse (returned to the runtime when an allocation would otherwise fail).  No reference counterpart; the
 * reference reports operator memory to its ThreadMemoryManager (core/src/memory/thread_memory_manager.h:45-58). */
int omni_gpu_memory_stats(int64_t *in_use_bytes, int64_t *cached_bytes);

/* Kernel timing: when enabled, every kernel the library launches is bracketed by HIP events on
 * the stream it runs on; omni_gpu_profile_get reports the launch count and summed duration (ms)
 * of the kernels whose name starts with `name_prefix`. */
int omni_gpu_profile_enable(int on);
int omni_gpu_profile_reset(void);
int omni_gpu_profile_get(const char *name_prefix, int64_t *launches, double *total_ms);

/* ---- hash aggregation ---------------------------------------------------------------------------
 * Replaces HashAggregationOperator's Emplace / ProcessGroup / Output path
 * (core/src/operator/aggregation/group_aggregation.cpp:320-406, 811-890, 892-955, 1486-1538).
 * The spec mirrors HashAggregationOperatorFactory's constructor arguments
 * (core/src/operator/aggregation/group_aggregation.h:233-245).  Raw input; final output, or the partial
 * output of agg_partial_out.  A final stage over partial input is composed by the operator layer from these
 * aggregates (omnioperator_amd/host/gpu_operators.cpp). */
typedef struct omni_gpu_agg_spec {
    int32_t num_keys;
    int32_t key_types[OMNI_GPU_MAX_KEYS];   /* any fixed-width type id above, BYTE included */
    int32_t num_aggs;
    int32_t agg_funcs[OMNI_GPU_MAX_AGGS];      /* OMNI_AGG_* */
    int32_t agg_in_types[OMNI_GPU_MAX_AGGS];   /* ignored for COUNT_ALL */
    int32_t agg_out_types[OMNI_GPU_MAX_AGGS];
    int32_t agg_in_scales[OMNI_GPU_MAX_AGGS];  /* decimal scales; 0 otherwise */
    int32_t agg_out_scales[OMNI_GPU_MAX_AGGS];
    int32_t overflow_as_null;                  /* OverflowConfig: NULL instead of error on overflow */
    int32_t strategy;                          /* OMNI_GPU_STRATEGY_* */
    int64_t capacity_hint;                     /* expected number of groups; 0 = unknown */
    /* Two-stage aggregation (outputPartial = true in the factory, group_aggregation.h:233-245): an AVG with
     * agg_partial_out[i] = 1 outputs its running double sum -- 0, not NULL, for a group whose inputs were all
     * NULL -- instead of the mean: the first column of AverageFlatIMAggregator's partial output
     * (average_flatim_aggregator.h:256-273); a COUNT_COLUMN over the same input gives the second.  A decimal SUM
     * with the flag outputs the sum column of SumSparkDecimalAggregator's partial output
     * (sum_spark_decimal_aggregator.h:246-301 with OUTPUT_PARTIAL): 0 for an empty group, NULL for an overflowed
     * one (the per-add or the rescale overflow) and never an error whatever overflow_as_null says; the isEmpty
     * column is a COUNT_COLUMN over the same input being 0.  The other functions' partial output equals their
     * final output, and the flag is ignored for them. */
    int32_t agg_partial_out[OMNI_GPU_MAX_AGGS];
    /* AggHashTableRule (core/src/util/policy.h:51: NORMAL or ARRAY; group_aggregation.cpp:320-360).  Under
     * OMNI_GPU_HASH_TABLE_ARRAY a single integer key whose first inserted rows span at most 2^21 - 2 values
     * gets a direct-index group map over that range widened by max(50, max - min) on each side
     * (VectorAnalyzer::CheckArrayMap, core/src/operator/hashmap/vector_analyzer.cpp:10-80); keys outside it use
     * the hash table.  Results do not depend on the rule. */
    int32_t hash_table_rule;
} omni_gpu_agg_spec;

int omni_gpu_agg_create(const omni_gpu_agg_spec *spec, void **handle);
/* Adds one VectorBatch: keys[num_keys] and aggs[num_aggs] (the entry of a COUNT_ALL aggregate
 * is ignored and may be zeroed).  Asynchronous on the handle's stream; the columns must stay
 * valid until the next synchronizing call on the handle. */
int omni_gpu_agg_add_batch(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                           int64_t rows);
/* Adds num_batches VectorBatches in one call: batch b's key columns are keys[b * num_keys ..], its aggregate
 * columns aggs[b * num_aggs ..] and its row count rows[b].  The batches are copied into one staging run of
 * columns by one launch (an aggregate input that is the same column as a key or an earlier aggregate in
 * every batch is staged once) and aggregated by one insert, instead of one launch per batch -- the
 * AddInput loop of HashAggregationOperator (group_aggregation.cpp:320-406) over many small batches.
 * Results are those of adding the batches one by one.  Same asynchrony and column lifetime as
 * omni_gpu_agg_add_batch. */
int omni_gpu_agg_add_batches(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                             const int64_t *rows, int32_t num_batches);
/* Blocks until every batch added so far has been consumed (its columns may then be reused). */
int omni_gpu_agg_wait(void *handle);
/* Number of groups so far, NULL-key group included (HashAggregationOperator::GetHashMapUniqueKeys,
 * group_aggregation.cpp:1168-1180).  Synchronizes the handle's stream. */
int omni_gpu_agg_num_groups(void *handle, int64_t *groups);
/* Writes groups [start, start+count) as output columns: keys first, then one column per
 * aggregate, in spec order (the output layout of Output<>, group_aggregation.cpp:1486-1538).
 * Synchronizes.  Returns OMNI_GPU_ERR_OVERFLOW if a decimal aggregate overflowed and
 * overflow_as_null == 0. */
int omni_gpu_agg_get_output(void *handle, int64_t start, int64_t count, omni_gpu_out_column *outs);
/* The table kind chosen so far: *array_map = 1 when the direct-index map of the ARRAY rule is in use, with its
 * key range [*range_min, *range_min + *range_cells) (the reference's VectorAnalyzer range). */
int omni_gpu_agg_table_info(void *handle, int32_t *array_map, int64_t *range_min, int64_t *range_cells);
int omni_gpu_agg_destroy(void *handle);

/* ---- batched stream of independent group-bys ----------------------------------------------------------
 * Replaces one HashAggregationOperator per VectorBatch (AddInput + GetOutput per batch,
 * group_aggregation.cpp:320-406, 454-474) for a stream of many small, uneven batches resident in HBM.
 * One launch per phase covers all batches: phase 1 aggregates each batch in an LDS table (one workgroup per
 * batch) and abandons the batches whose keys do not fit; phase 2 splits each abandoned batch by key hash into
 * partitions that fit an LDS table (one workgroup per partition); phase 3 aggregates exactly, in a private
 * HBM table, any batch one of whose partitions still overflowed.
 * No reference counterpart exists for the batching itself; per-batch results equal the reference operator's.
 * One key column of <= 8 bytes; any aggregate the per-batch path supports. */
int omni_gpu_agg_batches_create(const omni_gpu_agg_spec *spec, int64_t max_batch_rows, void **handle);
/* Launch tuning (the autotuner's knobs): threads per workgroup (256 or 512), LDS table bytes
 * ([8 KiB, 64 KiB]) and the LDS probe limit before a batch is handed to phase 2. */
int omni_gpu_agg_batches_tune(void *handle, int32_t block_threads, int32_t lds_bytes, int32_t max_probe);
/* Phase-2 tuning: LDS bytes of a partition table ([8 KiB, 64 KiB], default 64 KiB) and rows per partition
 * (0, the default: half the table's slots, so a partition overflows only far above its mean share; more
 * rows mean fewer passes over each batch and more batches left to phase 3).  Under flag bit 4 the table is the
 * workgroup's, and part_rows, when not 0, is the number of distinct keys a partition is sized for (default: 45%
 * of the table's slots, and a batch whose keys fit that many goes direct). */
int omni_gpu_agg_batches_tune_partitions(void *handle, int32_t part_lds_bytes, int32_t part_rows);
int omni_gpu_agg_batches_info(void *handle, int32_t *lds_slots, int64_t *pool_bytes);
/* Runs the stream.  keys[0] / aggs[i] are the concatenated device columns of all batches; batch b owns rows
 * [batch_offsets[b], batch_offsets[b+1]) (host array, batch_offsets[0] == 0, at most max_batch_rows rows
 * each).  batch_modes (host, optional): 0 = LDS table first, 2 = skip the per-batch LDS table and start
 * at phase 2 (e.g. from a predicted group count).  flags: bit 0 = one launch per batch and phase instead of
 * one launch per phase; bit 1 = capture the launch sequence in a hipGraph and replay it while the stream's
 * shape stays the same; bit 2 (without batch_modes, not with bit 0) = a sketch launch estimates each batch's
 * distinct keys from its first 1,024 rows and starts the batches estimated above the LDS table at phase 2;
 * bit 3 (same conditions) = the advisor network picks each batch's first phase
 * (omni_gpu_agg_batches_set_advisor); bit 4 (not with bit 0) = one workgroup per batch does the batch's whole
 * group-by in LDS (a batch that fits one table directly, a larger one by hash partitions it writes in partition
 * order and aggregates one after another; no device-wide atomics), in place of phases 1 and 2; with bit 3 the
 * advisor's predicted group count picks direct or partitioned and the partition count.  Batch b's groups go to output rows [batch_offsets[b], batch_offsets[b] + groups_per_batch[b]) of outs (key,
 * then aggregates); outs hold batch_offsets[num_batches] rows and NULL bitmaps of 4 * ceil(rows / 32) bytes.
 * Synchronizes. */
int omni_gpu_agg_batches_run(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                             const int64_t *batch_offsets, int32_t num_batches, const int32_t *batch_modes,
                             int32_t flags, omni_gpu_out_column *outs, int64_t *groups_per_batch);
/* The learned form of the stream: with flag bit 3 in omni_gpu_agg_batches_run, a per-batch sketch (an evenly
 * strided sample of at most 512 rows of each batch, the 16 features of omni_gpu_sketch) goes through the advisor
 * network `mlp` (an omni_gpu_mlp_create handle, trained on runs of this stream) predicts its group count, and a
 * batch predicted above half the per-batch LDS table's slots starts in the hash-partitioned passes.  The
 * launches stay on the device (no synchronisation), so they are captured with the rest under flag bit 1.  Results
 * do not depend on the choice.  mlp NULL detaches the advisor. */
int omni_gpu_agg_batches_set_advisor(void *handle, void *mlp);
/* Phase that finished each batch of the last run (1, 2 or 3; phase[num_batches], host). */
int omni_gpu_agg_batches_phase_of(void *handle, int32_t *phase);
/* The advisor's features of each batch of a stream (the sketch of flag bit 3): features[num_batches * 16] (host)
 * for the key column `key` (device) cut by batch_offsets (host).  Synchronizes. */
int omni_gpu_batch_features(const omni_gpu_column *key, const int64_t *batch_offsets, int32_t num_batches,
                            float *features);
/* Batches of the last run finished by the per-batch LDS table (phase 1) / by a later phase. */
int omni_gpu_agg_batches_stats(void *handle, int64_t *lds_batches, int64_t *hbm_batches);
/* Batches of the last run finished by each phase: 1 the per-batch LDS table, 2 the hash-partitioned LDS
 * passes, 3 the per-batch HBM table (the exact fallback; in the one-launch-per-batch form every batch that
 * overflowed phase 1). */
int omni_gpu_agg_batches_phases(void *handle, int64_t *phase1, int64_t *phase2, int64_t *phase3);
int omni_gpu_agg_batches_destroy(void *handle);

/* ---- hash join ------------------------------------------------------------------------------------
 * Replaces HashBuilderOperator + JoinHashTableVariants (core/src/operator/join/hash_builder.cpp:169-213,
 * join_hash_table_variants.cpp:254-283, 695-808) and LookupJoinOperator::ProbeBatchForInnerJoin
 * (core/src/operator/join/lookup_join.cpp:924-1035) with its output gather (:1657-1779). */
typedef struct omni_gpu_join_spec {
    int32_t key_type;                 /* single equality key (num_keys <= 1): INT, DATE32, SHORT, LONG, TIMESTAMP,
                                         DECIMAL64, DOUBLE (compared as its 64-bit pattern), BOOLEAN (0 or 1),
                                         BYTE, DECIMAL128 */
    int32_t num_payloads;             /* build-side output columns */
    int32_t payload_types[OMNI_GPU_MAX_AGGS];
    int32_t strategy;                 /* 0 = reference rule (array table if dense), 1 = hash only */
    /* Several equality key columns (buildHashCols of HashBuilderOperatorFactory, hash_builder.cpp:44-93): num_keys
     * in [2, OMNI_GPU_MAX_KEYS] fixed-width columns of key_types (0 or 1: the single key_type).  Keys match as the
     * reference's table compares them: BYTE / SHORT / INT / LONG / DECIMAL64 columns of at most 32, 64 or 128 bits in
     * total by their ComputeMultiColKey packing (join_hash_table_variants.cpp:105-146; each value is sign-extended
     * into the key before it is OR-ed in, so a negative later column overwrites the earlier ones and such tuples
     * match), other combinations by the bytes of all their values.  With num_keys >= 2 or a DECIMAL128 key, the
     * `key` argument of omni_gpu_join_build_add and the probe entry points points to num_keys columns (in
     * omni_gpu_join_probe_batches, batch b's columns are keys[b * num_keys ..]); a row with a NULL in any key column
     * matches nothing; the table is never an array table. */
    int32_t num_keys;
    int32_t key_types[OMNI_GPU_MAX_KEYS];
} omni_gpu_join_spec;

int omni_gpu_join_build_create(const omni_gpu_join_spec *spec, void **handle);
/* Appends one build VectorBatch (HashBuilderOperator::AddInput retains the batch): only its column
 * pointers are recorded; omni_gpu_join_build_finish gathers every batch into device-resident build
 * storage in one launch, so the caller's buffers must stay valid until it returns.
 * NULL keys are kept but never match. */
int omni_gpu_join_build_add(void *handle, const omni_gpu_column *key, const omni_gpu_column *payloads,
                            int64_t rows);
/* Builds the table (HashBuilderOperator::GetOutput).  Returns once every build batch has been copied into
 * the build storage, so the caller's build buffers may be released; the table build itself can still be
 * running on the handle's stream, and later calls on the handle are ordered after it. */
int omni_gpu_join_build_finish(void *handle);
/* Table facts after finish: *is_array (1 = direct-index array table), *max_dup (largest number of
 * build rows sharing one key), *table_slots. */
int omni_gpu_join_build_info(void *handle, int32_t *is_array, int64_t *max_dup, int64_t *table_slots);
/* Probe tuning (results do not depend on either knob).
 *   probe_group: table lookups each lane issues together before resolving any (1, 2, 4, 8 or 16; default 1).
 *   partitions:  radix-partitioned probe of hash tables: -1 = automatic (tables of at least 8 MiB probed
 *                with at least half as many rows as the table has slots are cut into regions of 8,192 slots),
 *                -2 = always with that region size, 0 = never, or a power of two in [8, 4096].  Probe rows
 *                are scattered into partition order; a region of <= 8,192 slots is staged in LDS by each
 *                workgroup that probes it, a larger one is probed from the L2 of the XCD that owns it.
 * No reference counterpart: the CPU probe loop (lookup_join.cpp:924-1035) looks rows up one at a time, and
 * its partitions (hash_builder.cpp:155-160) split the build side across operator instances instead. */
int omni_gpu_join_tune(void *handle, int32_t probe_group, int32_t partitions);
/* Partition count the last probe on this handle used (1 = direct probe, 0 = no probe yet). */
int omni_gpu_join_probe_info(void *handle, int32_t *partitions);

/* Probes one batch.  For each match writes out_probe_rows[k] = probe_row_base + i and, for each
 * payload column p, the build row's payload into out_payloads[p].values[k] (and its NULL bit).
 * Output order is unspecified (compare as a row set).  *num_matches receives the match count;
 * if it exceeds out_capacity, nothing beyond out_capacity is written and OMNI_GPU_ERR_OUTPUT_FULL
 * is returned.  A payload output whose `values` is NULL is skipped.  A payload output's `nulls` may
 * be NULL when that build column has no NULLs; otherwise it must hold 4 * ceil(out_capacity / 32)
 * bytes (it is cleared and written as 32-bit words).
 * Synchronizes. */
int omni_gpu_join_probe(void *handle, const omni_gpu_column *probe_key, int64_t rows,
                        int64_t probe_row_base, int64_t *out_probe_rows, omni_gpu_out_column *out_payloads,
                        int64_t out_capacity, int64_t *num_matches);

/* Probes many batches in ONE launch.  keys[b] holds batch b's key column with rows[b] rows (with several key
 * columns or a DECIMAL128 key: keys[b * num_keys ..]); probe row ids are numbered consecutively across batches
 * starting at 0 in batch order. */
int omni_gpu_join_probe_batches(void *handle, const omni_gpu_column *keys, const int64_t *rows,
                                int32_t num_batches, int64_t *out_probe_rows,
                                omni_gpu_out_column *out_payloads, int64_t out_capacity,
                                int64_t *num_matches);
/* ---- join types other than inner ----------------------------------------------------------------
 * Sets the join type before omni_gpu_join_build_finish (HashBuilderOperatorFactory(joinType, buildSide, ...),
 * core/src/operator/join/hash_builder.cpp:11-43; the default is OMNI_GPU_JOIN_INNER).  The probe entry points
 * then follow the reference's per-type probe loops (core/src/operator/join/lookup_join.cpp:1037-1438; a NULL
 * probe key matches nothing):
 *   LEFT / RIGHT with the build on the other side (or build_side UNKNOWN), FULL
 *        (ProbeBatchForOppositeSideOuterJoin, ProbeBatchForFullJoin): every match, and one output row with every
 *        payload NULL for each probe row without a match (payload outputs must carry null bitmaps);
 *   LEFT_SEMI (ProbeBatchForLeftSemiJoin): one output row per probe row with a match; payloads not written;
 *   LEFT_ANTI (ProbeBatchForLeftAntiJoin): one output row per probe row without one; payloads not written;
 *   LEFT / RIGHT with the build on the same side, FULL (ProbeBatchForSameSideOuterJoin, ...ForFullJoin):
 *        the matched build rows are also marked for omni_gpu_join_outer_output;
 *   EXISTENCE: probed with omni_gpu_join_probe_exists only.
 * They are probed like the inner join: the radix-partitioned window probe for large hash tables (omni_gpu_join_tune),
 * the direct probe for array tables and small ones. */
int omni_gpu_join_set_type(void *handle, int32_t join_type, int32_t build_side);
/* LookupJoinOutputBuilder::ConstructProbeColumns (core/src/operator/join/lookup_join.cpp:1853-1877) on the
 * device: out[c] row k = cols[c] row rows[k] (value and NULL bit) for k < n, rows being device int64 row ids
 * into the columns (the out_probe_rows of a probe over them).  Fixed-width columns; out[c].nulls, when not NULL,
 * must hold 4 * ceil(n / 32) bytes and is written in full as 32-bit words.  Synchronizes. */
int omni_gpu_join_gather_rows(const omni_gpu_column *cols, int32_t num_cols, const int64_t *rows, int64_t n,
                              omni_gpu_out_column *out);
/* Hash-table partition of each probe row of a join over several builder partitions (operatorCount > 1 in
 * HashBuilderOperatorFactory, hash_builder.cpp:155-160): PopulateProbeHashes (lookup_join.cpp:1572-1612; per row
 * h = 0, then h = 31 * h + HashUtil::HashValue(value) for each non-NULL key column) and
 * HashUtil::GetRawHashPartition(h, mask) (hash_util.h:146-154), into out_partitions[rows] (device).  Key columns
 * of SHORT, INT, DATE32, LONG, TIMESTAMP, DECIMAL64, DOUBLE, BOOLEAN or DECIMAL128.  Synchronizes. */
int omni_gpu_join_partition_ids(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, uint32_t mask,
                                uint32_t *out_partitions);
/* ProbeBatchForExistenceJoin (lookup_join.cpp:1393-1438): out_exists[i] = 1 when probe row i has a matching
 * build row, else 0 (also for a NULL key); rows bytes of device memory.  Works with any join type.
 * Synchronizes. */
int omni_gpu_join_probe_exists(void *handle, const omni_gpu_column *probe_key, int64_t rows, uint8_t *out_exists);
/* LookupOuterJoinOperator::GetOutput (core/src/operator/join/lookup_outer_join.cpp:155-320): after every probe,
 * writes each build row that no probe matched (NULL-key rows included) -- its build row id into
 * out_build_rows[k] (build rows numbered from 0 in build order; may be NULL) and its payloads into
 * out_payloads, which follow the rules of omni_gpu_join_probe.  The probe-side columns of these rows are NULL.
 * Only for build-preserving types.  Output order is unspecified; *num_rows receives the row count
 * (OMNI_GPU_ERR_OUTPUT_FULL beyond out_capacity).  Synchronizes. */
int omni_gpu_join_outer_output(void *handle, int64_t *out_build_rows, omni_gpu_out_column *out_payloads,
                               int64_t out_capacity, int64_t *num_rows);
int omni_gpu_join_destroy(void *handle);

/* ---- shuffle hash partitioning ----------------------------------------------------------------------------
 * Replaces HashUtil::ComputePartitionIds (core/src/operator/hash_util.cpp:40-82), the partition id of each row
 * for a hash exchange: a Murmur3 hash with seed 42 folded over the row's columns (HashInt of INT / DATE32 values,
 * HashLong of LONG / TIMESTAMP / DECIMAL64 / DOUBLE bits, HashInt of 0 or 1 for BOOLEAN, a NULL leaving the hash
 * unchanged; core/src/operator/util/mm3_util.h), then out_ids[i] = Pmod(hash, num_partitions).  Other column
 * types are rejected.  out_ids is a device array of rows int32.  Synchronizes. */
int omni_gpu_partition_ids(const omni_gpu_column *cols, int32_t num_cols, int64_t rows, int32_t num_partitions,
                           int32_t *out_ids);

/* ---- learned strategy / capacity advisor ----------------------------------------------------------------
 * No reference counterpart: the reference sizes its tables by growth only (BaseHashMap's Grower,
 * core/src/operator/hashmap/base_hash_map.h:255-292) and picks the table kind by a fixed rule.  These entry
 * points compute a sample sketch of key columns, and run / train a 16-64-64-4 fp32 MLP that predicts log2 of
 * the group count (output 0) and the fastest strategy (outputs 1..3: LDS-local, global-atomic,
 * radix-partitioned).  The prediction only sets a capacity hint and a strategy: results never depend on it. */
typedef struct omni_gpu_sketch {
    float features[16];     /* MLP input, each roughly in [0, 1]: log rows, log sample, log distinct, f1 / m,
                               f2 / m, max / m, heavy / m, distinct / m, log GEE and Chao1 estimates, log key
                               range, NULL share, key bytes, key columns, join flag, bias (m = sampled rows) */
    int64_t rows, sampled;  /* rows of the column, rows in the 1-in-k sample */
    int64_t distinct, f1, f2, max_count, heavy, nulls;  /* over the sample: distinct keys, keys seen once / twice,
                                                          the top key's count, rows of keys seen >= 4 times */
    int64_t key_min, key_max;  /* range of the first key column over the sample */
} omni_gpu_sketch;

/* Sketch of one or two integer-like key columns of `rows` rows (device columns): evenly strided sample of 1% of
 * the rows, at least 4,096 (all rows below that) and at most 65,536. */
int omni_gpu_sketch_keys(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                         omni_gpu_sketch *out);
/* omni_gpu_sketch_keys without any device allocation, for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops.  No counterpart in the reference.  The same sketch is computed on `stream` (a
 * hipStream_t; NULL = the null stream) into *out, a DEVICE pointer to an omni_gpu_sketch; nothing is copied to the
 * host and nothing synchronizes.  The sample's table, counts and statistics live in a fixed device array of the
 * library (1.5 MiB per GPU, held until the process ends; a sample has at most 65,536 rows, so every row count
 * fits it), which every call on a GPU reuses and clears with kernels: calls that overlap on one GPU race. */
int omni_gpu_sketch_keys_noalloc(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                                 void *stream, omni_gpu_sketch *out);
/* Sketch + MLP inference in one call on the MLP handle's stream (one synchronisation): y[4] receives the network's
 * outputs for the key columns (y[0] * 32 = predicted log2(groups + 1); y[1..3] strategy logits). */
int omni_gpu_advise(void *mlp, const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                    float *y);
/* MLP handle with He-uniform weights drawn from `seed`.  Parameters are the flattened W1[16x64] b1[64]
 * W2[64x64] b2[64] W3[64x4] b3[4] (row-major, y = x @ W + b); host arrays below. */
int omni_gpu_mlp_create(uint64_t seed, void **handle);
int omni_gpu_mlp_num_params(void);
int omni_gpu_mlp_set_params(void *handle, const float *params);
int omni_gpu_mlp_get_params(void *handle, float *params);
/* Y[n][4] = MLP(X[n][16]). */
int omni_gpu_mlp_forward(void *handle, const float *X, int32_t n, float *Y);
/* Gradient of loss = mean (y0 - t)^2 + mean cross-entropy(softmax(y1..3), c) over the n samples (c in 0..2). */
int omni_gpu_mlp_gradient(void *handle, const float *X, const float *t, const int32_t *c, int32_t n, float *grad,
                          float *loss);
/* Adam training (beta 0.9 / 0.999) on the GPU: `epochs` passes over the n samples in shuffled mini-batches;
 * epoch_loss (optional, `epochs` floats) receives each epoch's mean loss. */
int omni_gpu_mlp_train(void *handle, const float *X, const float *t, const int32_t *c, int32_t n, int32_t batch,
                       int32_t epochs, float lr, uint64_t seed, float *epoch_loss);
int omni_gpu_mlp_destroy(void *handle);
/* Seeded synthetic keys on the device (the advisor's training and evaluation runs): `rows` keys whose rank
 * follows a Zipf(s) law over `distinct` ranks (s = 0: uniform), mapped through a bijection.  kind & 7: 0 int64
 * into k64; 1 int32 into k32; 2 (int32, int64) pairs into (k32, k64).  kind & 8: row i gets rank i + 1 (a join
 * build side of unique keys); kind & 16: about 20% of rows get keys outside the rank range (probe misses).
 * v64 (optional): int64 values in [-1e6, 1e6]. */
int omni_gpu_generate_keys(int32_t kind, int64_t rows, double distinct, double zipf_s, uint64_t seed, void *k64,
                           void *k32, void *v64);

#ifdef __cplusplus
}
#endif
#endif /* OMNI_GPU_H */
