/* sqe.h -- C ABI of the MI355X sparse statevector engine (libsqe.so).
 *
 * Drop-in replacement for the GPU path behind the reference's circuit API
 * (ivanNovikau/SparQC, emulator/).  Each entry point names the reference
 * interface it replaces; a C++ class with the reference's method names sits on
 * top of this ABI in include/sparqc_circuit.hpp.
 *
 * Conventions
 *   - Qubit t is bit t of a uint64 basis key, qubit 0 the least significant bit;
 *     |0...0> is key 0 (emulator/device/device_table_FC.cuh:31-36,42-43).
 *   - Amplitudes are exchanged as separate FP64 real / imaginary arrays.
 *   - The engine owns all device memory; the caller owns host arrays, which
 *     sqe_load / sqe_export copy.
 *   - Calls enqueue work on the engine's own HIP stream and return; sqe_nnz,
 *     sqe_norm2, sqe_export and sqe_sync wait for it.  One engine is used by
 *     one host thread at a time.
 *   - Errors are status codes; no C++ exception crosses the ABI.
 *     sqe_last_error() returns the calling thread's last message.
 */
#ifndef SQE_H
#define SQE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sqe_engine sqe_engine;

typedef enum {
    SQE_OK = 0,
    SQE_EINVAL = -1,       /* bad argument (qubit index, duplicate key, ...)      */
    SQE_ENOMEM = -2,       /* device or host allocation failed                    */
    SQE_ETABLE_FULL = -3,  /* the state would exceed the configured HBM budget    */
    SQE_EHIP = -4,         /* HIP runtime error or kernel-reported fault          */
    SQE_EDEVICE = -5,      /* no usable GPU                                       */
    SQE_ENOTLOCAL = -6     /* batch: the circuit has cross-bucket gates (it
                              belongs on the single-circuit engine)              */
} sqe_status;

typedef enum { SQE_F64 = 0, SQE_F32 = 1 } sqe_prec;

/* gate classes: diagonal (keys unchanged, never pruned), permutation (keys
 * flipped, never pruned), branching (pairs mixed, outputs pruned) */
typedef enum { SQE_DIAG = 0, SQE_PERM = 1, SQE_BRANCH = 2 } sqe_class;

/* gate kinds and their coefficient conventions (p[]):
 *   X, Y, Z       none
 *   H             p[0] = 1/sqrt(2)
 *   PHASE         p[0..1] = cos(theta), sin(theta)           diag(1, e^{i theta})
 *   RZ, RX, RY    p[0..1] = cos(theta/2), sin(theta/2)
 *   U             p[0..7] = u00, u01, u10, u11 (re, im)       any 2x2, branching
 *   DIAG          p[0..3] = d0, d1 (re, im)                   diag(d0, d1)
 *   PERM          p[0..3] = u01, u10 (re, im)                 [[0,u01],[u10,0]]
 * (gate matrices: emulator/include/QLib.h:57-157; angle halving:
 *  emulator/device/table_manager.cuh:333-336) */
typedef enum {
    SQE_GATE_X = 0, SQE_GATE_Y = 1, SQE_GATE_Z = 2, SQE_GATE_H = 3, SQE_GATE_PHASE = 4,
    SQE_GATE_RZ = 5, SQE_GATE_RX = 6, SQE_GATE_RY = 7, SQE_GATE_U = 8, SQE_GATE_DIAG = 9,
    SQE_GATE_PERM = 10
} sqe_gate_kind;

/* One gate (88 bytes).  It acts on keys k with (k & ctrl_mask) == ctrl_val;
 * ctrl_mask must not contain the target and ctrl_val must be a subset of it. */
typedef struct sqe_gate {
    uint32_t kind;
    uint32_t target;
    uint64_t ctrl_mask;
    uint64_t ctrl_val;
    double p[8];
} sqe_gate;

typedef struct sqe_config {
    uint32_t nq;                /* qubits, 1..63                                       */
    uint32_t prec;              /* sqe_prec                                            */
    double prune_thr;           /* 0: 1e-14 (F64), 1e-7 (F32)                          */
    double max_load;            /* hash index load factor, 0: 0.5                      */
    uint64_t hbm_budget_bytes;  /* 0: free device memory less the reserve below        */
    int32_t device;             /* HIP device ordinal                                  */
    int32_t bucket_bits;        /* -1: planner chooses; else 0..4 (F64) / 0..5 (F32)   */
    uint32_t renorm_every;      /* renormalise in FP64 every N gates, 0: never         */
    uint32_t fuse_window;       /* max gates per fused sweep, 0: unlimited             */
    uint32_t timing;            /* 1: time every kernel launch with HIP events         */
    uint32_t mem_reserve_ppm;   /* with hbm_budget_bytes = 0: fraction of free device
                                   memory kept unused, parts per million (0: 10%)     */
    uint32_t fuse_blocks;       /* FP32 with 16- or 32-slot buckets (4 or 5 bucket
                                   qubits): fuse windows of at least 6 (1) or N
                                   (N >= 2) gates on bucket qubits into dense 16x16
                                   or 32x32 complex blocks applied with MFMA (prunes
                                   once per block); 0: off                            */
    uint32_t batch_eager;       /* batch: 1 launches each step of sqe_batch_run
                                   directly; 0 replays the run as one captured
                                   hipGraph (planned once per set of circuits)    */
    uint32_t renorm_mode;       /* FP32 with renorm_every: 0 stops at every
                                   renormalisation point (one launch per period);
                                   1 (default) applies a queued run of bucket-local
                                   gates in one pass that takes the FP64 norm at each
                                   renormalisation point and applies the scale
                                   afterwards (prune decisions then lack the pending
                                   scale, within 2^-11 of 1), checked: the run is
                                   redone step by step when a norm moved by more than
                                   2^-11; 2 as 1 with the check failing (tests the
                                   rerun)                                              */
    uint32_t frame_off;         /* 1: Pauli gates on bucket qubits update the amplitude
                                   registers instead of the per-bucket frame           */
    uint32_t reserved[2];
} sqe_config;

typedef struct sqe_stats {
    uint64_t gates;             /* gates applied                                       */
    uint64_t updates;           /* sum over gates of nonzeros entering the gate        */
    uint64_t nnz;               /* nonzero amplitudes (after the last sync)            */
    uint64_t buckets;           /* live buckets                                        */
    uint64_t capacity;          /* bucket capacity of one store                        */
    uint64_t bytes_device;      /* device bytes held by the engine                     */
    int32_t bucket_bits;        /* slots per bucket = 2^bucket_bits                    */
    uint32_t reruns;            /* one-pass renormalised runs rerun step by step       */
    uint64_t bucket_mask;       /* bucket qubits                                       */
    uint64_t sweeps;            /* fused local sweep launches                          */
    uint64_t sweep_ops;         /* gates executed inside sweeps                        */
    uint64_t pairs;             /* cross-bucket gate launches                          */
    uint64_t remaps;            /* re-bucketing passes                                 */
    uint64_t moved;             /* sum over sweep / cross-bucket launches of nonzeros
                                   read plus nonzeros written (algorithmic traffic)    */
    double sweep_ms;            /* summed sweep kernel time (timing = 1)               */
    double pair_ms;             /* summed cross-bucket kernel time (timing = 1)        */
    double other_ms;            /* summed time of all other kernels (timing = 1)       */
    uint64_t moved_pair;        /* the part of `moved` from cross-bucket launches      */
    uint64_t blocks;            /* fused dense blocks in launched sweep programs       */
    uint64_t block_gates;       /* gates those blocks hold                             */
} sqe_stats;

/* Fill *cfg with defaults for nq qubits in FP64. */
void sqe_config_default(sqe_config* cfg, uint32_t nq);

/* Replaces TableManager__::TableManager__ / device_analysis / reserve_on_device
 * (emulator/device/table_manager.cuh:84-209,227-288; emulator/device/table.cuh:30-45)
 * and the Mapping__ selection (emulator/device/table_manager.cuh:150-176). */
sqe_status sqe_create(const sqe_config* cfg, sqe_engine** out);

/* Replaces ~TableManager__ (emulator/device/table_manager.cuh:212-224). */
sqe_status sqe_destroy(sqe_engine* e);

/* Basis state |key> with amplitude 1.  Replaces Mapping__::init +
 * FC__::set_to_zero_quantum_state (emulator/device/mapping_GHT.cuh:28-32,
 * emulator/device/device_table_FC.cuh:39-49); key 0 gives |0...0>. */
sqe_status sqe_set_basis(sqe_engine* e, uint64_t key);

/* Bulk insert of n distinct keys with amplitudes (zero amplitudes are
 * skipped).  No reference counterpart (the reference can only start at
 * |0...0>); it is the insert/rebuild entry of the hash table. */
sqe_status sqe_load(sqe_engine* e, const uint64_t* keys, const double* re, const double* im, uint64_t n);

/* Apply a 2x2 gate u (row-major, re/im interleaved: u00, u01, u10, u11) of the
 * given class on `target` under controls.  Replaces TableManager__::gate_sq /
 * gate_sq_par and the Mapping__ virtual gate methods
 * (emulator/device/table_manager.cuh:311-340; emulator/device/mapping_GHT.cuh:38-48). */
sqe_status sqe_apply(sqe_engine* e, uint32_t target, uint64_t ctrl_mask, uint64_t ctrl_val, const double u[8],
                     sqe_class cls);

/* Apply n gate records in order (same replacement as sqe_apply). */
sqe_status sqe_apply_gates(sqe_engine* e, const sqe_gate* gates, uint64_t n);

/* Number of nonzero amplitudes (the reference's N_, emulator/device/table.cuh:19). */
sqe_status sqe_nnz(sqe_engine* e, uint64_t* n);

/* Sum of |a|^2 in FP64 (no reference counterpart). */
sqe_status sqe_norm2(sqe_engine* e, double* out);

/* sqe_norm2 without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (tight serving loops, code that
 * must not call hipMalloc).  Its per-workgroup partial sums go to a fixed
 * device array of 16384 doubles (128 KB per GPU, held until the process
 * ends) instead of engine-allocated memory, so states of at most
 * 16384 * 256 = 4194304 buckets are accepted (SQE_EINVAL beyond).  Gates must
 * not be queued (SQE_EINVAL: call sqe_sync first, since planning them may
 * allocate).  The array is shared by every engine on the GPU: calls that
 * overlap on one GPU (engines driven from several host threads) race. */
sqe_status sqe_norm2_noalloc(sqe_engine* e, double* out);

/* Copy the non-negligible entries (|re| >= thr or |im| >= thr) to host arrays
 * of capacity cap, in the engine's bucket order.  *n receives the entry count;
 * when it exceeds cap nothing is copied and SQE_EINVAL is returned.  Replaces
 * Table__::get_statevector (emulator/device/table.cuh:47-75). */
sqe_status sqe_export(sqe_engine* e, uint64_t* keys, double* re, double* im, uint64_t cap, uint64_t* n);

/* sqe_export restricted to keys with (key & mask) == value, filtered on the
 * device: the pattern view of the reference's Circuit__::get_state_full
 * (emulator/include/circuit.h:205-257), whose state_to_choose fixes some
 * qubits to 0 or 1 (-1: any). */
sqe_status sqe_export_where(sqe_engine* e, uint64_t mask, uint64_t value, uint64_t* keys, double* re, double* im,
                            uint64_t cap, uint64_t* n);

/* Measurement in the computational basis: `shots` keys drawn with
 * probability |a|^2 / norm2 over the non-negligible entries, into keys[shots].
 * Deterministic for a given state and seed (draw k uses the splitmix64 output
 * of seed + (k+1) * 0x9E3779B97F4A7C15, top 53 bits as a uniform in [0, 1)).
 * No counterpart in the reference (its readout is the whole state,
 * emulator/include/circuit.h:111-126); at most 2^30 shots per call. */
sqe_status sqe_sample(sqe_engine* e, uint64_t shots, uint64_t seed, uint64_t* keys);

/* Restart file: write the current state (keys and FP64 amplitudes of every
 * non-negligible entry, plus the gate count) to a flat binary file / replace
 * the engine's state with one.  Stands in for the reference's planned
 * _RESTART.hdf5 output (emulator/include/QLib.h:39-40, 612-823); layout: a
 * 64-byte header (magic "SQESTATE", version 1, nq, precision, count, gates),
 * then keys[count] (u64), re[count], im[count] (f64), little-endian. */
sqe_status sqe_checkpoint_save(sqe_engine* e, const char* path);
sqe_status sqe_checkpoint_load(sqe_engine* e, const char* path);

/* Wait for all enqueued work and surface device-side errors. */
sqe_status sqe_sync(sqe_engine* e);

/* Counters and timings (waits for enqueued work). */
sqe_status sqe_get_stats(sqe_engine* e, sqe_stats* out);

/* Move the state into the bucket layout the planner picks for this gate list
 * (without applying it): a staged state is inserted into the table, a resident
 * one re-bucketed.  Optional; the first flush of queued gates does the same. */
sqe_status sqe_prepare(sqe_engine* e, const sqe_gate* gates, uint64_t n);

/* Keep a device-resident copy of the current state / return to it.  No
 * reference counterpart (its _RESTART.hdf5 naming, emulator/include/QLib.h:39-40,
 * was never wired up). */
sqe_status sqe_snapshot_save(sqe_engine* e);
sqe_status sqe_snapshot_restore(sqe_engine* e);

/* Last error message of the calling thread ("" when none). */
const char* sqe_last_error(void);

/* ------------------------------------------------------------------------
 * Batch engine: many independent circuits resident on one GPU, advanced
 * together by persistent launches.  No reference counterpart -- the
 * reference runs one circuit per process (emulator/src/main.cu:87-148) and
 * has no request queue; this is the batched / serving form of the same
 * gate kernels.  Every gate of an admitted circuit must be bucket-local for
 * the bucket qubits the planner picks (its branching targets, at most 4):
 * sqe_batch_add rejects a circuit with any other gate (SQE_ENOTLOCAL), which
 * then runs on the single-circuit engine.
 * ------------------------------------------------------------------------ */
typedef struct sqe_batch sqe_batch;

typedef struct sqe_batch_circuit {
    uint32_t done;              /* 1 when every gate has been applied                  */
    int32_t bucket_bits;
    uint64_t gates_total;
    uint64_t gates_done;
    uint64_t updates;           /* sum over applied gates of nonzeros entering them    */
    uint64_t nnz;               /* nonzeros after the last step that touched it        */
    uint64_t buckets;
    uint64_t bytes;             /* pool bytes charged to this circuit                  */
} sqe_batch_circuit;

typedef struct sqe_batch_tuning {
    uint32_t block;             /* workgroup size of the batched sweep: 64/128/256     */
    uint32_t tiles_per_item;    /* 64-bucket tiles per wavefront work item, 1..64      */
    double load_factor;         /* hash index load factor used by sqe_batch_add        */
    double sweep_ms;            /* summed batched sweep time (timing = 1), or the
                                   best trial time from sqe_batch_autotune             */
    uint64_t launches;
    double load_ms;             /* sqe_batch_autotune: index part of the load path
                                   (clear, insert, scan numbering of every admitted
                                   circuit's buckets) at the chosen load factor        */
    uint32_t one_launch;        /* 1 (default): the circuits of 0-3 bucket qubits
                                   share one sweep launch per step; 0: one launch
                                   per bucket size, on concurrent graph branches       */
    uint32_t reserved;          /* 0 */
} sqe_batch_tuning;

/* cfg->nq is ignored (each circuit has its own); prec, prune_thr, max_load,
 * hbm_budget_bytes (admission budget), device, bucket_bits, renorm_every,
 * renorm_mode and timing apply to every circuit.  FP32 circuits with
 * renorm_every and renorm_mode 1 renormalise in one pass: an FP64 norm
 * checkpoint every renorm_every gates and one scale at the end, so a run is
 * one sweep launch per bucket size, a checkpoint fold and a scale launch; a
 * checkpoint norm more than 2^-11 from 1 fails the circuit (SQE_EDEVICE from
 * sqe_batch_info / export: admit it with renorm_mode 0, which stops at every
 * renormalisation point).  FP64 circuits run their whole program in one
 * sweep launch per bucket size. */
sqe_status sqe_batch_create(const sqe_config* cfg, sqe_batch** out);
sqe_status sqe_batch_destroy(sqe_batch* b);

/* Admit a circuit: nq qubits, n distinct keys with amplitudes, ng gates.
 * *id receives its handle.  SQE_ETABLE_FULL when its pool bytes would exceed
 * the budget (nothing is admitted; retry after sqe_batch_remove).  The state
 * is copied to the device asynchronously (the call does not wait for the
 * device); a duplicate key in it is reported for this circuit by
 * sqe_batch_info, sqe_batch_export and sqe_batch_export_poll (SQE_EINVAL). */
sqe_status sqe_batch_add(sqe_batch* b, uint32_t nq, const uint64_t* keys, const double* re, const double* im,
                         uint64_t n, const sqe_gate* gates, uint64_t ng, int64_t* id);

/* sqe_batch_add for a state already in page-locked host memory
 * (sqe_host_alloc; a serving front end that receives requests into such
 * buffers): the arrays are uploaded as they are, with no host-side pass over
 * them, so they must stay unchanged until the upload has completed (after
 * the next sqe_batch_sync, or once the circuit's export has completed).
 * Entries with a zero amplitude load as absent; a key with bits beyond nq is
 * reported for this circuit by sqe_batch_info / export (SQE_EINVAL), like a
 * duplicate key.  One upload when re and im follow keys at one pitch (the
 * three arrays of one sqe_host_alloc buffer).  No reference counterpart. */
sqe_status sqe_batch_add_pinned(sqe_batch* b, uint32_t nq, const uint64_t* keys, const double* re, const double* im,
                                uint64_t n, const sqe_gate* gates, uint64_t ng, int64_t* id);

/* Advance every unfinished circuit by up to max_gates program records
 * (0: to its next FP64 renormalisation point or its end). */
sqe_status sqe_batch_step(sqe_batch* b, uint32_t max_gates);

/* Step until every admitted circuit is done. */
sqe_status sqe_batch_run(sqe_batch* b);
sqe_status sqe_batch_sync(sqe_batch* b);
sqe_status sqe_batch_info(sqe_batch* b, int64_t id, sqe_batch_circuit* out);

/* Same contract as sqe_export, for one circuit of the batch. */
sqe_status sqe_batch_export(sqe_batch* b, int64_t id, uint64_t* keys, double* re, double* im, uint64_t cap,
                            uint64_t* n);

/* Release a circuit's pool space (finished or not). */
sqe_status sqe_batch_remove(sqe_batch* b, int64_t id);

/* Asynchronous export, for serving loops: enqueue the export of circuit id
 * (the entries the readout filter keeps, in bucket order) and its copy into
 * keys / re / im, then return without waiting.  The arrays must be
 * page-locked host memory from sqe_host_alloc (or device memory) of at least
 * the circuit's slot count entries: *bound (optional) receives that count,
 * and a smaller cap is refused with SQE_EINVAL before anything is enqueued.
 * The circuit may be removed right after this call; the arrays belong to the
 * export until sqe_batch_export_poll reports it done.  *ticket identifies
 * the export.  No reference counterpart (the reference's readout,
 * emulator/device/table.cuh:47-75, is synchronous). */
sqe_status sqe_batch_export_start(sqe_batch* b, int64_t id, uint64_t* keys, double* re, double* im, uint64_t cap,
                                  int64_t* ticket, uint64_t* bound);

/* *done = 1 once the export's entries are in the caller's arrays (wait = 1
 * waits for it; wait = 0 returns at once with *done = 0 while it runs).  When
 * done, *n receives the entry count and *info (optional) the circuit's
 * counters at export, the ticket is retired, and a duplicate key in the
 * circuit's load is reported as SQE_EINVAL. */
sqe_status sqe_batch_export_poll(sqe_batch* b, int64_t ticket, int32_t wait, int32_t* done, uint64_t* n,
                                 sqe_batch_circuit* info);

/* Ids of admitted circuits whose every gate has been enqueued (host state, no
 * wait); *n receives the count, SQE_EINVAL when it exceeds cap. */
sqe_status sqe_batch_finished(sqe_batch* b, int64_t* ids, uint64_t cap, uint64_t* n);

/* Admitted bytes, budget, and circuits not yet done. */
sqe_status sqe_batch_usage(sqe_batch* b, uint64_t* used, uint64_t* budget, uint64_t* active);

sqe_status sqe_batch_get_tuning(sqe_batch* b, sqe_batch_tuning* t);
sqe_status sqe_batch_set_tuning(sqe_batch* b, const sqe_batch_tuning* t);

/* Time each (workgroup size, tiles per item, one_launch) candidate on the
 * admitted circuits (fastest of three runs; their state is restored after
 * every run) and keep the
 * fastest; then time the index part of the load path (table clear, insert of
 * every admitted circuit's bucket bases, scan numbering into scratch) at each
 * candidate load factor (0.25 ... 0.75) and keep the fastest for later
 * admissions; *best (optional) receives the choice. */
sqe_status sqe_batch_autotune(sqe_batch* b, sqe_batch_tuning* best);

/* Keep a device-side copy of every admitted circuit's state, position and
 * counters / return to it (invalid after sqe_batch_add or sqe_batch_remove). */
sqe_status sqe_batch_snapshot_save(sqe_batch* b);
sqe_status sqe_batch_snapshot_restore(sqe_batch* b);

const char* sqe_batch_last_error(void);

/* Page-locked host memory for export (and load) arrays: device-to-host copies
 * into it run at the full host-link rate instead of through the runtime's
 * staging of pageable memory.  No counterpart in the reference (its readout
 * copies into a pageable vector, emulator/device/table.cuh:47-75).  Errors are
 * reported through sqe_batch_last_error. */
sqe_status sqe_host_alloc(uint64_t bytes, void** out);
sqe_status sqe_host_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* SQE_H */
