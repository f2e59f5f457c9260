/*
 * ccs.h — C ABI of the MI355X-native Cunningham-chain (first kind) candidate pipeline.
 *
 * The reference (nmicic/cunningham-chain-search) has no plugin or FFI layer: its
 * boundary is the in-file contract between the host driver, the filter kernel K1 and
 * the prover.  Each entry point below names the reference code it replaces.
 *   v13 = src/cuda/cc18_filter_cuda_CpC_v13.cu
 *   v16 = src/cuda/experimental/cc20_first_kind_immune_v16.cu
 *
 * Conventions: plain pointers and sizes only; device pointers are prefixed d_, host
 * pointers h_; streams are hipStream_t passed as void* (NULL = the context's stream).
 * Every call returns CCS_OK (0), a negative error, or CCS_OVERFLOW (> 0) when an output
 * buffer was too small — the true count is still reported.  Nothing calls exit().
 * One host thread per context.
 */
#ifndef CCS_H
#define CCS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCS_OK 0
#define CCS_OVERFLOW 1
#define CCS_ERR_INVALID (-1)
#define CCS_ERR_HIP (-2)
#define CCS_ERR_NOMEM (-3)
#define CCS_ERR_UNSUPPORTED (-4)
#define CCS_ERR_BUSY (-5)
#define CCS_DONE 2                  /* a campaign has no further batch (sweep complete, --tiles reached) */

typedef struct ccs_ctx ccs_ctx;       /* one GPU: stream, scratch buffers, timing events   */
typedef struct ccs_shape ccs_shape;   /* CRT shape + filter profile, device-resident tables */

/* 128-bit unsigned as two 64-bit halves */
typedef struct {
    uint64_t lo, hi;
} ccs_u128;

/* Shape parameters.  Candidates n = (tile*WP + k)*M + base with M = prod(crt_primes),
 * WP = prod(wheel_primes), base admissible mod M and k admissible mod the wheel primes
 * for chains of length `target`, and n = 5 (mod 6) (v13.cu:907-1059).  The filter
 * rejects n when a filter prime q divides p_i = 2^i (n+1) - 1 for i in [0, depth) if
 * q <= pos0_max_q, else for i in [1, depth) (v13.cu:793-847, L2/ext vs line sieve). */
typedef struct {
    uint32_t n_crt;
    const uint32_t* crt_primes;     /* v13 CRT_PRIMES {5,...,19}                       */
    uint32_t n_wheel;
    const uint32_t* wheel_primes;   /* v13 WHEEL_PRIMES {23,29,31}                     */
    int32_t target;                 /* v13 target_len                                   */
    int32_t depth;                  /* v13 sieve_len, 1..32                             */
    uint32_t n_filter;
    const uint32_t* filter_primes;  /* ascending, all above the largest wheel prime, < 2^18 */
    uint32_t n_stages;              /* reporting stages: stage s ends at filter index  */
    const uint32_t* stage_end;      /*   stage_end[s] (exclusive); may be 0/NULL        */
    uint32_t pos0_max_q;            /* reference: 97                                   */
    uint32_t flags;                 /* CCS_SHAPE_*                                      */
} ccs_shape_params;

#define CCS_SHAPE_NO_CLASSES 1u     /* build filter tables only (class-form jobs)      */

typedef struct {
    uint64_t M, WP, W;              /* W = 6*WP*M (family shapes: low 64 bits of D*2^E)  */
    uint32_t n_bases;
    uint32_t n_filter;
    uint64_t n_classes;             /* admissible (base, k) pairs = classes mod W       */
    uint32_t n_dense, n_tierb, n_deep;   /* how the filter primes are split on the GPU  */
    uint32_t rows_per_group, unit_j;
    uint64_t W_hi;                  /* high 64 bits of W (family shapes)                */
    int32_t family, exp_start;      /* family shape: n = Q*D*2^exp_start - 1            */
    ccs_u128 D;                     /* family shape: immune-factor product D            */
    uint32_t n_slots;               /* dense patterns (a pattern folds up to 3 primes)  */
    uint32_t n_tierb_lds;           /* stage-B primes whose bitsets sit in LDS          */
} ccs_shape_info;

/* ---- fingerprint families (the v16 Q-iteration front end, v16...cu:1-30, :556-924,
 *      :4405-4468, :5104-5227): p_0 = Q*D*2^E - 1, p_j = Q*D*2^(E+j) - 1.  Every prime
 *      dividing D divides p_j + 1 for all j (the family is immune there); the search
 *      variable is Q.  A family shape is the class W - 1 mod W = D*2^E (u128): its jobs
 *      are Q windows [u0, u0 + count) with a = 0, and survivors are (job, Q - u0). ---- */
typedef struct {
    ccs_u128 D;                     /* immune-factor product D_V (a pool entry)          */
    int32_t exp_start;              /* E (v16 --exp-start)                               */
    int32_t target;                 /* chain length of interest (top position = target-1) */
    int32_t depth;                  /* positions [0, depth) are sieved for every prime   */
    uint32_t n_filter;
    const uint32_t* filter_primes;  /* ascending odd primes < 2^18; those dividing D are  */
                                    /*   immune and dropped (v16 per-V active primes)    */
} ccs_family_params;

/* Survivor record — identical to v13 `Survivor` (v13.cu:207-212): tile_idx is relative
 * to the launch's tile_start. */
typedef struct {
    uint32_t base_idx, k_offset, tile_idx, _pad;
} ccs_survivor;

/* Class-form survivor (batched/served jobs): n = a + W*(u0 + u_off) of job job_idx.
 * The 8-byte analogue of v16 `QSurvivor` (v16...cu:103-108). */
typedef struct {
    uint32_t job_idx, u_off;
} ccs_job_survivor;

/* Class-form job ("fingerprint shape"): residue class a mod W and u in [u0, u0+count). */
typedef struct {
    uint64_t a;
    ccs_u128 u0;
    uint64_t count;
} ccs_job;

/* Prove result per survivor (same index as the survivor array): the outcome of the
 * GMP confirmation engine prover_worker (v13.cu:1794-1924); chain_len counts from the
 * true root reached by the backward walk. */
typedef struct {
    uint32_t survivor_idx;
    uint32_t chain_len;     /* 0 unless CCS_F_PRIME                              */
    uint32_t back_steps;    /* 0 = the survivor is the root                       */
    uint32_t flags;         /* CCS_F_*                                             */
    uint64_t root_lo, root_hi;
} ccs_chain_result;

/* chain statistics of prove results (v13 ProverStats, v13.cu:1740-1768) */
typedef struct {
    uint64_t primes;            /* prime survivors (ProverStats.primes)               */
    uint64_t chains;            /* chains followed from a true root                   */
    uint64_t non_roots;         /* primes whose (n-1)/2 is prime (walked back)        */
    uint32_t best_chain, _pad;
    uint64_t chain_dist[65];    /* chains by length (v13 g_chain_dist)                */
} ccs_confirm_stats;

#define CCS_F_BITS_OK 1u       /* passed the bit-size gate (v13.cu:1828)          */
#define CCS_F_TRIAL_OK 2u      /* no factor among 67..409 (v13.cu:1832-1838)      */
#define CCS_F_PRIME 4u         /* probable prime (Baillie-PSW or MR bases)        */
#define CCS_F_NONROOT 8u       /* (n-1)/2 prime: walked back (v13.cu:1849-1867)   */
#define CCS_F_OVERFLOW 16u     /* a chain value reached 2^127: not followed       */
#define CCS_F_UNSUPPORTED 32u  /* no Lucas parameter (perfect square): unreachable */

#define CCS_TEST_BPSW 0        /* mpz_probab_prime_p(n, 1) decision (v13.cu:1841) */
#define CCS_TEST_MR 1          /* strong Miller-Rabin over explicit bases          */

typedef struct {
    int32_t bits;              /* bit-size gate (--bits), 0 = off                  */
    int32_t test_kind;         /* CCS_TEST_*                                       */
    uint32_t n_bases;          /* for CCS_TEST_MR, <= 16                           */
    uint32_t bases[16];
} ccs_prove_params;

/* Filter statistics of the last launch on a context. */
typedef struct {
    uint64_t candidates;       /* lattice points in the range (v33.c:1900 local_wheel++) */
    uint64_t alive_dense;      /* after the dense stage                            */
    uint64_t alive_tierb;      /* after stage B                                    */
    uint64_t survivors;        /* after all filter primes                          */
    uint32_t dense_end, tierb_end;   /* filter-prime indices where the GPU stages end */
    float ms_sieve, ms_deep, ms_prove;   /* device time of each kernel (last call)   */
} ccs_filter_stats;

/* Kernel tiling knobs (chosen by the autotuner, or set explicitly). */
typedef struct {
    uint32_t n_dense;          /* filter primes handled by the dense pattern stage (<= 128, at most 64 patterns) */
    uint32_t n_tierb;          /* next primes handled by stage B (<= 1024, all < 16384) */
    uint32_t rows_per_group;   /* 1, 2 or 4 rows of 8192 candidates (64 lanes x 128) per step */
    uint32_t unit_j;           /* candidates per work unit, multiple of 8192, <= 2^20 */
    uint32_t blocks_per_cu;    /* persistent grid = CUs * blocks_per_cu             */
    uint32_t pair_max;         /* fold up to three adjacent dense primes into one  */
                               /*   pattern while their product <= pair_max (at    */
                               /*   most 2^18; 0 = one prime per pattern)           */
    uint32_t waves_per_block;  /* 1, 2, 4, 8 or 16 wave64s share one LDS table copy */
} ccs_tuning;

/* ---- context (replaces the module statics + streams of v13.cu:268-327) ---- */
int ccs_ctx_create(int device, ccs_ctx** out);
void ccs_ctx_destroy(ccs_ctx* ctx);
int ccs_device_count(int* out);
const char* ccs_last_error(void);

/* ---- device memory helpers (plain pointers for callers without a framework) ---- */
int ccs_dmalloc(ccs_ctx* ctx, void** d_ptr, uint64_t bytes);
int ccs_dfree(ccs_ctx* ctx, void* d_ptr);
int ccs_memcpy_h2d(ccs_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
int ccs_memcpy_d2h(ccs_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
int ccs_memset_d(ccs_ctx* ctx, void* d_dst, int value, uint64_t bytes);
int ccs_synchronize(ccs_ctx* ctx);   /* waits for every stream of the device */

/* ---- streams and events of the context's device (default priority), for callers with
 *      no HIP runtime of their own: the `stream` argument of the launch entry points takes
 *      these (NULL = the context's stream).  Lets the filter of one batch overlap the
 *      prove of the previous one, as the reference overlaps its GPU filter with the CPU
 *      prover threads (v13.cu:1967-2068, :3769-3960).  Launches on different streams of
 *      one context may run at the same time: each stream the context launches on gets its
 *      own launch scratch (stage-B buffer, counters, prove queue), held until
 *      ccs_stream_destroy or ccs_ctx_destroy; launches on one stream share it in order. ---- */
int ccs_stream_create(ccs_ctx* ctx, void** stream);
int ccs_stream_destroy(ccs_ctx* ctx, void* stream);
int ccs_event_create(ccs_ctx* ctx, void** event);
int ccs_event_destroy(ccs_ctx* ctx, void* event);
int ccs_event_record(ccs_ctx* ctx, void* event, void* stream);
int ccs_stream_wait_event(ccs_ctx* ctx, void* stream, void* event);
int ccs_event_synchronize(ccs_ctx* ctx, void* event);   /* host waits for the event */

/* ---- shape (replaces init_filter_tables v13.cu:793-901, generate_bases :907-1059,
 *      gpu_upload_tables :1065-1222) ---- */
int ccs_shape_build(ccs_ctx* ctx, const ccs_shape_params* params, const ccs_tuning* tuning,
                    ccs_shape** out);
/* Family shape (the v16 Q-iteration sieve: per-V forbid masks v16...cu:5104-5227 with the
 * active-prime lists :5229-5300): W = D*2^E, the filter tables of the primes that do not
 * divide D, positions [0, depth) for every prime (v16_forbidden_residues :1244-1262).
 * Job sets, batched launches, graphs and proves work on it as on any class-form shape;
 * plain tile launches, tile ranges and lattice decomposition do not. */
int ccs_family_shape_build(ccs_ctx* ctx, const ccs_family_params* params, const ccs_tuning* tuning,
                           ccs_shape** out);
/* Q window of a family for candidates of bits_lo..bits_hi bits (compute_q_windows_subband,
 * v16...cu:4405-4468): q_min = ceil(2^(bits_lo-1) / (D*2^E)),
 * q_max = floor((2^bits_hi - 1) / (D*2^E)); CCS_DONE when q_min > q_max. */
int ccs_family_window(ccs_u128 D, int32_t exp_start, int32_t bits_lo, int32_t bits_hi,
                      ccs_u128* q_min, ccs_u128* q_max);

/* A pool entry (v16 v16_pool_seed, v16...cu:137-144): a named D and its prime factors
 * (a prime power p^k appears k times). */
typedef struct {
    char name[64];
    ccs_u128 D;
    uint32_t primes[64];
    uint32_t n_primes, _pad;
} ccs_pool_entry;
/* Comma-separated pool list (v16 --seed-pool: parse_seed_pool :650-690, parse_pool_spec
 * :556-648, expand_auto_pool_one :828-880): "Z#" (primes <= Z), "Z#/X/Y" (with X, Y
 * omitted), "P*Q^k*R" (explicit, k <= 16), "auto:Z#:K" or "auto:Z#:K1-K2" (every way of
 * dropping K primes <= Z that are not in keep[]).  Appends to out[]; *n_out = entries
 * produced (CCS_OVERFLOW when more than cap). */
int ccs_pool_parse(const char* list, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                   uint32_t cap, uint32_t* n_out);
/* Pool file: one list per line; everything from the first '#' on a line is a comment
 * (parse_seed_pool_file, v16...cu:697-738). */
int ccs_pool_load(const char* path, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                  uint32_t cap, uint32_t* n_out);
/* first[i] = 1 iff values[i] does not occur at a lower index (the v16 dedup of p_0 over
 * (Q, V) framings of one chain, v16...cu:2478-2541). */
int ccs_dedup_u128(const ccs_u128* values, uint32_t count, uint8_t* first);
void ccs_shape_free(ccs_shape* shape);
int ccs_shape_query(const ccs_shape* shape, ccs_shape_info* out);
int ccs_shape_bases(const ccs_shape* shape, uint64_t* h_out, uint32_t cap);
/* residue class a mod W and mod-6 tile residue of every (base, k) class, base-major */
int ccs_shape_classes(const ccs_shape* shape, uint64_t* h_a, uint32_t* h_base_idx,
                      uint32_t* h_k, uint32_t* h_rho, uint64_t cap);
int ccs_shape_set_tuning(ccs_shape* shape, const ccs_tuning* tuning);
int ccs_shape_get_tuning(const ccs_shape* shape, ccs_tuning* out);
/* Autotuner: coordinate descent over the dense-prime cutoff, rows per group, work-unit
 * size and blocks per CU, each trial timed (HIP events, sieve + deep) on the sample tile
 * range; leaves the fastest tuning set on the shape.  At most max_trials timed trials. */
int ccs_autotune(ccs_ctx* ctx, ccs_shape* shape, ccs_u128 tile_start, uint32_t tile_count,
                 uint32_t max_trials, ccs_tuning* best, float* best_ms);
/* The same search timed on the batched launch over a sample of class-form jobs. */
int ccs_autotune_jobs(ccs_ctx* ctx, ccs_shape* shape, const ccs_job* h_jobs, uint32_t n_jobs,
                      uint32_t max_trials, ccs_tuning* best, float* best_ms);
uint64_t ccs_count_candidates(const ccs_shape* shape, ccs_u128 tile_start, uint32_t tile_count);

/* Campaign range: tiles [tile_min, tile_max) whose candidates have `bits` bits, narrowed
 * to the binary prefix (prefix_bits = 0: none) and split into `lanes` (0 = no split);
 * replaces compute_prefix_tile_range (v13.cu:450-518) and the --bits / --prefix-lanes
 * range setup (v13.cu:3474-3573). */
int ccs_tile_range(const ccs_shape* shape, int32_t bits, ccs_u128 prefix_value, int32_t prefix_bits,
                   uint32_t lanes, uint32_t lane_id, ccs_u128* tile_min, ccs_u128* tile_max);

/* "0b1..." -> value and bit count, with the reference's errors (parse_prefix, v13.cu:408-448) */
int ccs_parse_prefix(const char* str, ccs_u128* value, int32_t* bits);

/* Where n sits on the shape's lattice (the --verify-root diagnostic, v13.cu:2921-3066):
 * base = n mod M, k = (n - base)/M, tile = k / WP, k_offset = k mod WP; on_lattice when the
 * base is admissible, (base, k_offset) is an admissible class and tile = rho (mod 6).
 * Needs a shape built with its class table. */
typedef struct {
    uint32_t on_lattice;       /* all checks below hold                              */
    uint32_t base_found;       /* n mod M is one of the shape's bases                */
    uint32_t class_found;      /* (base, k_offset) is an admissible class            */
    uint32_t bucket_ok;        /* tile = rho(base, k_offset) (mod 6)                 */
    uint32_t n_mod6;           /* 5 for the first kind                               */
    uint32_t base_idx, k_offset, _pad;
    uint64_t base;
    ccs_u128 tile;
} ccs_decomposition;
int ccs_shape_decompose(const ccs_shape* shape, ccs_u128 n, ccs_decomposition* out);

/* ---- campaign driver state (replaces the v13 main loop's tile walk, stop rules and
 *      checkpoint file: v13.cu:3584-3667, :3719-3765, :3960-3970, :520-664).  Host-only;
 *      the caller runs the filter / prove launches for the batches it hands out. ---- */
typedef struct ccs_campaign ccs_campaign;
#define CCS_WALK_INHERIT 0          /* --prefix-mode not given: sequential with --start, else random */
#define CCS_WALK_SEQUENTIAL 1
#define CCS_WALK_RANDOM 2
typedef struct {
    int32_t bits, target, depth;   /* --bits, --target, --depth (checkpoint compatibility) */
    const char* prefix;            /* --prefix "0b..." or NULL                              */
    int32_t walk;                  /* --prefix-mode: CCS_WALK_*                            */
    uint32_t lanes, lane_id;       /* --prefix-lanes, --prefix-lane-id                      */
    int32_t has_start;             /* --start given                                        */
    int32_t has_seed;              /* --seed given (else /dev/urandom; a given seed makes   */
    uint64_t seed;                 /*   the random walk reproducible)                       */
    ccs_u128 start;
    uint32_t batch_tiles, _pad;    /* --batch                                               */
    uint64_t max_tiles;            /* --tiles (0 = unlimited)                               */
} ccs_campaign_params;
typedef struct {
    ccs_u128 tile_min, tile_max, current;
    uint64_t seed, batches, total_tiles, total_candidates, total_survivors;
    int32_t random_mode, done;
    ccs_confirm_stats chains;
} ccs_campaign_state;

int ccs_campaign_create(const ccs_shape* shape, const ccs_campaign_params* params, ccs_campaign** out);
/* the same over an explicit tile range (no shape needed) */
int ccs_campaign_create_range(ccs_u128 tile_min, ccs_u128 tile_max, const ccs_campaign_params* params,
                              ccs_campaign** out);
void ccs_campaign_free(ccs_campaign* c);
/* next batch [*tile_start, *tile_start + *tile_count): CCS_OK, or CCS_DONE */
int ccs_campaign_next(ccs_campaign* c, ccs_u128* tile_start, uint32_t* tile_count);
/* account a finished batch, in the order next() handed them out */
int ccs_campaign_record(ccs_campaign* c, ccs_u128 tile_start, uint32_t tile_count, uint64_t candidates,
                        uint64_t survivors, const ccs_chain_result* res, uint32_t n_res);
int ccs_campaign_get_state(const ccs_campaign* c, ccs_campaign_state* out);
/* CC18_CUDA_v10 checkpoint, written to path.tmp then renamed over path; the tile saved is
 * the first one whose batch is not recorded yet */
int ccs_campaign_save(const ccs_campaign* c, const char* path);
/* continue sequentially from a checkpoint (bits, target, depth, prefix and lanes must match) */
int ccs_campaign_resume(ccs_campaign* c, const char* path);

/* ---- plain filter: K1 over tiles [tile_start, tile_start + tile_count), all bases
 *      (replaces prepare_batch v13.cu:1232-1259 + cc18_filter_kernel :1551-1714 and its
 *      launch :3769-3796).  Survivors go to d_out in unspecified order; *d_count (device)
 *      receives the true count of the launch's survivors, unless its stage-B buffer
 *      overflowed (ccs_stream_status).  Asynchronous on `stream`. ---- */
int ccs_filter_launch(ccs_ctx* ctx, const ccs_shape* shape, ccs_u128 tile_start,
                      uint32_t tile_count, ccs_survivor* d_out, uint32_t cap, uint32_t* d_count,
                      void* stream);
/* The same launch with no device allocation (no counterpart in the reference; for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its stage-B buffer (2^20 entries), counters and statistics are one
 * fixed array of this library in each device's memory (8 MiB + 256 B, held until the process
 * ends); the caller passes d_out and d_count as for ccs_filter_launch.  Tile ranges whose
 * expected stage-B survivors (twice the expectation + 2^16) exceed the buffer are rejected
 * with CCS_ERR_INVALID; a launch whose stage-B survivors still overflow it sets *d_count to
 * 0xFFFFFFFF.  Every call on one device shares that device's array: calls that overlap on one
 * device (other streams, other contexts, other threads) race. */
int ccs_filter_launch_noalloc(ccs_ctx* ctx, const ccs_shape* shape, ccs_u128 tile_start,
                              uint32_t tile_count, ccs_survivor* d_out, uint32_t cap, uint32_t* d_count,
                              void* stream);
/* Synchronous host form with statistics. */
int ccs_filter(ccs_ctx* ctx, const ccs_shape* shape, ccs_u128 tile_start, uint32_t tile_count,
               ccs_survivor* h_out, uint32_t cap, uint32_t* h_count, ccs_filter_stats* stats);
int ccs_last_stats(ccs_ctx* ctx, ccs_filter_stats* out);

/* Stage-B overflow of asynchronous launches (the reference warns and clamps on survivor
 * overflow, v13.cu:3846-3852; v16 aborts, v16...cu:5729-5743).  The candidates that pass the
 * dense stage and stage B wait in a device buffer sized from the shape's expected survival
 * (twice the expectation + 2^20 entries) before the deep stage; a launch with more of them
 * drops the excess, and its *d_count then undercounts.  Every such launch is recorded in a
 * sticky status of the stream's (or job set's) launch scratch.  ccs_stream_status /
 * ccs_jobset_status wait for the stream, return the status of the launches since the last
 * call and clear it: CCS_OK when none overflowed, CCS_OVERFLOW when some did. */
typedef struct {
    uint32_t overflowed_launches;   /* launches whose stage-B survivors exceeded the buffer  */
    uint32_t worst_stage_b;         /* the largest true stage-B count among them             */
    uint64_t dropped;               /* stage-B survivors lost over those launches            */
    uint32_t stage_b_cap;           /* entries of the buffer                                 */
    uint32_t _pad;
} ccs_launch_status;
int ccs_stream_status(ccs_ctx* ctx, void* stream, ccs_launch_status* out);
/* Test knob (no reference counterpart): launches on the context use at most `entries` of
 * their stage-B buffer (0 = all of it), so tests can force the overflow path. */
int ccs_debug_limit_stage_b(ccs_ctx* ctx, uint32_t entries);
/* Test knob (no reference counterpart): the deep stage of launches on the context takes
 * `per_wave` stage-B records per wave batch (1: one at a time; 2..64: batches of that
 * many; 0 = chosen from the launch's count, the default), so tests can run either form
 * on any input.  Results do not depend on it. */
int ccs_debug_deep_batch(ccs_ctx* ctx, uint32_t per_wave);

/* ---- prove (replaces cc18_prove_kernel v13.cu:1462-1538 and the GMP prover pool
 *      prover_worker :1794-1924): one ccs_chain_result per survivor.  The launch reads
 *      min(*d_count, cap) survivors; cap <= 4096 runs one kernel with a wave per survivor
 *      (every chain position tested at once: short latency), a larger cap a gate kernel
 *      and a chain-walk kernel with a lane per survivor (little work per survivor for
 *      large batches).  Both give the same results. ---- */
int ccs_prove_launch(ccs_ctx* ctx, const ccs_shape* shape, ccs_u128 tile_start,
                     const ccs_survivor* d_surv, const uint32_t* d_count, uint32_t cap,
                     const ccs_prove_params* params, ccs_chain_result* d_out, void* stream);
int ccs_prove(ccs_ctx* ctx, const ccs_shape* shape, ccs_u128 tile_start,
              const ccs_survivor* h_surv, uint32_t n, const ccs_prove_params* params,
              ccs_chain_result* h_out);

/* The reference's chain result record (v13 ChainResult, v13.cu:218-223): a chain root among
 * a launch's survivors, named by its Survivor fields, and the chain length from it. */
typedef struct {
    uint32_t base_idx, k_offset, tile_idx, chain_len;
} ccs_chain_root;

/* The v13 K2 output (cc18_prove_kernel, v13.cu:1462-1538) from prove results: one
 * ccs_chain_root per prime survivor whose (n-1)/2 is not prime (back_steps 0), chain_len =
 * its length forward, in unspecified order.  Reads min(*d_count, cap) survivors and their
 * results (same index); *d_root_count = the true number of roots (at most out_cap written),
 * *d_prime_count (optional) = the prime survivors (v13 gpu_primes_count).  Both counts are
 * set by the launch.  A non-root survivor, which the GMP engine walks back to its true root,
 * is left out here, as K2 leaves it out. */
int ccs_chain_roots_launch(ccs_ctx* ctx, const ccs_survivor* d_surv, const ccs_chain_result* d_res,
                           const uint32_t* d_count, uint32_t cap, ccs_chain_root* d_out, uint32_t out_cap,
                           uint32_t* d_root_count, uint32_t* d_prime_count, void* stream);

/* Format accepted chains like the v13 results file (v13.cu:2030-2033):
 * "[TARGET ][NON-ROOT ]CC<len> 0x<HEX> <bits>\n" for len >= log_threshold or >= target.
 * Returns the number of bytes needed (excluding NUL). */
int ccs_format_results(const ccs_chain_result* res, uint32_t n, int32_t log_threshold,
                       int32_t target, char* buf, uint64_t buf_size, uint64_t* needed);

/* ---- host confirmation engine (replaces the CPU prover pool prover_worker v13.cu:1794-1924
 *      and its merge prove_survivors_mt :1967-2068): the same semantics on host threads with
 *      the library's own u128 Baillie-PSW instead of GMP.  Pure CPU work; usable without a
 *      GPU, e.g. to re-confirm the GPU's chains ---- */

/* one result per value, same fields as the GPU prove; nthreads 0 = hardware threads
 * (at most 16) */
int ccs_confirm_values(const ccs_u128* h_n, uint32_t count, const ccs_prove_params* params,
                       uint32_t nthreads, ccs_chain_result* h_out, ccs_confirm_stats* stats);
/* survivors of a plain launch, n = ((T0 + tile_idx)*WP + k)*M + base (v13.cu:1819-1826) */
int ccs_confirm_host(const ccs_shape* shape, ccs_u128 tile_start, const ccs_survivor* h_surv,
                     uint32_t count, const ccs_prove_params* params, uint32_t nthreads,
                     ccs_chain_result* h_out, ccs_confirm_stats* stats);
/* add prove results (GPU or host) to a stats record */
int ccs_confirm_stats_add(const ccs_chain_result* res, uint32_t count, ccs_confirm_stats* stats);

/* ---- kernel timing: while collecting, every sieve / deep / prove launch on the context
 *      records HIP events on the stream it runs on (the reference times K1 with CUDA events,
 *      v13.cu:3786-3796) ---- */
typedef struct {
    uint32_t n_sieve, n_deep, n_prove;
    double ms_sieve, ms_deep, ms_prove;   /* summed device time per kernel kind */
} ccs_kernel_times;

int ccs_profile_start(ccs_ctx* ctx);
int ccs_profile_stop(ccs_ctx* ctx, ccs_kernel_times* out);   /* synchronises the stream */

/* ---- batched form: many class-form jobs in one persistent, work-stealing launch
 *      (replaces the v16 per-(variant, chunk) launch loop v16...cu:5611-5965) ---- */
typedef struct ccs_jobset ccs_jobset;   /* jobs + their unit prefix, uploaded once */
typedef struct ccs_graph ccs_graph;     /* a captured batched launch               */

int ccs_jobset_create(ccs_ctx* ctx, const ccs_shape* shape, const ccs_job* h_jobs, uint32_t n_jobs,
                      ccs_jobset** out);
void ccs_jobset_free(ccs_jobset* js);
/* replace the jobs of a set (at most as many as it was created with, within its stage-B
 * buffer): the serving loop's per-slice job table */
int ccs_jobset_update(ccs_ctx* ctx, ccs_jobset* js, const ccs_job* h_jobs, uint32_t n_jobs, void* stream);
/* A job set over several shapes (no reference counterpart: it replaces v16's one launch per
 * variant, v16...cu:5611-5965, and the served form's one launch per depth profile): jobs[i]
 * belongs to shapes[job_shape[i]], jobs grouped by shape (job_shape non-decreasing); the
 * shapes share rows_per_group.  Launch, per-job loop, graph, prove and top gate work on it
 * like on a one-shape set; survivors name the job, whose shape gives W. */
int ccs_groupset_create(ccs_ctx* ctx, const ccs_shape* const* shapes, uint32_t n_shapes,
                        const ccs_job* h_jobs, const uint32_t* job_shape, uint32_t n_jobs,
                        ccs_jobset** out);
int ccs_groupset_update(ccs_ctx* ctx, ccs_jobset* js, const ccs_job* h_jobs, const uint32_t* job_shape,
                        uint32_t n_jobs, void* stream);
/* one persistent work-stealing launch over all jobs (kernels only: graph-capturable) */
int ccs_jobset_launch(ccs_ctx* ctx, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                      uint32_t* d_count, void* stream);
/* the plain per-shape launch loop over the same jobs (comparator; v16...cu:5611-5965) */
int ccs_jobset_launch_each(ccs_ctx* ctx, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                           uint32_t* d_count, void* stream);
/* counters of the last launch on the set (synchronises) */
int ccs_jobset_stats(ccs_ctx* ctx, ccs_jobset* js, uint32_t* d_count, ccs_filter_stats* out);
/* stage-B overflow status of the set's launches since the last call, once `stream` is done
 * (see ccs_stream_status) */
int ccs_jobset_status(ccs_ctx* ctx, ccs_jobset* js, void* stream, ccs_launch_status* out);
int ccs_graph_create(ccs_ctx* ctx, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                     uint32_t* d_count, ccs_graph** out);
int ccs_graph_launch(ccs_ctx* ctx, ccs_graph* graph, void* stream);
void ccs_graph_free(ccs_graph* graph);
/* K1.5 top-position gate of a family set (v16_k15_topmr_kernel, v16...cu:2185-2296): keeps
 * the survivors whose p_top = Q*D*2^(E+top_pos) - 1 passes strong Miller-Rabin to the first
 * `reps` (2 or 3) bases of {2, 3, 5}; compacted into d_out in unspecified order, *d_out_count
 * = their true count.  Every job's p_top must stay below 2^127. */
int ccs_family_top_gate_launch(ccs_ctx* ctx, ccs_jobset* js, const ccs_job_survivor* d_in,
                               const uint32_t* d_in_count, uint32_t in_cap, int32_t top_pos,
                               int32_t reps, ccs_job_survivor* d_out, uint32_t out_cap,
                               uint32_t* d_out_count, void* stream);
int ccs_prove_jobset_launch(ccs_ctx* ctx, ccs_jobset* js, const ccs_job_survivor* d_surv,
                            const uint32_t* d_count, uint32_t cap, const ccs_prove_params* params,
                            ccs_chain_result* d_out, void* stream);
int ccs_batch_launch(ccs_ctx* ctx, const ccs_shape* shape, const ccs_job* h_jobs, uint32_t n_jobs,
                     ccs_job_survivor* d_out, uint32_t cap, uint32_t* d_count, void* stream);
int ccs_batch(ccs_ctx* ctx, const ccs_shape* shape, const ccs_job* h_jobs, uint32_t n_jobs,
              ccs_job_survivor* h_out, uint32_t cap, uint32_t* h_count, ccs_filter_stats* stats);
int ccs_prove_jobs(ccs_ctx* ctx, const ccs_shape* shape, const ccs_job* h_jobs, uint32_t n_jobs,
                   const ccs_job_survivor* h_surv, uint32_t n, const ccs_prove_params* params,
                   ccs_chain_result* h_out);

/* ---- served form: continuous-batching request loop (no counterpart in the reference;
 *      closest is the v16 dispatch ring + prove pool, v16...cu:2938-3279, :5594-5965).
 *      The context is owned by the server's thread until ccs_server_shutdown; submit,
 *      poll, results and drain may be called from any one client thread.  The shapes must
 *      not be retuned while a server uses them. ---- */
typedef struct ccs_server ccs_server;

#define CCS_SERVER_RECOMPUTE_PHASES 1u  /* no resident residue state: every work unit computes
                                           its sieve phases from its u (the comparison mode) */

typedef struct {
    uint64_t mem_budget_bytes;      /* device bytes for admitted jobs' resident state: a job  */
                                    /*   is admitted while a slot of it is free (counters.job_bytes) */
    uint32_t max_active;            /* admitted jobs at most (and slice table size)         */
    uint32_t flags;                 /* CCS_SERVER_*                                         */
    uint32_t max_slice_survivors;   /* survivor buffer per slice and depth group            */
    uint32_t group_width_ratio;     /* depth profiles share a launch per slice while their */
                                    /*   stage-B register widths differ by at most this     */
                                    /*   factor (0: no limit, one launch per row layout;    */
                                    /*   1: one launch per kernel instance)                 */
    uint64_t slice_candidates;      /* target candidates per slice over all admitted jobs   */
    uint64_t min_chunk;             /* least candidates of a job per slice                  */
    uint64_t max_chunk;             /* largest piece of a job in one launch (<= 2^32); a    */
                                    /* job's share of a slice may span several pieces       */
    ccs_prove_params prove;         /* prove settings for every job                         */
} ccs_server_params;

typedef struct {
    uint64_t tag;                   /* caller's identifier                                  */
    int32_t depth;                  /* selects the shape (filter profile) of that depth     */
    int32_t bits;                   /* informational: the width the range was drawn from    */
    uint64_t a;                     /* residue class mod W of the depth's shape              */
    ccs_u128 u0;                    /* n = a + W*(u0 + j), j in [0, count)                   */
    uint64_t count;                 /* below 2^48; sliced by the server                      */
} ccs_request;

#define CCS_JOB_INCOMPLETE 1u       /* survivors of the job were lost to a full stage-B or */
                                    /*   slice buffer: its results are a subset             */
typedef struct {
    uint64_t job_id, tag, candidates;
    uint32_t n_survivors, n_primes, best_len, flags;   /* flags: CCS_JOB_*                  */
    double t_submit, t_admit, t_done;   /* steady-clock seconds */
} ccs_job_done;

typedef struct {
    uint64_t candidates, slices, launches;
    uint64_t resident_bytes;        /* device bytes of the admitted jobs' resident state     */
    uint32_t active, pending;
    uint64_t job_bytes;             /* resident device state of one admitted job: its sieve  */
                                    /*   phases and its rows in both lanes' slice tables     */
    uint64_t fixed_bytes;           /* the server's slice survivor and result buffers        */
    uint32_t slots;                 /* jobs the memory budget admits at once                 */
    uint32_t incomplete;            /* jobs finished with CCS_JOB_INCOMPLETE                 */
    double busy_s;                  /* seconds with admitted work or slices in flight        */
} ccs_server_counters;

int ccs_server_create(ccs_ctx* ctx, const ccs_shape* const* shapes, const int32_t* depths,
                      uint32_t n_shapes, const ccs_server_params* params, ccs_server** out);
int ccs_server_submit(ccs_server* server, const ccs_request* req, uint64_t* job_id);
/* finished jobs since the last poll; CCS_OVERFLOW when one of them is CCS_JOB_INCOMPLETE */
int ccs_server_poll(ccs_server* server, ccs_job_done* h_out, uint32_t cap, uint32_t* n);
/* survivors of a finished job as u offsets within the request (n = a + W*(u0 + u_off)),
 * with their prove results (survivor_idx indexes h_u_off) */
int ccs_server_job_results(ccs_server* server, uint64_t job_id, uint64_t* h_u_off,
                           ccs_chain_result* h_res, uint32_t cap, uint32_t* n);
/* waits until every submitted job has finished; CCS_OVERFLOW when some job so far finished
 * CCS_JOB_INCOMPLETE */
int ccs_server_drain(ccs_server* server);
int ccs_server_stats(ccs_server* server, ccs_server_counters* out);
int ccs_server_shutdown(ccs_server* server);

#ifdef __cplusplus
}
#endif

#endif /* CCS_H */
