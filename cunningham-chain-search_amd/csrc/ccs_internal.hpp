// Internal structures shared by the host runtime (ccs_host.cpp) and the gfx950
// kernels (ccs_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ccs.h"

namespace ccs {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t u8;

// Candidates are organised as residue classes a (mod W) and a run of u: n = a + W*u.
// For the reference lattice n = (tile*WP + k)*M + base (v13.cu:1479-1492) a class is
// (base, k, rho) with tile = 6u + rho, W = 6*WP*M and a = rho*WP*M + k*M + base.

constexpr int kRowBits = 8192;      // one row = 64 lanes x 128 candidates
constexpr int kRowWords = 4;        // kill words (32 candidates each) per lane and row
constexpr int kMaxDense = 64;       // dense primes held one per lane
constexpr int kDeepBuckets = 16;    // forbidden-residue table of a deep prime: 16 buckets
constexpr int kDeepBucketSlots = 4; // of four residues (one 16-byte load)
constexpr int kDeepTableWords = kDeepBuckets * kDeepBucketSlots;
static_assert(kDeepBuckets == 16, "the bucket of r is the top four bits of r * mult");
constexpr int kTierBGroup = 64;     // stage-B primes tested per wave step (one per lane)
constexpr int kMaxTierB = 1024;     // stage-B primes: at most 16 groups of 64
constexpr unsigned kTierBMaxQ = 16384;   // stage-B primes are below this (14 bits)
constexpr unsigned kTierBOffBits = 18;   // bitset word offsets of stage-B primes (< 2^18)
// After the LDS tables (at a 16-byte boundary): the sieve's scheduler words, then, in the
// instances that queue their dense survivors (one and two rows per lane step: the shallow,
// survivor-rich shapes), one candidate queue per wave.  A queue is tested once it holds
// kQueueTest candidates and a word row adds at most 64 at a time, so it holds at most
// kQueueTest + 63; then up to three padding entries and a discard slot.  The four-row
// instances (the deep, VALU-bound shapes) take their survivors one at a time in scalar code.
constexpr unsigned kSchedWords = 8;
constexpr unsigned kQueueTest = 16;
constexpr unsigned kQueueDiscard = kQueueTest + 63 + 3;
constexpr unsigned kQueueWords = 84;    // > kQueueDiscard, a multiple of four
constexpr unsigned kMaxWaves = 16;      // waves per block
static_assert(kQueueDiscard < kQueueWords && kQueueWords % 4 == 0 && kSchedWords % 4 == 0, "queue layout");
__host__ __device__ constexpr bool survivor_queue(int rows_per_group) { return rows_per_group < 4; }
__host__ __device__ constexpr unsigned lds_ctl_words(int rows_per_group) {
    return kSchedWords + (survivor_queue(rows_per_group) ? kMaxWaves * kQueueWords : 0u);
}
// Launch counters of a scratch (16 dwords): words [0, kLaunchWords) are cleared before every
// filter launch ([0] stage-B count, [2..3] unit cursor); words from kStatusWord on are the
// sticky overflow status, cleared only when the caller reads it: [+0] launches whose stage-B
// survivors exceeded the buffer, [+1] the largest such true count, [+2..+3] stage-B survivors
// dropped (u64).
constexpr unsigned kLaunchWords = 8;
constexpr unsigned kStatusWord = 8;
// prove launches over at most this many survivors run as one kernel with a wave per
// survivor (every chain position at once); larger ones as a gate kernel (a lane per
// survivor) and a chain kernel (a wave per prime survivor)
constexpr unsigned kProveWaveMax = 4096;

// Per-modulus constants for exact residues of u64/u128 values: y = sum limb_k*2^(32k)
// folded with c_k = 2^(32k) mod q (< 2^52 for q < 2^18), then reduced with a double
// reciprocal.  q is a prime, or the product of two primes for a paired dense slot.
struct PrimeMod {
    u32 q;
    u32 winv;       // W^-1 mod q (W is prime to q)
    u32 c32, c64, c96;
    u32 pad;
    double rcp;     // 1.0 / q
};

// Dense stage: kill pattern over u, periodic with period q, stored as an extended bit
// string in LDS (16-byte aligned) so that a row (64 lanes x 128 bits) starting at any phase
// s reads the aligned blocks of dwords from ((s>>5) & ~3) + 4*lane, 8 dwords per lane, with
// no wrap-around.
// A dense slot is one prime, or two primes folded into one pattern of period q = q1*q2
// (bit x killed iff x mod q1 or x mod q2 is killed).  The slot's phase is taken modulo q
// directly (W is prime to q), with the slot's own PrimeMod record: pmod[slot].
struct DensePrime {
    u32 q;          // period of the slot's pattern (q1 or q1*q2)
    u32 lds_off;    // dword offset of the extended pattern in LDS
    u32 step;       // (kRowBits * RG) mod q: phase advance per row group
    u32 pad;
};

// Stage B: per-candidate test with a q-bit forbidden bitset in LDS.
struct TierBPrime {
    u32 q;
    u32 lds_off;    // dword offset of the q-bit bitset: in LDS for the first
                    // SieveArgs::nb_lds primes, in SieveArgs::tierb_far for the others
    float rcp;      // 1/q rounded up, for x mod q with x < 2^21
    u32 pad;
};

// Deep stage (K1c): the forbidden residues of n mod q of deep prime i in a table of
// kDeepBuckets buckets of kDeepBucketSlots residues at i * kDeepTableWords: residue r is in
// bucket (r * mult) >> 28 (32-bit product), unused slots hold 0xFFFFFFFF, and the host picks
// the prime's odd multiplier so that no bucket overflows (a prime has at most 32 forbidden
// residues), so one 16-byte load and four compares decide r.  After the nc tables come nc
// 128-bit masks (4 words each): bit (r & 127) is set for every forbidden residue r, so a
// clear bit proves r allowed without the table load (a scattered L2 line per lane).  Shapes
// whose stage-B survivors are many (expected survival above kDeepMaskSurvival) read the
// masks and load the table only where the bit is set; the others load it directly, as the
// extra load costs more than it saves when the survivors die within a few steps.
struct DeepPrime {
    u32 q;
    u32 c32, c64, c96;   // 2^(32k) mod q
    double rq;           // 1/q, correctly rounded
    u32 mult;            // bucket multiplier of the forbidden-residue table
    u32 _pad;
};
constexpr double kDeepMaskSurvival = 1e-6;
constexpr unsigned kDeepBatch = 64;   // deep-stage records per wave batch (survivor-rich launches)
constexpr unsigned kDeepBatchMin = 12;       // smaller batches run slower than one record per wave
constexpr unsigned kDeepFillWaves = 8192;    // resident waves of the deep kernel (256 CUs x 32)

struct ClassRec {   // plain lattice class (base_idx, k, rho) and its residue a mod W
    u64 a;
    u32 base_idx;
    u32 k;
    u32 rho;
    u32 pad;
};

struct JobRec {     // class-form job: n = a + W*(u0 + j), j in [0, count); a and W are u128
    u64 a, a_hi;
    u64 u0_lo, u0_hi;
    u64 count;
    u64 W, W_hi;      // the modulus of the job's shape
    u32 shape;        // the job's shape within a group launch (0 otherwise)
    u32 unit_j;       // candidates per work unit of this job (a launch's unit_j overrides)
    // resident residue state (served jobs): the phases of the job's sieve primes at some u_r
    // are kept in slot ph_slot of the launch's phase arena, and u0 = u_r + ph_off; kNoSlot =
    // none (phases are computed from u0)
    u64 ph_off;
    u32 ph_slot;
    u32 pad;
};
constexpr u32 kNoSlot = 0xffffffffu;

// Resident phases of one job: (u_r + a*W^-1) mod q for every dense slot and stage-B prime of
// its shape, in the order of the shape's PrimeMod table (the dense slots, then stage B).
struct PhaseInit {
    const PrimeMod* pmod;     // the job's shape's residue constants
    u32 n;                    // dense slots + stage-B primes of the shape
    u32 slot;                 // slot in the phase arena
    u64 a, a_hi;
    u64 u_lo, u_hi;           // u_r
};

struct MidRec {     // stage-B survivor: owner (class or job index) and j
    u32 owner;
    u32 j;
};

struct PlainSrc {
    const ClassRec* cls;
    u64 nclasses;
    u64 U0_lo, U0_hi;   // floor(T0 / 6)
    u32 off;            // T0 - 6*U0
    u32 tile_count;
    u32 chunks_per_class;
    u32 pad;
};

struct JobSrc {
    const JobRec* jobs;
    const u64* unit_prefix;   // [n_jobs + 1] first unit index of each job
    u32 n_jobs;
    u32 job_base;             // added to the job index recorded as a survivor's owner
};

// One shape's sieve tables (device pointers).  A launch sieves one shape, or several
// (a group: a pool of families, the depth profiles of a server) whose jobs are grouped by
// shape; a block then loads each shape's tables into LDS as its work reaches that shape.
struct ShapeTab {
    const u32* lds_image;     // dense patterns + tier-B bitsets, copied to LDS
    u32 lds_words;
    u32 nd;                   // dense slots (one or two primes each)
    u32 nb;                   // tier-B primes
    u32 nb_lds;               // tier-B primes with bitsets in LDS (whole groups of 64)
    const DensePrime* dense;
    const TierBPrime* tierb;
    const u32* tierb_far;     // bitsets of stage-B primes >= nb_lds (lds_off indexes here)
    const PrimeMod* pmod;     // [nd + nb] residue constants: dense slot moduli, then tier-B primes
    u32 zero_lds, zero_far;   // a zero word in the LDS image / the device bitsets
};

struct SieveArgs {
    u64 W, W_hi;              // the shape modulus (u128; one-shape launches)
    ShapeTab tab;             // the tables of a one-shape launch
    const ShapeTab* group;    // [n_group] tables of a group launch (null: one shape)
    u32 n_group;
    u32 lds_ctl;              // word offset of the scheduler words (lds_ctl_words) in LDS
    const u64* shape_units;   // [n_group + 1] first unit of each shape's jobs (group launch)
    u32 unit_j;               // candidates per work unit (job launches: 0 = each job's own)
    u32 one;                  // 1 (see the dense loop)
    const u32* phases;        // phase arena of resident jobs (JobRec::ph_slot), or null
    u32 ph_stride;            // words per slot
    // work
    u64 n_units;
    u64 unit_base;            // added to the cursor value: first unit of this launch
    u64* unit_counter;        // work-stealing cursor (zeroed before launch)
    int mode;                 // 0 = plain, 1 = jobs
    PlainSrc plain;
    JobSrc jobsrc;
    // output
    MidRec* mid;
    u32 mid_cap;
    u32* mid_count;
    unsigned long long* stats; // [0] alive after dense stage, [1] tested in stage B, [2] stage-B
                               // survivors accumulated by the deep kernel (may be null)
};

// One shape's deep-stage tables (group launches index them by the job's shape).
struct DeepTab {
    const DeepPrime* deep;
    const u32* hash;
    u32 nc;
    u32 masked;               // read the masks after the hashes
};

struct DeepArgs {
    u64 W, W_hi;              // the shape modulus (u128; plain launches)
    const DeepPrime* deep;
    const u32* hash;
    u32 nc;
    u32 masked;               // read the masks after the hashes (plain / one-shape launches)
    const DeepTab* group;     // per-shape deep tables of a group launch (null: one shape)
    int mode;
    PlainSrc plain;
    JobSrc jobsrc;
    const MidRec* mid;
    const u32* mid_count;
    u32 mid_cap;
    void* out;                // plain: ccs_survivor[16 B]; jobs: ccs_job_survivor[8 B]
    u32 out_cap;
    u32* out_count;
    unsigned long long* stats; // [2] += stage-B survivors of this launch (may be null)
    u32* status;               // sticky overflow status (kStatusWord layout, may be null)
    u32 per_wave;              // records per wave batch (0: from the launch's count)
};

// Prove: GMP confirmation engine semantics (v13.cu:1794-1924) per survivor.
struct ProveArgs {
    int mode;                  // 0 = plain survivors, 1 = job survivors
    // plain reconstruction n = ((T0 + tile_idx)*WP + k)*M + base
    u64 T0_lo, T0_hi;
    u64 WP, M;
    const u64* bases;
    // jobs reconstruction n = a + W*(u0 + u_off) from the job records
    const JobRec* jobs;
    const void* surv;
    const u32* surv_count;
    u32 surv_cap;
    int bits;                  // bit-size gate, 0 = off
    int test_kind;             // 0 = Baillie-PSW (GMP equivalent), 1 = strong MR over bases
    int n_bases;
    u32 bases_mr[16];
    void* out;                 // ccs_chain_result per survivor
    u32* chain_queue;          // [surv_cap] indices of prime survivors (scratch)
    u32* chain_count;          // queue length (scratch, cleared by the launcher)
};

// K1.5 family top-position gate
struct TopGateArgs {
    const JobRec* jobs;        // p_0 = a + W*u of each job (family shapes: W - 1 + W*u)
    const void* in;            // ccs_job_survivor[in_cap]
    const u32* in_count;
    u32 in_cap;
    int top_pos;               // 0..63
    int reps;                  // bases {2, 3, 5}[0..reps)
    void* out;                 // ccs_job_survivor[out_cap]
    u32 out_cap;
    u32* out_count;
};

// the v13 K2 output: chain roots of a plain launch's prove results
struct RootsArgs {
    const ccs_survivor* surv;
    const ccs_chain_result* res;
    const u32* count;
    u32 cap;
    ccs_chain_root* out;
    u32 out_cap;
    u32* out_count;
    u32* primes;              // prime survivors (may be null)
};

// sets the text returned by ccs_last_error() and returns code (ccs_host.cpp)
int set_error(int code, const char* msg);

// kernel launchers (ccs_kernels.hip)
hipError_t launch_sieve(const SieveArgs& a, int rg, int nbv, int grid, int block, size_t lds_bytes,
                        hipStream_t s);
hipError_t launch_deep(const DeepArgs& a, u32 max_entries, hipStream_t s);
hipError_t launch_prove(const ProveArgs& a, u32 max_entries, hipStream_t s);
hipError_t launch_zero(u32* p, size_t n_words, hipStream_t s);
// zeroes up to three ranges of at most 256 words each (null pointers are skipped)
hipError_t launch_zero3(u32* a, u32 na, u32* b, u32 nb, u32* c, u32 nc, hipStream_t s);
hipError_t launch_top_gate(const TopGateArgs& a, hipStream_t s);
hipError_t launch_chain_roots(const RootsArgs& a, hipStream_t s);
hipError_t launch_mark_overflow(const u32* status, u32* count, hipStream_t s);
hipError_t launch_phase_init(const PhaseInit* d_jobs, u32 n_jobs, u32* phases, u32 stride, hipStream_t s);

// ---- served form hooks (ccs_host.cpp, for ccs_server.cpp) ----
// words of resident phase state per job of a shape (dense slots + stage-B primes)
u32 shape_phase_words(const ccs_shape* s);
// the expected fraction of a shape's candidates that survive every filter prime
double shape_final_survival(const ccs_shape* s);
// stage-B register width (groups of 64 primes) of the sieve kernel instance that runs a
// shape with n_tierb stage-B primes: one group per 64 primes, widths above 8 rounded up
// to 12 or 16 (the instances compiled)
inline u32 sieve_stage_b_width(u32 n_tierb) {
    u32 w = n_tierb ? (n_tierb + 63) / 64 : 1;
    if (w > 8) w = w <= 12 ? 12 : 16;
    return w;
}
// the residue constants of a shape (device, dense slots then stage B)
const PrimeMod* shape_pmod(const ccs_shape* s);
// a group-set update whose jobs carry resident phases: job i reads slot slots[i] of the arena,
// whose phases are at u0_i - offs[i]
// the sticky overflow status words of a job set's launch scratch (kStatusWord layout); word
// 0 counts the launches whose stage-B survivors overflowed, and only grows
const u32* jobset_status_words(const ccs_jobset* js);
int groupset_update_resident(ccs_ctx* c, ccs_jobset* js, const ccs_job* jobs, const uint32_t* job_shape,
                             const uint32_t* slots, const uint64_t* offs, uint32_t n_jobs, const u32* d_phases,
                             u32 stride, hipStream_t st);

}  // namespace ccs
