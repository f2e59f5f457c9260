// Host runtime of the MI355X Cunningham-chain pipeline: contexts, shape/table building,
// kernel launches and the C ABI declared in include/ccs.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ccs.h"
#include "ccs_internal.hpp"

namespace ccs {
hipError_t upload_prove_constants(const PrimeMod* small18, const PrimeMod* trial62);
}

using namespace ccs;
typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

namespace ccs {
// error reporting for the other translation units of the library
int set_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace ccs

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(CCS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

extern "C" const char* ccs_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

// Launch scratch of one stream.  Launches queued on different streams of a context may
// run at the same time (the sieve of batch k+1 filling the tail of batch k's sieve and its
// deep stage), so every stream a context launches on gets its own stage-B buffer, counters,
// statistics and prove queue; launches on one stream run in order and share them.
struct StreamScratch {
    hipStream_t st = nullptr;
    MidRec* mid = nullptr;
    u32 mid_cap = 0;
    u32* counters = nullptr;            // launch counters + sticky status (kLaunchWords / kStatusWord)
    unsigned long long* stats = nullptr;   // [0] dense alive [1] stage-B tested
    void* chain = nullptr;              // prove chain queue: [0] count, [16..] indices
    u64 chain_bytes = 0;
};

struct ccs_ctx {
    int device = 0;
    int n_cu = 256;
    u32 max_lds_words = 16384;   // LDS per workgroup, in dwords
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    std::vector<StreamScratch> scratch;   // [0] belongs to `stream`
    void* out = nullptr;                // host-form output staging
    u64 out_bytes = 0;
    void* res = nullptr;                // prove result staging
    u64 res_bytes = 0;
    JobRec* jobs = nullptr;
    u64* unit_prefix = nullptr;
    u32 jobs_cap = 0;
    ccs_filter_stats last = {};
    bool timed = false;
    bool last_timed = false;            // events 0..2 hold the last launch's timing
    // kernel-time collection (ccs_profile_start/stop): event pairs per launch
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> marks;   // (kind, index of the begin event)
    u32 stage_b_limit = 0;              // test knob: stage-B buffer entries used (0 = all)
    u32 deep_per_wave = 0;              // test knob: deep-stage records per wave (0 = by count)
    unsigned char* noalloc = nullptr;   // this device's copy of the no-allocation scratch
};

// begin/end events of one kernel launch while profiling (kind: 0 sieve, 1 deep, 2 prove)
static hipEvent_t prof_event(ccs_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

static void prof_begin(ccs_ctx* c, int kind, hipStream_t st) {
    if (!c->profiling) return;
    size_t idx = c->ev_used;
    hipEvent_t a = prof_event(c);
    if (!a) return;
    (void)hipEventRecord(a, st);
    c->marks.push_back({kind, idx});
}

static void prof_end(ccs_ctx* c, hipStream_t st) {
    if (!c->profiling || c->marks.empty()) return;
    hipEvent_t b = prof_event(c);
    if (b) (void)hipEventRecord(b, st);
}

static int ctx_check(ccs_ctx* c) {
    if (!c) return fail(CCS_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->device));
    return CCS_OK;
}

static void free_scratch(StreamScratch& x) {
    (void)hipFree(x.mid);
    (void)hipFree(x.counters);
    (void)hipFree(x.stats);
    (void)hipFree(x.chain);
    x = StreamScratch();
}

// the scratch of stream st, made on the first launch on it (nullptr after an error)
static StreamScratch* stream_scratch(ccs_ctx* c, hipStream_t st) {
    for (auto& x : c->scratch)
        if (x.st == st) return &x;
    StreamScratch x;
    x.st = st;
    if (hipMalloc((void**)&x.counters, 64) != hipSuccess || hipMalloc((void**)&x.stats, 64) != hipSuccess ||
        hipMemset(x.counters, 0, 64) != hipSuccess) {
        free_scratch(x);
        fail(CCS_ERR_NOMEM, "launch counters");
        return nullptr;
    }
    c->scratch.push_back(x);
    return &c->scratch.back();
}

static u64 mod_pow(u64 b, u64 e, u64 m) {
    u128 r = 1 % m, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (u64)r;
}

// Smallest float >= (1/q)(1 + 2^-23).  For x < 2^21 the f32 product x * rcp, rounded
// to nearest, then truncated, is exactly floor(x / q): the relative excess of rcp plus
// the product's rounding (at most 2^-22 + 2^-24) stays below 1/x, too small to reach the
// next integer, and it is never below x / q.  Checked for every prime < 16384 and every
// x < 2^21 by the test suite.
static float rcp_up(u32 q) {
    const double d = (1.0 / (double)q) * (1.0 + std::ldexp(1.0, -23));
    float f = (float)d;
    while ((double)f < d) f = std::nextafter(f, INFINITY);
    return f;
}

// a^-1 mod m for gcd(a, m) = 1 (extended Euclid)
static u64 mod_inv(u64 a, u64 m) {
    int64_t t = 0, nt = 1, r = (int64_t)m, nr = (int64_t)(a % m);
    while (nr) {
        const int64_t qq = r / nr, tt = t - qq * nt, rr = r - qq * nr;
        t = nt; nt = tt; r = nr; nr = rr;
    }
    return (u64)(t < 0 ? t + (int64_t)m : t);
}

static PrimeMod make_primemod(u32 q, u128 W) {
    PrimeMod pm;
    pm.q = q;
    pm.winv = W ? (u32)mod_inv((u64)(W % q), q) : 0;
    pm.c32 = (u32)(((u128)1 << 32) % q);
    pm.c64 = (u32)(((u128)1 << 64) % q);
    pm.c96 = (u32)(((u128)1 << 96) % q);
    pm.pad = 0;
    pm.rcp = 1.0 / (double)q;
    return pm;
}

// ---- the no-allocation launch (ccs_filter_launch_noalloc): its stage-B buffer, counters
// and statistics live in this file-scope device array; each device holds its own copy until
// the process ends.  Host code uses the address hipGetSymbolAddress gives for the current
// device (kept in the context), never the array's name.
constexpr u32 kNoAllocMid = 1u << 20;        // stage-B entries (8 MiB)
constexpr size_t kNoAllocHead = 256;         // counters (64 B) and statistics (64 B) first
static __device__ __align__(256) unsigned char scratch[kNoAllocHead + (size_t)kNoAllocMid * sizeof(MidRec)];

// once per context, with its device current (ccs_ctx_create)
static int noalloc_setup(ccs_ctx* c) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CCS_ERR_HIP, std::string("no-allocation scratch address: ") + hipGetErrorString(e));
    c->noalloc = (unsigned char*)p;
    return CCS_OK;
}

extern "C" int ccs_device_count(int* out) {
    if (!out) return fail(CCS_ERR_INVALID, "null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return CCS_OK;
}

extern "C" int ccs_ctx_create(int device, ccs_ctx** out) {
    if (!out) return fail(CCS_ERR_INVALID, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CCS_ERR_HIP, "no HIP device");
    if (device < 0 || device >= n) return fail(CCS_ERR_INVALID, "bad device index");
    ccs_ctx* c = new ccs_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(CCS_ERR_HIP, hipGetErrorString(e));
    }
    c->n_cu = prop.multiProcessorCount;
    c->max_lds_words = (u32)(prop.sharedMemPerBlock / 4);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(CCS_ERR_HIP, "stream create");
    }
    for (auto& ev : c->ev) (void)hipEventCreate(&ev);
    if (!stream_scratch(c, c->stream)) {
        const std::string msg = g_err;
        ccs_ctx_destroy(c);
        return fail(CCS_ERR_NOMEM, msg);
    }
    if (noalloc_setup(c) != CCS_OK) {
        const std::string msg = g_err;
        ccs_ctx_destroy(c);
        return fail(CCS_ERR_HIP, msg);
    }
    // prove constants: small primes 2..61, trial primes 67..409
    static const u32 small[18] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    PrimeMod ps[18], pt[62];
    for (int i = 0; i < 18; ++i) ps[i] = make_primemod(small[i], 0);
    int k = 0;
    for (u32 q = 67; k < 62; ++q) {
        bool pr = true;
        for (u32 d = 2; d * d <= q; ++d)
            if (q % d == 0) { pr = false; break; }
        if (pr) pt[k++] = make_primemod(q, 0);
    }
    if (upload_prove_constants(ps, pt) != hipSuccess) {
        ccs_ctx_destroy(c);
        return fail(CCS_ERR_HIP, "prove constants");
    }
    *out = c;
    return CCS_OK;
}

extern "C" void ccs_ctx_destroy(ccs_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto& x : c->scratch) free_scratch(x);
    (void)hipFree(c->out);
    (void)hipFree(c->res);
    (void)hipFree(c->jobs);
    (void)hipFree(c->unit_prefix);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->ev_pool) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int ccs_dmalloc(ccs_ctx* c, void** p, uint64_t bytes) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!p) return fail(CCS_ERR_INVALID, "null");
    HIPCHK(hipMalloc(p, bytes ? bytes : 1));
    return CCS_OK;
}
extern "C" int ccs_dfree(ccs_ctx* c, void* p) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipFree(p));
    return CCS_OK;
}
extern "C" int ccs_memcpy_h2d(ccs_ctx* c, void* d, const void* h, uint64_t bytes) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCS_OK;
}
extern "C" int ccs_memcpy_d2h(ccs_ctx* c, void* h, const void* d, uint64_t bytes) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCS_OK;
}
extern "C" int ccs_memset_d(ccs_ctx* c, void* d, int v, uint64_t bytes) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(d, v, bytes, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCS_OK;
}
extern "C" int ccs_synchronize(ccs_ctx* c) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipDeviceSynchronize());
    return CCS_OK;
}

extern "C" int ccs_stream_create(ccs_ctx* c, void** stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!stream) return fail(CCS_ERR_INVALID, "null argument");
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *stream = (void*)st;
    return CCS_OK;
}

extern "C" int ccs_stream_destroy(ccs_ctx* c, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!stream) return fail(CCS_ERR_INVALID, "null stream");
    if ((hipStream_t)stream == c->stream) return fail(CCS_ERR_INVALID, "the context's own stream");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    for (size_t i = 1; i < c->scratch.size(); ++i)
        if (c->scratch[i].st == (hipStream_t)stream) {
            free_scratch(c->scratch[i]);
            c->scratch.erase(c->scratch.begin() + (long)i);
            break;
        }
    HIPCHK(hipStreamDestroy((hipStream_t)stream));
    return CCS_OK;
}

extern "C" int ccs_event_create(ccs_ctx* c, void** event) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!event) return fail(CCS_ERR_INVALID, "null argument");
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = (void*)e;
    return CCS_OK;
}

extern "C" int ccs_event_destroy(ccs_ctx* c, void* event) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!event) return fail(CCS_ERR_INVALID, "null event");
    HIPCHK(hipEventDestroy((hipEvent_t)event));
    return CCS_OK;
}

extern "C" int ccs_event_record(ccs_ctx* c, void* event, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!event) return fail(CCS_ERR_INVALID, "null event");
    HIPCHK(hipEventRecord((hipEvent_t)event, stream ? (hipStream_t)stream : c->stream));
    return CCS_OK;
}

extern "C" int ccs_event_synchronize(ccs_ctx* c, void* event) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!event) return fail(CCS_ERR_INVALID, "null event");
    HIPCHK(hipEventSynchronize((hipEvent_t)event));
    return CCS_OK;
}

extern "C" int ccs_stream_wait_event(ccs_ctx* c, void* stream, void* event) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!event) return fail(CCS_ERR_INVALID, "null event");
    HIPCHK(hipStreamWaitEvent(stream ? (hipStream_t)stream : c->stream, (hipEvent_t)event, 0));
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// shape
// ---------------------------------------------------------------------------

struct ccs_shape {
    int device = 0;
    u32 max_lds_words = 16384;
    std::vector<u32> crt, wheel, fp, stage_end;
    int target = 0, depth = 0;
    u32 pos0_max_q = 97;
    u32 flags = 0;
    u64 M = 1, WP = 1, W = 1;
    u64 W_hi = 0;                            // the modulus is u128 for family shapes
    bool family = false;                     // n = Q*D*2^E - 1 (ccs_family_shape_build)
    u128 D = 0;
    int exp_start = 0;
    std::vector<u64> bases;
    std::vector<ClassRec> classes;
    u64 per_rho[6] = {0, 0, 0, 0, 0, 0};
    std::vector<std::vector<u32>> forb;      // n-space forbidden residues per filter prime
    ccs_tuning tune = {};
    // device tables
    ClassRec* d_cls = nullptr;
    u64* d_bases = nullptr;
    u32* d_lds_image = nullptr;
    DensePrime* d_dense = nullptr;
    TierBPrime* d_tierb = nullptr;
    u32* d_tierb_far = nullptr;          // stage-B bitsets of the primes after the first 64
    PrimeMod* d_pmod = nullptr;
    DeepPrime* d_deep = nullptr;
    u32* d_hash = nullptr;
    u32 lds_words = 0, nd = 0, nd_primes = 0, nb = 0, nc = 0;   // nd: dense slots
    u32 zero_lds = 0, zero_far = 0;          // zero words of the LDS image / device bitsets
    u32 nb_lds = 0;                          // stage-B primes with their bitsets in LDS
    double est_survival_b = 1.0;             // expected fraction alive after stage B
    u32 deep_masked = 0;                     // the deep stage reads the forbidden masks
    double est_survival_d = 1.0;             // expected fraction alive after the dense stage
    double est_survival_f = 1.0;             // expected fraction alive after every filter prime
};

static bool is_prime_u32(u32 q) {
    if (q < 2) return false;
    for (u32 d = 2; (u64)d * d <= q; ++d)
        if (q % d == 0) return false;
    return true;
}

// F_q: residues (2^-i - 1) mod q for i in [lo, min(hi, q-1)) (v13.cu:771-787)
static u128 wfull(const ccs_shape* s) { return ((u128)s->W_hi << 64) | s->W; }

static std::vector<u32> forbidden(u32 q, int lo, int hi) {
    std::vector<u32> out;
    if (hi > (int)q - 1) hi = (int)q - 1;
    for (int i = lo; i < hi; ++i) {
        u32 inv2i = (u32)mod_pow(mod_pow(2, i, q), q - 2, q);
        u32 r = (inv2i + q - 1) % q;
        if (std::find(out.begin(), out.end(), r) == out.end()) out.push_back(r);
    }
    return out;
}

static void free_device_tables(ccs_shape* s) {
    (void)hipFree(s->d_cls);
    (void)hipFree(s->d_bases);
    (void)hipFree(s->d_lds_image);
    (void)hipFree(s->d_dense);
    (void)hipFree(s->d_tierb);
    (void)hipFree(s->d_tierb_far);
    (void)hipFree(s->d_pmod);
    (void)hipFree(s->d_deep);
    (void)hipFree(s->d_hash);
    s->d_cls = nullptr;
    s->d_bases = nullptr;
    s->d_lds_image = nullptr;
    s->d_dense = nullptr;
    s->d_tierb = nullptr;
    s->d_tierb_far = nullptr;
    s->d_pmod = nullptr;
    s->d_deep = nullptr;
    s->d_hash = nullptr;
}

// words of a dense pattern of period q in LDS: the phase's word ((q-1) >> 5 at most, rounded
// down to a multiple of 4) + 8 words per lane-window + 256 words per further row, padded to
// a multiple of 4 so that every pattern starts 16-byte aligned
static u32 dense_ext_words(u32 q, u32 rg) { return (((q - 1) >> 5) + 256u * rg + 4u + 3u) & ~3u; }

// A dense slot: up to three consecutive dense primes sharing one pattern, whose period is
// their product (at most pair_max <= 2^18, the bound of the slot's residue constants).
struct DenseSlot {
    u32 first, count;   // filter-prime indices [first, first + count)
};

// dense slots: consecutive dense primes folded in threes, else pairs, while the product
// stays <= pair_max (config 2 at 2^16: 37*41*43 = 65231, then 16 pairs)
static std::vector<DenseSlot> dense_slots(const ccs_shape* s, const ccs_tuning* t) {
    std::vector<DenseSlot> out;
    for (u32 i = 0; i < t->n_dense;) {
        u32 c = 1;
        u64 prod = s->fp[i];
        while (c < 3 && i + c < t->n_dense && prod * s->fp[i + c] <= t->pair_max) prod *= s->fp[i + c++];
        out.push_back({i, c});
        i += c;
    }
    return out;
}

static u32 slot_period(const ccs_shape* s, DenseSlot sl) {
    u32 p = 1;
    for (u32 k = 0; k < sl.count; ++k) p *= s->fp[sl.first + k];
    return p;
}

// default split: dense primes below 128 (at most 64), stage B up to 256 primes below 2^12
static int validate_tuning(const ccs_shape* s, ccs_tuning* t);

// Stage-B primes whose bitsets go to LDS: whole groups of 64, as many as fit beside the
// dense patterns in the block's share of the CU's LDS (the rest stay in device memory);
// ~0u when even the dense patterns and the first group do not fit.
static u32 tierb_in_lds(const ccs_shape* s, const ccs_tuning* t) {
    u32 words = 0;
    for (auto sl : dense_slots(s, t)) words += dense_ext_words(slot_period(s, sl), t->rows_per_group);
    const u32 bpc = t->blocks_per_cu ? t->blocks_per_cu : 1;
    // the zero word, the alignment of the control words and the control words follow
    const u32 ctl = lds_ctl_words((int)t->rows_per_group) + 4;
    const u32 cap = s->max_lds_words - ctl;
    const u32 budget = std::min(cap, std::max(s->max_lds_words / bpc, std::min(s->max_lds_words, 16384u)) - ctl);
    u32 n = 0;
    while (n < t->n_tierb) {
        const u32 end = std::min(n + (u32)kTierBGroup, t->n_tierb);
        u32 w = 0;
        for (u32 i = n; i < end; ++i) w += (s->fp[t->n_dense + i] + 31) / 32;
        if (words + w > budget) break;
        words += w;
        n = end;
    }
    if (words > cap || (t->n_tierb && n == 0)) return ~0u;
    return n;
}

// Starting point measured on the 31# depth-18 workload (config 2) and the 37# depth-19 one
// (config 3) by sweeps of the sieve with 128-bit dense reads: dense primes below 240 (320 at
// shallow depths) folded up to period 2^16, four rows of 8192 candidates per step, stage B
// sized by the expected survival (below), 2^20-candidate units, one block of 16 waves per
// CU.  Shrunk until the tables fit the device's LDS (four rows take 4 KiB of pattern
// extension per slot, so dense-heavy shapes fall back to two rows).
static void default_tuning(const ccs_shape* s, ccs_tuning* t) {
    u32 nd = 0;
    while (nd < s->fp.size() && nd < 2u * (u32)kMaxDense && s->fp[nd] < 240) ++nd;
    // At a shallow depth many candidates survive the dense primes, and stage B's cost per
    // survivor dwarfs a dense pattern's per candidate: when more than 0.25 % survive the
    // dense primes, the dense stage takes the primes below 320 as well (measured on the 31#
    // profiles: depths 12 and 14 -24 %).
    double sd = 1.0;
    for (u32 i = 0; i < nd; ++i) sd *= 1.0 - (double)s->forb[i].size() / s->fp[i];
    if (sd > 2.5e-3)
        while (nd < s->fp.size() && nd < 2u * (u32)kMaxDense && s->fp[nd] < 320) ++nd;
    // Stage B: the smallest multiple of 64 primes (at least 128) after which at most 1e-6 of
    // the candidates are expected to survive, up to 1024 primes below 16384; the deep stage
    // takes the rest (with its masks above 1e-6 survival).  Measured per 2^36 candidates:
    // config 2 320 primes 4.16 ms per step (256: 4.23, 192: 4.27, 128: 4.36); config 3 192
    // primes 20.75 ms per 2^38 (320: 21.28); the shallow 31# depths 12 and 14 take 1024.
    u32 nb = 0;
    double surv = 1.0;
    for (u32 i = 0; i < nd; ++i) surv *= 1.0 - (double)s->forb[i].size() / s->fp[i];
    while (nd + nb < s->fp.size() && nb < (u32)kMaxTierB && s->fp[nd + nb] < kTierBMaxQ) {
        surv *= 1.0 - (double)s->forb[nd + nb].size() / s->fp[nd + nb];
        ++nb;
        if (nb >= 128 && nb % 64 == 0 && surv <= 1e-6) break;
    }
    t->n_dense = nd;
    t->n_tierb = nb;
    t->rows_per_group = 4;
    t->unit_j = 1u << 20;
    t->blocks_per_cu = 1;
    t->pair_max = 1u << 16;
    t->waves_per_block = 16;
    for (;;) {
        ccs_tuning tt = *t;
        if (validate_tuning(s, &tt) == CCS_OK) return;
        if (t->rows_per_group > 1) t->rows_per_group >>= 1;
        else if (t->pair_max > 4096) t->pair_max >>= 2;
        else if (t->n_tierb > 16) t->n_tierb -= 16;
        else if (t->n_dense > 8) --t->n_dense;
        else return;
    }
}

static int validate_tuning(const ccs_shape* s, ccs_tuning* t) {
    if (t->rows_per_group != 1 && t->rows_per_group != 2 && t->rows_per_group != 4)
        return fail(CCS_ERR_INVALID, "rows_per_group must be 1, 2 or 4");
    if (t->n_dense > 2u * kMaxDense) return fail(CCS_ERR_INVALID, "n_dense > 128");
    if (t->pair_max > (1u << 18)) return fail(CCS_ERR_INVALID, "pair_max > 2^18");
    const auto slots = dense_slots(s, t);
    if (slots.size() > (size_t)kMaxDense) return fail(CCS_ERR_INVALID, "more than 64 dense patterns");
    if (t->n_tierb > (u32)kMaxTierB) return fail(CCS_ERR_INVALID, "n_tierb > 1024");
    if (t->n_dense + t->n_tierb > s->fp.size())
        return fail(CCS_ERR_INVALID, "n_dense + n_tierb exceeds the filter primes");
    if (t->unit_j == 0 || t->unit_j % kRowBits || t->unit_j > (1u << 20))
        return fail(CCS_ERR_INVALID, "unit_j must be a multiple of 8192, <= 2^20");
    for (u32 i = 0; i < t->n_dense + t->n_tierb; ++i)
        if (s->fp[i] >= kTierBMaxQ) return fail(CCS_ERR_INVALID, "dense/stage-B primes must be < 16384");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 16) return fail(CCS_ERR_INVALID, "blocks_per_cu");
    if (t->waves_per_block == 0) t->waves_per_block = 4;
    if (t->waves_per_block != 1 && t->waves_per_block != 2 && t->waves_per_block != 4 &&
        t->waves_per_block != 8 && t->waves_per_block != 16)
        return fail(CCS_ERR_INVALID, "waves_per_block must be 1, 2, 4, 8 or 16");
    if (tierb_in_lds(s, t) == ~0u) return fail(CCS_ERR_INVALID, "dense + stage-B tables exceed LDS");
    return CCS_OK;
}

// Build the device tables for the current tuning.
static int upload_tables(ccs_shape* s) {
    const ccs_tuning& t = s->tune;
    const u32 nd = t.n_dense, nb = t.n_tierb, nfp = (u32)s->fp.size(), nc = nfp - nd - nb;
    const u32 rg = t.rows_per_group;
    const auto slots = dense_slots(s, &t);
    const u32 nb_lds = tierb_in_lds(s, &t);
    if (nb_lds == ~0u) return fail(CCS_ERR_INVALID, "dense + stage-B tables exceed LDS");
    std::vector<u32> image;
    std::vector<DensePrime> dense(slots.empty() ? 1 : slots.size());
    std::vector<TierBPrime> tierb(nb ? nb : 1);
    std::vector<PrimeMod> pmod(slots.size() + nb ? slots.size() + nb : 1);
    // u-space kill sets: u + (a * W^-1) = f * W^-1 (mod q)  for f in F_q
    auto uset = [&](u32 i) {
        const u32 q = s->fp[i];
        const u64 winv = mod_pow((u64)(wfull(s) % q), q - 2, q);
        std::vector<u8> kill(q, 0);
        for (u32 f : s->forb[i]) kill[(u64)f * winv % q] = 1;
        return kill;
    };
    for (size_t k = 0; k < slots.size(); ++k) {
        const DenseSlot sl = slots[k];
        const u32 Q = slot_period(s, sl);
        std::vector<std::vector<u8>> ks;
        std::vector<u32> qs;
        for (u32 m = 0; m < sl.count; ++m) {
            ks.push_back(uset(sl.first + m));
            qs.push_back(s->fp[sl.first + m]);
        }
        const u32 words = dense_ext_words(Q, rg);
        DensePrime d;
        d.q = Q;
        d.lds_off = (u32)image.size();
        d.step = (u32)(((u64)kRowBits * rg) % Q);
        d.pad = 0;
        dense[k] = d;
        pmod[k] = make_primemod(Q, wfull(s));
        for (u32 w = 0; w < words; ++w) {
            u32 v = 0;
            for (u32 b = 0; b < 32; ++b) {
                const u64 x = (u64)w * 32 + b;
                bool kill = false;
                for (u32 m = 0; m < sl.count; ++m) kill = kill || ks[m][x % qs[m]];
                if (kill) v |= 1u << b;
            }
            image.push_back(v);
        }
    }
    // stage-B bitsets: the first groups of 64 primes (the first is tested for every dense
    // survivor) in LDS as far as it holds them, the later groups in device memory
    std::vector<u32> far_bits;
    for (u32 i = 0; i < nb; ++i) {
        const u32 q = s->fp[nd + i];
        std::vector<u8> kill = uset(nd + i);
        std::vector<u32>& dst = i < nb_lds ? image : far_bits;
        TierBPrime b;
        b.q = q;
        b.lds_off = (u32)dst.size();
        b.rcp = rcp_up(q);
        b.pad = 0;
        tierb[i] = b;
        for (u32 w = 0; w < (q + 31) / 32; ++w) {
            u32 v = 0;
            for (u32 bb = 0; bb < 32; ++bb) {
                u32 x = w * 32 + bb;
                if (x < q && kill[x]) v |= 1u << bb;
            }
            dst.push_back(v);
        }
        pmod[slots.size() + i] = make_primemod(q, wfull(s));
    }
    // a zero word in each memory: the stage-B lanes past the last prime probe it
    s->zero_lds = (u32)image.size();
    image.push_back(0);
    s->zero_far = (u32)far_bits.size();
    far_bits.push_back(0);
    // the sieve packs a stage-B prime's bitset offset into kTierBOffBits bits beside q
    if (far_bits.size() > (1u << kTierBOffBits) || image.size() > (1u << kTierBOffBits))
        return fail(CCS_ERR_INVALID, "stage-B bitsets exceed 2^18 words: fewer stage-B primes");
    std::vector<DeepPrime> deep(nc ? nc : 1);
    // nc bucket tables (one empty one when nc = 0), then the nc masks
    const size_t hash_words = (size_t)(nc ? nc : 1) * kDeepTableWords;
    std::vector<u32> hash(hash_words + (size_t)nc * 4, 0xffffffffu);
    std::fill(hash.begin() + hash_words, hash.end(), 0u);
    for (u32 i = 0; i < nc; ++i) {
        const u32 idx = nd + nb + i;
        const PrimeMod pm = make_primemod(s->fp[idx], wfull(s));
        DeepPrime d;
        d.q = pm.q;
        d.c32 = pm.c32;
        d.c64 = pm.c64;
        d.c96 = pm.c96;
        d.rq = 1.0 / (double)pm.q;
        d._pad = 0;
        // the first odd multiplier of a fixed sequence that leaves no bucket over-full
        const std::vector<u32>& F = s->forb[idx];
        if (F.size() > (size_t)kDeepTableWords / 2) return fail(CCS_ERR_INVALID, "too many forbidden residues");
        d.mult = 0;
        for (u32 k = 0; k < 65536 && !d.mult; ++k) {
            const u32 m = (k * 0x9E3779B9u + 0x7F4A7C15u) | 1u;
            u32 fill[kDeepBuckets] = {};
            bool ok = true;
            for (u32 f : F) ok = ok && ++fill[(f * m) >> 28] <= (u32)kDeepBucketSlots;
            if (ok) d.mult = m;
        }
        if (!d.mult) return fail(CCS_ERR_INVALID, "no bucket multiplier for a deep prime");
        u32* tab = &hash[(size_t)i * kDeepTableWords];
        u32* mask = &hash[(size_t)nc * kDeepTableWords + (size_t)i * 4];
        for (u32 f : F) {
            u32* b = tab + ((f * d.mult) >> 28) * kDeepBucketSlots;
            u32 k = 0;
            while (b[k] != 0xffffffffu) ++k;   // < kDeepBucketSlots: the multiplier fits them
            b[k] = f;
            mask[(f >> 5) & 3] |= 1u << (f & 31);
        }
        deep[i] = d;
    }

    // expected survival after the GPU's stage B (sizes the stage-B survivor buffer)
    double surv = 1.0;
    for (u32 i = 0; i < nd + nb; ++i) {
        surv *= 1.0 - (double)s->forb[i].size() / s->fp[i];
        if (i + 1 == nd) s->est_survival_d = surv;
    }
    if (nd == 0) s->est_survival_d = 1.0;
    s->est_survival_b = surv;
    for (size_t i = nd + nb; i < s->fp.size(); ++i) surv *= 1.0 - (double)s->forb[i].size() / s->fp[i];
    s->est_survival_f = surv;
    s->deep_masked = surv > kDeepMaskSurvival ? 1u : 0u;

    free_device_tables(s);
    HIPCHK(hipSetDevice(s->device));
    auto up = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(dst, bytes ? bytes : 16);
        if (e != hipSuccess) return e;
        if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        return e;
    };
    HIPCHK(up((void**)&s->d_lds_image, image.data(), image.size() * 4));
    HIPCHK(up((void**)&s->d_dense, dense.data(), dense.size() * sizeof(DensePrime)));
    HIPCHK(up((void**)&s->d_tierb, tierb.data(), tierb.size() * sizeof(TierBPrime)));
    HIPCHK(up((void**)&s->d_tierb_far, far_bits.data(), far_bits.size() * 4));
    HIPCHK(up((void**)&s->d_pmod, pmod.data(), pmod.size() * sizeof(PrimeMod)));
    HIPCHK(up((void**)&s->d_deep, deep.data(), deep.size() * sizeof(DeepPrime)));
    HIPCHK(up((void**)&s->d_hash, hash.data(), hash.size() * 4));
    HIPCHK(up((void**)&s->d_bases, s->bases.data(), s->bases.size() * 8));
    HIPCHK(up((void**)&s->d_cls, s->classes.data(), s->classes.size() * sizeof(ClassRec)));
    s->lds_words = (u32)image.size();
    s->nd = (u32)slots.size();
    s->nd_primes = nd;
    s->nb = nb;
    s->nb_lds = nb_lds;
    s->nc = nc;
    return CCS_OK;
}

extern "C" int ccs_shape_build(ccs_ctx* c, const ccs_shape_params* p, const ccs_tuning* tuning,
                               ccs_shape** out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!p || !out) return fail(CCS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (p->depth < 1 || p->depth > 32) return fail(CCS_ERR_INVALID, "depth must be in 1..32");
    if (p->target < 1 || p->target > 64) return fail(CCS_ERR_INVALID, "target must be in 1..64");
    if (p->n_crt > 16 || p->n_wheel > 8) return fail(CCS_ERR_INVALID, "too many CRT/wheel primes");
    ccs_shape* s = new ccs_shape();
    s->device = c->device;
    s->max_lds_words = c->max_lds_words;
    s->target = p->target;
    s->depth = p->depth;
    s->pos0_max_q = p->pos0_max_q;
    s->flags = p->flags;
    u32 last = 3;
    for (u32 i = 0; i < p->n_crt; ++i) {
        u32 q = p->crt_primes[i];
        if (!is_prime_u32(q) || q <= last) { delete s; return fail(CCS_ERR_INVALID, "CRT primes must be ascending primes > 3"); }
        last = q;
        s->crt.push_back(q);
        s->M *= q;
    }
    for (u32 i = 0; i < p->n_wheel; ++i) {
        u32 q = p->wheel_primes[i];
        if (!is_prime_u32(q) || q <= last) { delete s; return fail(CCS_ERR_INVALID, "wheel primes must be ascending primes above the CRT primes"); }
        last = q;
        s->wheel.push_back(q);
        s->WP *= q;
    }
    if ((u128)6 * s->WP * s->M >= ((u128)1 << 63)) { delete s; return fail(CCS_ERR_INVALID, "6*WP*M must be < 2^63"); }
    s->W = 6 * s->WP * s->M;
    for (u32 i = 0; i < p->n_filter; ++i) {
        u32 q = p->filter_primes[i];
        if (!is_prime_u32(q) || q <= last || q >= (1u << 18)) {
            delete s;
            return fail(CCS_ERR_INVALID, "filter primes must be ascending primes above the wheel primes and < 2^18");
        }
        last = q;
        s->fp.push_back(q);
        s->forb.push_back(forbidden(q, q <= p->pos0_max_q ? 0 : 1, p->depth));
    }
    for (u32 i = 0; i < p->n_stages && p->stage_end; ++i) s->stage_end.push_back(p->stage_end[i]);

    // CRT bases mod M admissible at target (v13.cu:907-935)
    std::vector<std::vector<u8>> cforb;
    for (u32 q : s->crt) {
        std::vector<u8> f(q, 0);
        for (u32 r : forbidden(q, 0, s->target)) f[r] = 1;
        cforb.push_back(f);
    }
    for (u64 r = 0; r < s->M; ++r) {
        bool ok = true;
        for (size_t i = 0; i < s->crt.size() && ok; ++i) ok = !cforb[i][r % s->crt[i]];
        if (ok) s->bases.push_back(r);
    }
    if (!(p->flags & CCS_SHAPE_NO_CLASSES)) {
        // wheel offsets per base (v13.cu:977-1008) and the mod-6 tile residue (v13.cu:963-975)
        std::vector<std::vector<u8>> wforb;
        for (u32 q : s->wheel) {
            std::vector<u8> f(q, 0);
            for (u32 r : forbidden(q, 0, s->target)) f[r] = 1;
            wforb.push_back(f);
        }
        const u64 WPM = s->WP * s->M;
        for (size_t b = 0; b < s->bases.size(); ++b) {
            const u64 base = s->bases[b];
            std::vector<u32> bq(s->wheel.size()), mq(s->wheel.size());
            for (size_t i = 0; i < s->wheel.size(); ++i) {
                bq[i] = (u32)(base % s->wheel[i]);
                mq[i] = (u32)(s->M % s->wheel[i]);
            }
            for (u64 k = 0; k < s->WP; ++k) {
                bool ok = true;
                for (size_t i = 0; i < s->wheel.size() && ok; ++i) {
                    const u32 q = s->wheel[i];
                    ok = !wforb[i][(bq[i] + (k % q) * mq[i]) % q];
                }
                if (!ok) continue;
                u32 rho = 6;
                for (u32 t = 0; t < 6; ++t)
                    if ((t * (WPM % 6) + (k % 6) * (s->M % 6) + base % 6) % 6 == 5) { rho = t; break; }
                if (rho == 6) continue;   // cannot happen: WP*M is a unit mod 6
                ClassRec cr;
                cr.a = rho * WPM + k * s->M + base;
                cr.base_idx = (u32)b;
                cr.k = (u32)k;
                cr.rho = rho;
                cr.pad = 0;
                s->classes.push_back(cr);
                s->per_rho[rho]++;
            }
        }
    }
    ccs_tuning t;
    if (tuning) t = *tuning;
    else default_tuning(s, &t);
    rc = validate_tuning(s, &t);
    if (rc) { delete s; return rc; }
    s->tune = t;
    rc = upload_tables(s);
    if (rc) { free_device_tables(s); delete s; return rc; }
    *out = s;
    return CCS_OK;
}

// Family shape: candidates n = Q*D*2^E - 1 as the class W - 1 mod W = D*2^E.  Forbidden
// sets at every position [0, depth) of every kept prime, as v16 builds its per-V masks
// (v16...cu:5104-5227: Q = inv(D*2^(E+j)) mod q, which is n = 2^-j - 1 mod q in n-space).
extern "C" int ccs_family_shape_build(ccs_ctx* c, const ccs_family_params* p, const ccs_tuning* tuning,
                                      ccs_shape** out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!p || !out || (p->n_filter && !p->filter_primes)) return fail(CCS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (p->depth < 1 || p->depth > 32) return fail(CCS_ERR_INVALID, "depth must be in 1..32");
    if (p->target < 1 || p->target > 64) return fail(CCS_ERR_INVALID, "target must be in 1..64");
    if (p->exp_start < 0 || p->exp_start > 126) return fail(CCS_ERR_INVALID, "exp_start must be in 0..126");
    const u128 D = ((u128)p->D.hi << 64) | p->D.lo;
    if (D == 0) return fail(CCS_ERR_INVALID, "D must be >= 1");
    if (D >> (127 - p->exp_start)) return fail(CCS_ERR_INVALID, "D*2^E must be below 2^127");
    const u128 W = D << p->exp_start;
    ccs_shape* s = new ccs_shape();
    s->device = c->device;
    s->max_lds_words = c->max_lds_words;
    s->target = p->target;
    s->depth = p->depth;
    s->pos0_max_q = 0xffffffffu;
    s->flags = CCS_SHAPE_NO_CLASSES;
    s->family = true;
    s->D = D;
    s->exp_start = p->exp_start;
    s->W = (u64)W;
    s->W_hi = (u64)(W >> 64);
    u32 last = 2;
    for (u32 i = 0; i < p->n_filter; ++i) {
        const u32 q = p->filter_primes[i];
        if (!is_prime_u32(q) || q <= last || q >= (1u << 18)) {
            delete s;
            return fail(CCS_ERR_INVALID, "filter primes must be ascending odd primes < 2^18");
        }
        last = q;
        if (D % q == 0) continue;   // q | D: p_j = -1 (mod q) for every j, nothing to sieve
        s->fp.push_back(q);
        s->forb.push_back(forbidden(q, 0, p->depth));
    }
    ccs_tuning t;
    if (tuning) t = *tuning;
    else default_tuning(s, &t);
    rc = validate_tuning(s, &t);
    if (rc) { delete s; return rc; }
    s->tune = t;
    rc = upload_tables(s);
    if (rc) { free_device_tables(s); delete s; return rc; }
    *out = s;
    return CCS_OK;
}

extern "C" void ccs_shape_free(ccs_shape* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    free_device_tables(s);
    delete s;
}

extern "C" int ccs_shape_set_tuning(ccs_shape* s, const ccs_tuning* t) {
    if (!s || !t) return fail(CCS_ERR_INVALID, "null");
    ccs_tuning tt = *t;
    int rc = validate_tuning(s, &tt);
    if (rc) return rc;
    s->tune = tt;
    return upload_tables(s);
}

extern "C" int ccs_shape_get_tuning(const ccs_shape* s, ccs_tuning* out) {
    if (!s || !out) return fail(CCS_ERR_INVALID, "null");
    *out = s->tune;
    return CCS_OK;
}

extern "C" int ccs_shape_query(const ccs_shape* s, ccs_shape_info* o) {
    if (!s || !o) return fail(CCS_ERR_INVALID, "null");
    o->M = s->M;
    o->WP = s->WP;
    o->W = s->W;
    o->W_hi = s->W_hi;
    o->family = s->family ? 1 : 0;
    o->exp_start = s->exp_start;
    o->D.lo = (u64)s->D;
    o->D.hi = (u64)(s->D >> 64);
    o->n_bases = (u32)s->bases.size();
    o->n_filter = (u32)s->fp.size();
    o->n_classes = s->classes.size();
    o->n_dense = s->nd_primes;
    o->n_tierb = s->nb;
    o->n_deep = s->nc;
    o->rows_per_group = s->tune.rows_per_group;
    o->unit_j = s->tune.unit_j;
    o->n_slots = s->nd;
    o->n_tierb_lds = s->nb_lds;
    return CCS_OK;
}

extern "C" int ccs_shape_bases(const ccs_shape* s, uint64_t* out, uint32_t cap) {
    if (!s || !out) return fail(CCS_ERR_INVALID, "null");
    u32 n = std::min<u32>(cap, (u32)s->bases.size());
    memcpy(out, s->bases.data(), (size_t)n * 8);
    return n < s->bases.size() ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_shape_classes(const ccs_shape* s, uint64_t* a, uint32_t* b, uint32_t* k, uint32_t* rho,
                                 uint64_t cap) {
    if (!s) return fail(CCS_ERR_INVALID, "null");
    u64 n = std::min<u64>(cap, s->classes.size());
    for (u64 i = 0; i < n; ++i) {
        if (a) a[i] = s->classes[i].a;
        if (b) b[i] = s->classes[i].base_idx;
        if (k) k[i] = s->classes[i].k;
        if (rho) rho[i] = s->classes[i].rho;
    }
    return n < s->classes.size() ? CCS_OVERFLOW : CCS_OK;
}

static u128 mk128(ccs_u128 v) { return ((u128)v.hi << 64) | v.lo; }

// candidates of a tile range: the v33 counting convention (cc_gmp_v33_03.c:1900)
extern "C" uint64_t ccs_count_candidates(const ccs_shape* s, ccs_u128 ts, uint32_t tile_count) {
    if (!s) return 0;
    const u128 T0 = mk128(ts);
    const u32 r0 = (u32)(T0 % 6);
    u64 total = 0;
    for (u32 m = 0; m < 6; ++m) {
        const u64 first = (m + 6 - r0) % 6;
        const u64 cnt = tile_count > first ? ((u64)tile_count - first + 5) / 6 : 0;
        total += cnt * s->per_rho[m];
    }
    return total;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------

static int ensure_mid(StreamScratch* x, u64 want) {
    if (want < (1u << 16)) want = 1u << 16;
    if (want > (1u << 28)) want = 1u << 28;
    if (x->mid_cap >= want) return CCS_OK;
    // the stream's earlier launches may still read the old buffer
    HIPCHK(hipStreamSynchronize(x->st));
    (void)hipFree(x->mid);
    x->mid = nullptr;
    x->mid_cap = 0;
    HIPCHK(hipMalloc((void**)&x->mid, want * sizeof(MidRec)));
    x->mid_cap = (u32)want;
    return CCS_OK;
}

// stage-B buffer, counters ([0] stage-B count, [2..3] unit cursor) and statistics of a launch
struct Scratch {
    MidRec* mid;
    u32 mid_cap;
    u32* counters;
    unsigned long long* stats;
};

static Scratch to_scratch(const StreamScratch* x) { return Scratch{x->mid, x->mid_cap, x->counters, x->stats}; }

static Scratch noalloc_scratch(const ccs_ctx* c) {
    Scratch x;
    x.counters = (u32*)c->noalloc;
    x.stats = (unsigned long long*)(c->noalloc + 64);
    x.mid = (MidRec*)(c->noalloc + kNoAllocHead);
    x.mid_cap = kNoAllocMid;
    return x;
}

// the stage-B capacity a launch may use: all of the buffer, or less under the test knob
static Scratch limited(const ccs_ctx* c, Scratch x) {
    if (c->stage_b_limit && c->stage_b_limit < x.mid_cap) x.mid_cap = c->stage_b_limit;
    return x;
}

static ShapeTab shape_tab(const ccs_shape* s) {
    ShapeTab t;
    t.lds_image = s->d_lds_image;
    t.lds_words = s->lds_words;
    t.nd = s->nd;
    t.nb = s->nb;
    t.nb_lds = s->nb_lds;
    t.dense = s->d_dense;
    t.tierb = s->d_tierb;
    t.tierb_far = s->d_tierb_far;
    t.pmod = s->d_pmod;
    t.zero_lds = s->zero_lds;
    t.zero_far = s->zero_far;
    return t;
}

static DeepTab deep_tab(const ccs_shape* s) {
    DeepTab t;
    t.deep = s->d_deep;
    t.hash = s->d_hash;
    t.nc = s->nc;
    t.masked = s->deep_masked;
    return t;
}

static void fill_sieve_common(const ccs_shape* s, const Scratch& x, SieveArgs& a) {
    memset(&a, 0, sizeof(a));
    a.W = s->W;
    a.W_hi = s->W_hi;
    a.tab = shape_tab(s);
    a.group = nullptr;
    a.n_group = 1;
    a.lds_ctl = (s->lds_words + 3) & ~3u;
    a.one = 1;
    a.unit_j = s->tune.unit_j;
    a.phases = nullptr;
    a.ph_stride = 0;
    a.unit_counter = (u64*)(x.counters + 2);
    a.mid = x.mid;
    a.mid_cap = x.mid_cap;
    a.mid_count = x.counters + 0;
    a.stats = x.stats;
}

// zero the launch counters, then sieve + deep.  zero_out also clears the caller's count
// (the per-job loop appends several launches into one output).
// (fixed: the launch runs in the no-allocation scratch: its status words are cleared with
// the counters, and no timing events are recorded)
static int run_filter_kernels(ccs_ctx* c, const ccs_shape* s, const Scratch& x, SieveArgs& sa, DeepArgs& da,
                              u64 n_units, u32* d_count, hipStream_t st, bool zero_out = true,
                              bool zero_stats = true, u32 nb_max = 0, bool fixed = false) {
    if (n_units >> 32) return fail(CCS_ERR_INVALID, "a launch takes fewer than 2^32 work units");
    // fill kernels rather than memset nodes, so the sequence can be graph-captured
    if (launch_zero3(x.counters, fixed ? 16u : kLaunchWords, zero_stats ? (u32*)x.stats : nullptr, 16,
                     zero_out ? d_count : nullptr, 1, st) != hipSuccess)
        return fail(CCS_ERR_HIP, "zero");
    const bool timed = c->timed && !fixed, prof = c->profiling;
    if (fixed) c->profiling = false;
    sa.n_units = n_units;
    // a group launch takes its stage-B register width from its widest shape (nb_max)
    const u32 nbm = nb_max > s->nb ? nb_max : s->nb;
    const int nbv = nbm ? (int)((nbm + 63) / 64) : 1;
    const u32 wpb = s->tune.waves_per_block;
    const int block = 64 * (int)wpb;
    const size_t lds = (size_t)(sa.lds_ctl + lds_ctl_words((int)s->tune.rows_per_group)) * 4;
    u64 want_grid = (u64)c->n_cu * s->tune.blocks_per_cu;
    const u64 blocks_needed = (n_units + wpb - 1) / wpb;
    if (want_grid > blocks_needed) want_grid = blocks_needed ? blocks_needed : 1;
    c->last_timed = timed;
    if (timed) HIPCHK(hipEventRecord(c->ev[0], st));
    prof_begin(c, 0, st);
    hipError_t e = launch_sieve(sa, (int)s->tune.rows_per_group, nbv, (int)want_grid, block, lds, st);
    if (e != hipSuccess) {
        c->profiling = prof;
        return fail(CCS_ERR_HIP, std::string("sieve launch: ") + hipGetErrorString(e));
    }
    prof_end(c, st);
    if (timed) HIPCHK(hipEventRecord(c->ev[1], st));
    da.mid = x.mid;
    da.mid_count = x.counters + 0;
    da.mid_cap = x.mid_cap;
    da.out_count = d_count;
    da.stats = x.stats;
    da.status = x.counters + kStatusWord;
    da.per_wave = c->deep_per_wave;
    prof_begin(c, 1, st);
    e = launch_deep(da, x.mid_cap, st);
    c->profiling = prof;
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("deep launch: ") + hipGetErrorString(e));
    prof_end(c, st);
    if (timed) HIPCHK(hipEventRecord(c->ev[2], st));
    return CCS_OK;
}

// fixed: the no-allocation scratch to run in (null: the stream's own launch scratch)
static int filter_launch_impl(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, uint32_t tile_count,
                              ccs_survivor* d_out, uint32_t cap, uint32_t* d_count, hipStream_t st,
                              const Scratch* fixed = nullptr) {
    if (!s || !d_out || !d_count) return fail(CCS_ERR_INVALID, "null argument");
    if (s->classes.empty()) return fail(CCS_ERR_INVALID, "shape built without classes");
    const u128 T0 = mk128(ts);
    const u128 U0 = T0 / 6;
    const u32 off = (u32)(T0 - U0 * 6);
    const u64 jmax = ((u64)tile_count + off + 5) / 6;
    const u64 cpc = (jmax + s->tune.unit_j - 1) / s->tune.unit_j;
    const u64 n_units = cpc * s->classes.size();
    if (n_units >> 32) return fail(CCS_ERR_INVALID, "launch has 2^32 work units or more: fewer tiles per launch");
    const u64 cand = ccs_count_candidates(s, ts, tile_count);
    Scratch x;
    if (fixed) {
        const double want = cand * s->est_survival_b * 2.0 + (1u << 16);
        if (want > fixed->mid_cap)
            return fail(CCS_ERR_INVALID, "tile range too large for the no-allocation scratch: " +
                                             std::to_string((u64)want) + " stage-B entries wanted, " +
                                             std::to_string(fixed->mid_cap) + " held");
        x = *fixed;
    } else {
        StreamScratch* ss = stream_scratch(c, st);
        if (!ss) return CCS_ERR_NOMEM;
        int rc = ensure_mid(ss, (u64)(cand * s->est_survival_b * 2.0) + (1u << 20));
        if (rc) return rc;
        x = limited(c, to_scratch(ss));
    }
    SieveArgs sa;
    fill_sieve_common(s, x, sa);
    sa.mode = 0;
    sa.plain.cls = s->d_cls;
    sa.plain.nclasses = s->classes.size();
    sa.plain.U0_lo = (u64)U0;
    sa.plain.U0_hi = (u64)(U0 >> 64);
    sa.plain.off = off;
    sa.plain.tile_count = tile_count;
    sa.plain.chunks_per_class = (u32)cpc;
    DeepArgs da;
    memset(&da, 0, sizeof(da));
    da.W = s->W;
    da.W_hi = s->W_hi;
    da.deep = s->d_deep;
    da.hash = s->d_hash;
    da.nc = s->nc;
    da.masked = s->deep_masked;
    da.mode = 0;
    da.plain = sa.plain;
    da.out = d_out;
    da.out_cap = cap;
    c->last = ccs_filter_stats();
    c->last.candidates = cand;
    c->last.dense_end = s->nd_primes;
    c->last.tierb_end = s->nd_primes + s->nb;
    const int rc = run_filter_kernels(c, s, x, sa, da, n_units, d_count, st, true, true, 0, fixed != nullptr);
    if (rc || !fixed) return rc;
    if (launch_mark_overflow(x.counters + kStatusWord, d_count, st) != hipSuccess)
        return fail(CCS_ERR_HIP, "overflow mark launch");
    return CCS_OK;
}

extern "C" int ccs_filter_launch(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, uint32_t tile_count,
                                 ccs_survivor* d_out, uint32_t cap, uint32_t* d_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    c->timed = false;
    return filter_launch_impl(c, s, ts, tile_count, d_out, cap, d_count,
                              stream ? (hipStream_t)stream : c->stream);
}

extern "C" int ccs_filter_launch_noalloc(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, uint32_t tile_count,
                                         ccs_survivor* d_out, uint32_t cap, uint32_t* d_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!c->noalloc) return fail(CCS_ERR_HIP, "no-allocation scratch unavailable");
    c->timed = false;
    const Scratch x = noalloc_scratch(c);
    return filter_launch_impl(c, s, ts, tile_count, d_out, cap, d_count,
                              stream ? (hipStream_t)stream : c->stream, &x);
}

static int ensure_buf(void** p, u64* have, u64 want) {
    if (*have >= want) return CCS_OK;
    (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIPCHK(hipMalloc(p, want));
    *have = want;
    return CCS_OK;
}

static int collect_stats(ccs_ctx* c, u32 count, const Scratch& x) {
    unsigned long long st[3] = {0, 0, 0};
    u32 midc = 0;
    HIPCHK(hipMemcpy(st, x.stats, sizeof(st), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&midc, x.counters, 4, hipMemcpyDeviceToHost));
    c->last.alive_dense = st[0];
    c->last.alive_tierb = st[2];
    c->last.survivors = count;
    float ms = 0;
    if (c->last_timed) {
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->last.ms_sieve = ms;
        if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) c->last.ms_deep = ms;
    }
    if (midc > x.mid_cap) return fail(CCS_ERR_NOMEM, "stage-B survivor buffer overflow");
    return CCS_OK;
}

extern "C" int ccs_filter(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, uint32_t tile_count,
                          ccs_survivor* h_out, uint32_t cap, uint32_t* h_count, ccs_filter_stats* stats) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!h_count) return fail(CCS_ERR_INVALID, "null count");
    rc = ensure_buf(&c->out, &c->out_bytes, (u64)(cap ? cap : 1) * sizeof(ccs_survivor) + 256);
    if (rc) return rc;
    u32* d_count = (u32*)((char*)c->out + (u64)(cap ? cap : 1) * sizeof(ccs_survivor));
    c->timed = true;
    rc = filter_launch_impl(c, s, ts, tile_count, (ccs_survivor*)c->out, cap, d_count, c->stream);
    c->timed = false;
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    u32 n = 0;
    HIPCHK(hipMemcpy(&n, d_count, 4, hipMemcpyDeviceToHost));
    const u32 keep = std::min(n, cap);
    if (keep && h_out) HIPCHK(hipMemcpy(h_out, c->out, (size_t)keep * sizeof(ccs_survivor), hipMemcpyDeviceToHost));
    *h_count = n;
    rc = collect_stats(c, n, limited(c, to_scratch(stream_scratch(c, c->stream))));
    if (stats) *stats = c->last;
    if (rc) return rc;
    return n > cap ? CCS_OVERFLOW : CCS_OK;
}

// read and clear the sticky overflow status of a scratch, once `st` has finished
static int read_status(ccs_ctx* c, const Scratch& x, hipStream_t st, ccs_launch_status* out) {
    HIPCHK(hipStreamSynchronize(st));
    u32 w[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpy(w, x.counters + kStatusWord, sizeof(w), hipMemcpyDeviceToHost));
    HIPCHK(hipMemset(x.counters + kStatusWord, 0, sizeof(w)));
    std::memset(out, 0, sizeof(*out));
    out->overflowed_launches = w[0];
    out->worst_stage_b = w[1];
    out->dropped = (u64)w[2] | ((u64)w[3] << 32);
    out->stage_b_cap = limited(c, x).mid_cap;
    return w[0] ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_stream_status(ccs_ctx* c, void* stream, ccs_launch_status* out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!out) return fail(CCS_ERR_INVALID, "null argument");
    const hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    StreamScratch* ss = stream_scratch(c, st);
    if (!ss) return CCS_ERR_NOMEM;
    return read_status(c, to_scratch(ss), st, out);
}

extern "C" int ccs_debug_limit_stage_b(ccs_ctx* c, uint32_t entries) {
    if (!c) return fail(CCS_ERR_INVALID, "null context");
    c->stage_b_limit = entries;
    return CCS_OK;
}

extern "C" int ccs_debug_deep_batch(ccs_ctx* c, uint32_t per_wave) {
    if (!c) return fail(CCS_ERR_INVALID, "null context");
    if (per_wave > kDeepBatch) return fail(CCS_ERR_INVALID, "deep batch above its maximum");
    c->deep_per_wave = per_wave;
    return CCS_OK;
}

extern "C" int ccs_last_stats(ccs_ctx* c, ccs_filter_stats* out) {
    if (!c || !out) return fail(CCS_ERR_INVALID, "null");
    *out = c->last;
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// prove
// ---------------------------------------------------------------------------

// point a prove launch at a chain-queue scratch buffer of at least cap entries
static int prove_scratch(void** buf, u64* have, u32 cap, ProveArgs& a, hipStream_t st) {
    const u64 want = ((u64)cap + 64) * 4;
    if (*have < want) HIPCHK(hipStreamSynchronize(st));   // earlier proves may still use it
    int rc = ensure_buf(buf, have, want);
    if (rc) return rc;
    a.chain_count = (u32*)*buf;
    a.chain_queue = (u32*)*buf + 16;
    return CCS_OK;
}

static int fill_prove(const ccs_shape* s, const ccs_prove_params* p, ProveArgs& a) {
    memset(&a, 0, sizeof(a));
    a.WP = s->WP;
    a.M = s->M;
    a.bases = s->d_bases;
    a.bits = p ? p->bits : 0;
    a.test_kind = p ? p->test_kind : CCS_TEST_BPSW;
    if (a.test_kind != CCS_TEST_BPSW && a.test_kind != CCS_TEST_MR) return fail(CCS_ERR_INVALID, "test_kind");
    if (a.test_kind == CCS_TEST_MR) {
        if (!p || p->n_bases < 1 || p->n_bases > 16) return fail(CCS_ERR_INVALID, "MR needs 1..16 bases");
        a.n_bases = (int)p->n_bases;
        for (u32 i = 0; i < p->n_bases; ++i) {
            if (p->bases[i] < 2) return fail(CCS_ERR_INVALID, "MR base < 2");
            a.bases_mr[i] = p->bases[i];
        }
    }
    return CCS_OK;
}

extern "C" int ccs_prove_launch(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, const ccs_survivor* d_surv,
                                const uint32_t* d_count, uint32_t cap, const ccs_prove_params* p,
                                ccs_chain_result* d_out, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || !d_surv || !d_count || !d_out) return fail(CCS_ERR_INVALID, "null argument");
    if (s->bases.empty()) return fail(CCS_ERR_INVALID, "tile survivors need a CRT shape (this one has no bases)");
    ProveArgs a;
    rc = fill_prove(s, p, a);
    if (rc) return rc;
    a.mode = 0;
    a.T0_lo = ts.lo;
    a.T0_hi = ts.hi;
    a.surv = d_surv;
    a.surv_count = d_count;
    a.surv_cap = cap;
    a.out = d_out;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    StreamScratch* ss = stream_scratch(c, st);
    if (!ss) return CCS_ERR_NOMEM;
    rc = prove_scratch(&ss->chain, &ss->chain_bytes, cap, a, st);
    if (rc) return rc;
    if (c->timed) HIPCHK(hipEventRecord(c->ev[3], st));
    prof_begin(c, 2, st);
    hipError_t e = launch_prove(a, cap, st);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("prove launch: ") + hipGetErrorString(e));
    prof_end(c, st);
    if (c->timed) HIPCHK(hipEventRecord(c->ev[4], st));
    return CCS_OK;
}

extern "C" int ccs_chain_roots_launch(ccs_ctx* c, const ccs_survivor* d_surv, const ccs_chain_result* d_res,
                                      const uint32_t* d_count, uint32_t cap, ccs_chain_root* d_out, uint32_t out_cap,
                                      uint32_t* d_root_count, uint32_t* d_prime_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!d_surv || !d_res || !d_count || !d_out || !d_root_count) return fail(CCS_ERR_INVALID, "null argument");
    RootsArgs a;
    a.surv = d_surv;
    a.res = d_res;
    a.count = d_count;
    a.cap = cap;
    a.out = d_out;
    a.out_cap = out_cap;
    a.out_count = d_root_count;
    a.primes = d_prime_count;
    const hipError_t e = launch_chain_roots(a, stream ? (hipStream_t)stream : c->stream);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("chain roots launch: ") + hipGetErrorString(e));
    return CCS_OK;
}

static int prove_host_common(ccs_ctx* c, ProveArgs& a, const void* h_surv, size_t rec_bytes, uint32_t n,
                             ccs_chain_result* h_out) {
    if (n == 0) return CCS_OK;
    const u64 in_bytes = (u64)n * rec_bytes;
    int rc = ensure_buf(&c->res, &c->res_bytes, in_bytes + 256 + (u64)n * sizeof(ccs_chain_result));
    if (rc) return rc;
    char* base = (char*)c->res;
    u32* d_count = (u32*)(base + in_bytes);
    ccs_chain_result* d_res = (ccs_chain_result*)(base + in_bytes + 256);
    HIPCHK(hipMemcpyAsync(base, h_surv, in_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_count, &n, 4, hipMemcpyHostToDevice, c->stream));
    a.surv = base;
    a.surv_count = d_count;
    a.surv_cap = n;
    a.out = d_res;
    StreamScratch* ss = stream_scratch(c, c->stream);
    if (!ss) return CCS_ERR_NOMEM;
    rc = prove_scratch(&ss->chain, &ss->chain_bytes, n, a, c->stream);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    hipError_t e = launch_prove(a, n, c->stream);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("prove launch: ") + hipGetErrorString(e));
    HIPCHK(hipEventRecord(c->ev[4], c->stream));
    HIPCHK(hipMemcpyAsync(h_out, d_res, (size_t)n * sizeof(ccs_chain_result), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) c->last.ms_prove = ms;
    return CCS_OK;
}

extern "C" int ccs_prove(ccs_ctx* c, const ccs_shape* s, ccs_u128 ts, const ccs_survivor* h_surv, uint32_t n,
                         const ccs_prove_params* p, ccs_chain_result* h_out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || (n && (!h_surv || !h_out))) return fail(CCS_ERR_INVALID, "null argument");
    for (u32 i = 0; i < n; ++i)
        if (h_surv[i].base_idx >= s->bases.size()) return fail(CCS_ERR_INVALID, "survivor base_idx out of range");
    ProveArgs a;
    rc = fill_prove(s, p, a);
    if (rc) return rc;
    a.mode = 0;
    a.T0_lo = ts.lo;
    a.T0_hi = ts.hi;
    return prove_host_common(c, a, h_surv, sizeof(ccs_survivor), n, h_out);
}

// ---------------------------------------------------------------------------
// result formatting (v13.cu:2016-2061)
// ---------------------------------------------------------------------------

static std::string hex128(u64 lo, u64 hi) {
    char buf[40];
    if (hi) snprintf(buf, sizeof(buf), "0x%llX%016llX", (unsigned long long)hi, (unsigned long long)lo);
    else snprintf(buf, sizeof(buf), "0x%llX", (unsigned long long)lo);
    return buf;
}

static int bitlen64x2(u64 lo, u64 hi) {
    if (hi) return 128 - __builtin_clzll(hi);
    return lo ? 64 - __builtin_clzll(lo) : 0;
}

extern "C" int ccs_format_results(const ccs_chain_result* res, uint32_t n, int32_t log_threshold, int32_t target,
                                  char* buf, uint64_t buf_size, uint64_t* needed) {
    if (n && !res) return fail(CCS_ERR_INVALID, "null");
    std::string out;
    for (u32 i = 0; i < n; ++i) {
        const ccs_chain_result& r = res[i];
        if (!(r.flags & CCS_F_PRIME)) continue;
        const int len = (int)r.chain_len;
        if (len < log_threshold && len < target) continue;
        if (len >= target) out += "TARGET ";
        if (r.flags & CCS_F_NONROOT) out += "NON-ROOT ";
        out += "CC" + std::to_string(len) + " " + hex128(r.root_lo, r.root_hi) + " " +
               std::to_string(bitlen64x2(r.root_lo, r.root_hi)) + "\n";
    }
    if (needed) *needed = out.size();
    if (buf && buf_size) {
        const size_t k = std::min<size_t>(out.size(), buf_size - 1);
        memcpy(buf, out.data(), k);
        buf[k] = 0;
        if (k < out.size()) return CCS_OVERFLOW;
    }
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// batched form: job sets (uploaded once), one persistent launch, per-job loop, graphs
// ---------------------------------------------------------------------------

struct ccs_jobset {
    int device = 0;
    const ccs_shape* shape = nullptr;
    u32 n_jobs = 0;
    JobRec* d_jobs = nullptr;
    u64* d_prefix = nullptr;
    std::vector<u64> prefix;
    u64 n_units = 0, candidates = 0;
    u32 unit_j = 0;                 // candidates per work unit of this set (set_unit)
    u32 cap_jobs = 0;
    JobRec* staging = nullptr;      // pinned upload staging of ccs_jobset_update
    u64* staging_prefix = nullptr;
    hipEvent_t staged = nullptr;    // the last staging upload has been consumed
    std::vector<u64> counts;        // candidates of each job (host copy)
    std::vector<JobRec> hjobs;      // the device job records (host copy)
    u64* d_each = nullptr;          // per-job {0, units} pairs of the one-job-per-launch loop
    std::vector<u32> each_unit;     // unit size of each job launched on its own
    bool each_valid = false;
    void* chain = nullptr;          // prove chain queue of launches on this set
    u64 chain_bytes = 0;
    Scratch x = {nullptr, 0, nullptr, nullptr};
    // group sets: jobs of several shapes, grouped by shape
    std::vector<const ccs_shape*> group;
    ShapeTab* d_group = nullptr;
    DeepTab* d_deep_group = nullptr;
    u64* d_shape_units = nullptr;   // [n_group + 1] first unit of each shape's jobs
    u64* staging_units = nullptr;
    std::vector<u32> job_shape;
    u32 lds_ctl = 0, nb_max = 0;
    const u32* d_phases = nullptr;  // resident phase arena of the jobs (served form), or null
    u32 ph_stride = 0;
};

extern "C" void ccs_jobset_free(ccs_jobset* js) {
    if (!js) return;
    (void)hipSetDevice(js->device);
    (void)hipFree(js->d_jobs);
    (void)hipFree(js->d_prefix);
    (void)hipFree(js->x.mid);
    (void)hipFree(js->x.counters);
    (void)hipFree(js->x.stats);
    (void)hipFree(js->chain);
    (void)hipFree(js->d_each);
    (void)hipFree(js->d_group);
    (void)hipFree(js->d_deep_group);
    (void)hipFree(js->d_shape_units);
    if (js->staging_units) (void)hipHostFree(js->staging_units);
    if (js->staging) (void)hipHostFree(js->staging);
    if (js->staging_prefix) (void)hipHostFree(js->staging_prefix);
    if (js->staged) (void)hipEventDestroy(js->staged);
    delete js;
}

// The device record of a job.  CRT shapes: n = a + W*u with the job's class a < W.
// Family shapes: every job is the class W - 1, n = W*Q - 1 = (W - 1) + W*(Q - 1), so `a`
// must be 0, the device u0 is Q0 - 1 (survivor offsets stay Q - Q0), Q starts at 1 or
// above and the last candidate stays below 2^127.
static int job_record(const ccs_shape* s, const ccs_job& j, JobRec& r) {
    if (j.count > (1ull << 32)) return fail(CCS_ERR_INVALID, "job count must be <= 2^32");
    const u128 u0 = ((u128)j.u0.hi << 64) | j.u0.lo;
    if (s->family) {
        if (j.a != 0) return fail(CCS_ERR_INVALID, "family jobs have no residue: a must be 0 (the class is W - 1)");
        const u128 W = wfull(s), a = W - 1;
        const u128 lim = ((u128)1 << 127) / W;   // W*Q - 1 < 2^127 for Q <= lim
        if (j.count && (u0 == 0 || u0 > lim || (u128)(j.count - 1) > lim - u0))
            return fail(CCS_ERR_INVALID, "family job: Q must be >= 1 and W*Q - 1 < 2^127");
        // n = (W - 1) + W*u = W*(u + 1) - 1: the device u runs from Q0 - 1
        const u128 d0 = u0 ? u0 - 1 : 0;
        r.a = (u64)a;
        r.a_hi = (u64)(a >> 64);
        r.u0_lo = (u64)d0;
        r.u0_hi = (u64)(d0 >> 64);
    } else {
        if (j.a >= s->W) return fail(CCS_ERR_INVALID, "job residue a must be < W");
        r.a = j.a;
        r.a_hi = 0;
        r.u0_lo = j.u0.lo;
        r.u0_hi = j.u0.hi;
    }
    r.count = j.count;
    r.W = s->W;
    r.W_hi = s->W_hi;
    r.shape = 0;
    r.unit_j = 0;
    r.ph_off = 0;
    r.ph_slot = kNoSlot;
    r.pad = 0;
    return CCS_OK;
}

// Work-unit size of a job set: the shape's unit_j, halved (down to kMinSetUnit) while the
// set would give a full-chip launch fewer than kUnitsPerWave units per wave.  A small set
// (a served slice) then still spreads over every wave and the launch's tail stays short;
// a large one keeps the tuned size.  Units only divide the work: any size gives the same
// survivors.
template <class Count>
static u32 set_unit(const ccs_ctx* c, const ccs_shape* s, u32 n_jobs, Count count) {
    constexpr u32 kMinSetUnit = 32768;
    constexpr u64 kUnitsPerWave = 8;
    const u64 waves = (u64)c->n_cu * s->tune.blocks_per_cu * s->tune.waves_per_block;
    u32 unit = s->tune.unit_j;
    while (unit > kMinSetUnit) {
        u64 units = 0;
        for (u32 i = 0; i < n_jobs && units < kUnitsPerWave * waves; ++i) units += (count(i) + unit - 1) / unit;
        if (units >= kUnitsPerWave * waves) break;
        const u32 half = (unit / 2) / kRowBits * kRowBits;
        if (half < kMinSetUnit) break;
        unit = half;
    }
    return unit;
}

static u32 set_unit(const ccs_ctx* c, const ccs_shape* s, const ccs_job* jobs, u32 n_jobs) {
    return set_unit(c, s, n_jobs, [&](u32 i) { return jobs[i].count; });
}

extern "C" int ccs_jobset_create(ccs_ctx* c, const ccs_shape* s, const ccs_job* jobs, uint32_t n_jobs,
                                 ccs_jobset** out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || !jobs || !out || n_jobs == 0) return fail(CCS_ERR_INVALID, "null argument or no jobs");
    *out = nullptr;
    std::vector<JobRec> jr(n_jobs);
    ccs_jobset* js = new ccs_jobset();
    js->device = c->device;
    js->shape = s;
    js->n_jobs = n_jobs;
    js->prefix.resize((size_t)n_jobs + 1);
    for (u32 i = 0; i < n_jobs; ++i) {
        if ((rc = job_record(s, jobs[i], jr[i]))) { delete js; return rc; }
    }
    js->hjobs = jr;
    js->unit_j = set_unit(c, s, jobs, n_jobs);
    for (u32 i = 0; i < n_jobs; ++i) {
        js->prefix[i] = js->n_units;
        js->n_units += (jobs[i].count + js->unit_j - 1) / js->unit_j;
        js->candidates += jobs[i].count;
        js->counts.push_back(jobs[i].count);
    }
    js->prefix[n_jobs] = js->n_units;
    js->cap_jobs = n_jobs;
    u64 want = (u64)(js->candidates * s->est_survival_b * 2.0) + (1u << 20);
    if (want > (1u << 28)) want = 1u << 28;
    hipError_t e = hipMalloc((void**)&js->d_jobs, (size_t)n_jobs * sizeof(JobRec));
    if (e == hipSuccess) e = hipMalloc((void**)&js->d_prefix, ((size_t)n_jobs + 1) * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.mid, want * sizeof(MidRec));
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.counters, 64);
    if (e == hipSuccess) e = hipMemset(js->x.counters, 0, 64);
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.stats, 64);
    if (e == hipSuccess) e = hipMemcpy(js->d_jobs, jr.data(), jr.size() * sizeof(JobRec), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(js->d_prefix, js->prefix.data(), js->prefix.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ccs_jobset_free(js);
        return fail(CCS_ERR_HIP, std::string("jobset upload: ") + hipGetErrorString(e));
    }
    js->x.mid_cap = (u32)want;
    *out = js;
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// group sets: jobs of several shapes in one launch (the sieve kernel loads each shape's
// tables as its work reaches that shape's jobs)
// ---------------------------------------------------------------------------

// Work-unit sizes of a group set.  The shapes of a group can differ several-fold in sieve
// time per candidate (the dense survivors go one by one through stage B; depth 12 leaves
// ~30x more of them than depth 20), so each shape gets units of about equal cost:
// cost per candidate ~ 1 + 900 * (dense-stage survival), the ratio of the per-survivor to
// the per-candidate time measured on the 31# depth profiles.  The cheapest shape takes the
// size set_unit picks for the whole set's weighted work, the others proportionally
// smaller units (powers of two from 2^14).
static double sieve_weight(const ccs_shape* s) { return 1.0 + 900.0 * s->est_survival_d; }

static void group_units_per_job(const ccs_ctx* c, ccs_jobset* js, const ccs_job* jobs, u32 n_jobs, JobRec* rec) {
    double wmin = 1e300;
    for (const ccs_shape* g : js->group) wmin = std::min(wmin, sieve_weight(g));
    const u32 base = set_unit(c, js->shape, n_jobs, [&](u32 i) {
        return (u64)((double)jobs[i].count * sieve_weight(js->group[js->job_shape[i]]) / wmin);
    });
    js->unit_j = base;
    std::vector<u32> per(js->group.size());
    for (size_t k = 0; k < per.size(); ++k) {
        const double want = (double)base * wmin / sieve_weight(js->group[k]);
        u32 u = 1u << 14;
        while ((double)u * 2 <= want && u < base && u < (1u << 20)) u *= 2;
        per[k] = u;
    }
    for (u32 i = 0; i < n_jobs; ++i) rec[i].unit_j = per[js->job_shape[i]];
}

// the first unit of each shape's jobs (jobs are grouped by shape)
static void group_units(const ccs_jobset* js, std::vector<u64>& units) {
    const u32 ng = (u32)js->group.size();
    units.assign((size_t)ng + 1, js->n_units);
    u32 k = 0;
    for (u32 i = 0; i < js->n_jobs; ++i)
        while (k <= js->job_shape[i]) units[k++] = js->prefix[i];
    while (k <= ng) units[k++] = js->n_units;
}

static int group_check(const ccs_shape* const* shapes, uint32_t n_shapes, const uint32_t* job_shape,
                       uint32_t n_jobs) {
    if (!shapes || !n_shapes || !job_shape) return fail(CCS_ERR_INVALID, "null argument");
    for (u32 k = 0; k < n_shapes; ++k) {
        if (!shapes[k]) return fail(CCS_ERR_INVALID, "null shape");
        if (shapes[k]->device != shapes[0]->device) return fail(CCS_ERR_INVALID, "shapes on different devices");
        if (shapes[k]->tune.rows_per_group != shapes[0]->tune.rows_per_group)
            return fail(CCS_ERR_INVALID, "the shapes of a group need the same rows_per_group");
    }
    for (u32 i = 0; i < n_jobs; ++i) {
        if (job_shape[i] >= n_shapes) return fail(CCS_ERR_INVALID, "job_shape out of range");
        if (i && job_shape[i] < job_shape[i - 1]) return fail(CCS_ERR_INVALID, "jobs must be grouped by shape (job_shape non-decreasing)");
    }
    return CCS_OK;
}

extern "C" int ccs_groupset_create(ccs_ctx* c, const ccs_shape* const* shapes, uint32_t n_shapes,
                                   const ccs_job* jobs, const uint32_t* job_shape, uint32_t n_jobs,
                                   ccs_jobset** out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!jobs || !out || !n_jobs) return fail(CCS_ERR_INVALID, "null argument or no jobs");
    *out = nullptr;
    if ((rc = group_check(shapes, n_shapes, job_shape, n_jobs))) return rc;
    std::vector<JobRec> jr(n_jobs);
    for (u32 i = 0; i < n_jobs; ++i) {
        if ((rc = job_record(shapes[job_shape[i]], jobs[i], jr[i]))) return rc;
        jr[i].shape = job_shape[i];
    }
    ccs_jobset* js = new ccs_jobset();
    js->device = c->device;
    js->shape = shapes[0];
    js->n_jobs = n_jobs;
    js->group.assign(shapes, shapes + n_shapes);
    js->job_shape.assign(job_shape, job_shape + n_jobs);
    double surv = 0;
    for (u32 k = 0; k < n_shapes; ++k) {
        js->lds_ctl = std::max(js->lds_ctl, (shapes[k]->lds_words + 3) & ~3u);
        js->nb_max = std::max(js->nb_max, shapes[k]->nb);
    }
    group_units_per_job(c, js, jobs, n_jobs, jr.data());
    js->hjobs = jr;
    js->prefix.resize((size_t)n_jobs + 1);
    for (u32 i = 0; i < n_jobs; ++i) {
        js->prefix[i] = js->n_units;
        js->n_units += (jobs[i].count + jr[i].unit_j - 1) / jr[i].unit_j;
        js->candidates += jobs[i].count;
        js->counts.push_back(jobs[i].count);
        surv += (double)jobs[i].count * shapes[job_shape[i]]->est_survival_b;
    }
    js->prefix[n_jobs] = js->n_units;
    js->cap_jobs = n_jobs;
    std::vector<ShapeTab> tabs(n_shapes);
    std::vector<DeepTab> dtabs(n_shapes);
    for (u32 k = 0; k < n_shapes; ++k) {
        tabs[k] = shape_tab(shapes[k]);
        dtabs[k] = deep_tab(shapes[k]);
    }
    std::vector<u64> units;
    group_units(js, units);
    u64 want = (u64)(surv * 2.0) + (1u << 20);
    if (want > (1u << 28)) want = 1u << 28;
    hipError_t e = hipMalloc((void**)&js->d_jobs, (size_t)n_jobs * sizeof(JobRec));
    if (e == hipSuccess) e = hipMalloc((void**)&js->d_prefix, ((size_t)n_jobs + 1) * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&js->d_group, n_shapes * sizeof(ShapeTab));
    if (e == hipSuccess) e = hipMalloc((void**)&js->d_deep_group, n_shapes * sizeof(DeepTab));
    if (e == hipSuccess) e = hipMalloc((void**)&js->d_shape_units, ((size_t)n_shapes + 1) * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.mid, want * sizeof(MidRec));
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.counters, 64);
    if (e == hipSuccess) e = hipMemset(js->x.counters, 0, 64);
    if (e == hipSuccess) e = hipMalloc((void**)&js->x.stats, 64);
    if (e == hipSuccess) e = hipMemcpy(js->d_jobs, jr.data(), jr.size() * sizeof(JobRec), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(js->d_prefix, js->prefix.data(), js->prefix.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(js->d_group, tabs.data(), tabs.size() * sizeof(ShapeTab), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(js->d_deep_group, dtabs.data(), dtabs.size() * sizeof(DeepTab), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(js->d_shape_units, units.data(), units.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ccs_jobset_free(js);
        return fail(CCS_ERR_HIP, std::string("group set upload: ") + hipGetErrorString(e));
    }
    js->x.mid_cap = (u32)want;
    *out = js;
    return CCS_OK;
}

// new jobs (at most as many as the set was created with) for the same shapes; asynchronous
// on `st`.  With slots, job i reads its phases from slot slots[i] of the arena d_phases,
// where they are the phases at u0_i - offs[i].
static int groupset_update_impl(ccs_ctx* c, ccs_jobset* js, const ccs_job* h_jobs, const uint32_t* job_shape,
                                const uint32_t* slots, const uint64_t* offs, uint32_t n_jobs, const u32* d_phases,
                                u32 stride, hipStream_t st) {
    int rc;
    if (!js || !h_jobs || !n_jobs) return fail(CCS_ERR_INVALID, "null argument or no jobs");
    if (js->group.empty()) return fail(CCS_ERR_INVALID, "not a group set (use ccs_jobset_update)");
    if (n_jobs > js->cap_jobs) return fail(CCS_ERR_INVALID, "more jobs than the set was created with");
    if ((rc = group_check(js->group.data(), (u32)js->group.size(), job_shape, n_jobs))) return rc;
    if (!js->staging) {
        HIPCHK(hipHostMalloc((void**)&js->staging, (size_t)js->cap_jobs * sizeof(JobRec)));
        HIPCHK(hipHostMalloc((void**)&js->staging_prefix, ((size_t)js->cap_jobs + 1) * 8));
        HIPCHK(hipHostMalloc((void**)&js->staging_units, (js->group.size() + 1) * 8));
        HIPCHK(hipEventCreateWithFlags(&js->staged, hipEventDisableTiming));
    } else {
        HIPCHK(hipEventSynchronize(js->staged));
    }
    double surv = 0;
    for (u32 i = 0; i < n_jobs; ++i) {
        JobRec tmp;
        if ((rc = job_record(js->group[job_shape[i]], h_jobs[i], tmp))) return rc;
        surv += (double)h_jobs[i].count * js->group[job_shape[i]]->est_survival_b;
    }
    if ((u64)(surv * 2.0) + (1u << 16) > js->x.mid_cap)
        return fail(CCS_ERR_INVALID, "jobs exceed the stage-B buffer of the set");
    js->prefix.resize((size_t)n_jobs + 1);
    js->counts.resize(n_jobs);
    js->hjobs.resize(n_jobs);
    js->job_shape.assign(job_shape, job_shape + n_jobs);
    js->each_valid = false;
    for (u32 i = 0; i < n_jobs; ++i) {
        (void)job_record(js->group[job_shape[i]], h_jobs[i], js->staging[i]);
        js->staging[i].shape = job_shape[i];
        if (slots) {
            js->staging[i].ph_slot = slots[i];
            js->staging[i].ph_off = offs[i];
        }
    }
    js->d_phases = slots ? d_phases : nullptr;
    js->ph_stride = slots ? stride : 0;
    group_units_per_job(c, js, h_jobs, n_jobs, js->staging);
    u64 units = 0, cand = 0;
    for (u32 i = 0; i < n_jobs; ++i) {
        js->counts[i] = h_jobs[i].count;
        js->hjobs[i] = js->staging[i];
        js->prefix[i] = js->staging_prefix[i] = units;
        units += (h_jobs[i].count + js->staging[i].unit_j - 1) / js->staging[i].unit_j;
        cand += h_jobs[i].count;
    }
    js->prefix[n_jobs] = js->staging_prefix[n_jobs] = units;
    js->n_units = units;
    js->candidates = cand;
    js->n_jobs = n_jobs;
    std::vector<u64> su;
    group_units(js, su);
    std::memcpy(js->staging_units, su.data(), su.size() * 8);
    HIPCHK(hipMemcpyAsync(js->d_jobs, js->staging, (size_t)n_jobs * sizeof(JobRec), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(js->d_prefix, js->staging_prefix, ((size_t)n_jobs + 1) * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(js->d_shape_units, js->staging_units, su.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(js->staged, st));
    return CCS_OK;
}

extern "C" int ccs_groupset_update(ccs_ctx* c, ccs_jobset* js, const ccs_job* h_jobs, const uint32_t* job_shape,
                                   uint32_t n_jobs, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    return groupset_update_impl(c, js, h_jobs, job_shape, nullptr, nullptr, n_jobs, nullptr, 0,
                                stream ? (hipStream_t)stream : c->stream);
}

namespace ccs {
u32 shape_phase_words(const ccs_shape* s) { return s->nd + s->nb; }
double shape_final_survival(const ccs_shape* s) { return s->est_survival_f; }
const u32* jobset_status_words(const ccs_jobset* js) { return js->x.counters + kStatusWord; }
const PrimeMod* shape_pmod(const ccs_shape* s) { return s->d_pmod; }
int groupset_update_resident(ccs_ctx* c, ccs_jobset* js, const ccs_job* jobs, const uint32_t* job_shape,
                             const uint32_t* slots, const uint64_t* offs, uint32_t n_jobs, const u32* d_phases,
                             u32 stride, hipStream_t st) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!slots || !offs || !d_phases) return fail(CCS_ERR_INVALID, "null argument");
    return groupset_update_impl(c, js, jobs, job_shape, slots, offs, n_jobs, d_phases, stride, st);
}
}  // namespace ccs

// replace the jobs of a set (n_jobs <= the count it was created with; same shape) with an
// asynchronous upload on `stream`; h_jobs must stay valid until the stream reaches it
extern "C" int ccs_jobset_update(ccs_ctx* c, ccs_jobset* js, const ccs_job* h_jobs, uint32_t n_jobs,
                                 void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !h_jobs || n_jobs == 0) return fail(CCS_ERR_INVALID, "null argument or no jobs");
    if (n_jobs > js->cap_jobs) return fail(CCS_ERR_INVALID, "more jobs than the set was created with");
    if (!js->group.empty()) return fail(CCS_ERR_INVALID, "a group set takes ccs_groupset_update");
    const ccs_shape* s = js->shape;
    if (!js->staging) {
        HIPCHK(hipHostMalloc((void**)&js->staging, (size_t)js->cap_jobs * sizeof(JobRec)));
        HIPCHK(hipHostMalloc((void**)&js->staging_prefix, ((size_t)js->cap_jobs + 1) * 8));
        HIPCHK(hipEventCreateWithFlags(&js->staged, hipEventDisableTiming));
    } else {
        HIPCHK(hipEventSynchronize(js->staged));   // the previous upload has left the staging
    }
    u64 units = 0, cand = 0;
    for (u32 i = 0; i < n_jobs; ++i) {
        JobRec tmp;
        if ((rc = job_record(s, h_jobs[i], tmp))) return rc;
        cand += h_jobs[i].count;
    }
    if ((u64)(cand * s->est_survival_b * 2.0) + (1u << 16) > js->x.mid_cap)
        return fail(CCS_ERR_INVALID, "jobs exceed the stage-B buffer of the set");
    js->prefix.resize((size_t)n_jobs + 1);
    js->unit_j = set_unit(c, s, h_jobs, n_jobs);
    js->d_phases = nullptr;
    js->ph_stride = 0;
    js->counts.resize(n_jobs);
    js->hjobs.resize(n_jobs);
    js->each_valid = false;
    for (u32 i = 0; i < n_jobs; ++i) {
        js->counts[i] = h_jobs[i].count;
        (void)job_record(s, h_jobs[i], js->staging[i]);
        js->hjobs[i] = js->staging[i];
        js->prefix[i] = js->staging_prefix[i] = units;
        units += (h_jobs[i].count + js->unit_j - 1) / js->unit_j;
    }
    js->prefix[n_jobs] = js->staging_prefix[n_jobs] = units;
    js->n_units = units;
    js->candidates = cand;
    js->n_jobs = n_jobs;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    HIPCHK(hipMemcpyAsync(js->d_jobs, js->staging, (size_t)n_jobs * sizeof(JobRec), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(js->d_prefix, js->staging_prefix, ((size_t)n_jobs + 1) * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(js->staged, st));
    return CCS_OK;
}

static void jobset_args(const ccs_ctx* c, const ccs_jobset* js, SieveArgs& sa, DeepArgs& da, void* d_out, u32 cap) {
    const ccs_shape* s = js->shape;
    fill_sieve_common(s, limited(c, js->x), sa);
    sa.unit_j = js->unit_j;
    sa.phases = js->group.empty() ? nullptr : js->d_phases;   // (resident jobs are group-set jobs)
    sa.ph_stride = js->ph_stride;
    sa.mode = 1;
    sa.jobsrc.jobs = js->d_jobs;
    sa.jobsrc.unit_prefix = js->d_prefix;
    sa.jobsrc.n_jobs = js->n_jobs;
    memset(&da, 0, sizeof(da));
    da.W = s->W;
    da.W_hi = s->W_hi;
    da.deep = s->d_deep;
    da.hash = s->d_hash;
    da.nc = s->nc;
    da.masked = s->deep_masked;
    da.mode = 1;
    if (!js->group.empty()) {
        sa.unit_j = 0;                 // each job's own unit size
        sa.group = js->d_group;
        sa.n_group = (u32)js->group.size();
        sa.shape_units = js->d_shape_units;
        sa.lds_ctl = js->lds_ctl;
        da.group = js->d_deep_group;
    }
    da.jobsrc = sa.jobsrc;
    da.out = d_out;
    da.out_cap = cap;
}

static void jobset_stats_begin(ccs_ctx* c, const ccs_jobset* js) {
    c->last = ccs_filter_stats();
    c->last.candidates = js->candidates;
    c->last.dense_end = js->shape->nd_primes;
    c->last.tierb_end = js->shape->nd_primes + js->shape->nb;
}

// one persistent, work-stealing launch over every job of the set
extern "C" int ccs_jobset_launch(ccs_ctx* c, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                                 uint32_t* d_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !d_out || !d_count) return fail(CCS_ERR_INVALID, "null argument");
    SieveArgs sa;
    DeepArgs da;
    jobset_args(c, js, sa, da, d_out, cap);
    jobset_stats_begin(c, js);
    return run_filter_kernels(c, js->shape, limited(c, js->x), sa, da, js->n_units, d_count,
                              stream ? (hipStream_t)stream : c->stream, true, true, js->nb_max);
}

// the plain per-shape loop: one launch per job (the v16 dispatch pattern, v16...cu:5611-5965),
// appending into the same output
extern "C" int ccs_jobset_launch_each(ccs_ctx* c, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                                      uint32_t* d_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !d_out || !d_count) return fail(CCS_ERR_INVALID, "null argument");
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    // each job launched as a set of its own would be: its own unit size and {0, units} prefix
    if (!js->each_valid) {
        if (!js->d_each) HIPCHK(hipMalloc((void**)&js->d_each, (size_t)js->cap_jobs * 16));
        std::vector<u64> pairs(2 * (size_t)js->n_jobs);
        js->each_unit.resize(js->n_jobs);
        for (u32 i = 0; i < js->n_jobs; ++i) {
            const u64 n = js->counts[i];
            js->each_unit[i] = set_unit(c, js->shape, 1, [&](u32) { return n; });
            pairs[2 * i] = 0;
            pairs[2 * i + 1] = (n + js->each_unit[i] - 1) / js->each_unit[i];
        }
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpy(js->d_each, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice));
        js->each_valid = true;
    }
    jobset_stats_begin(c, js);
    if (launch_zero3(d_count, 1, (u32*)js->x.stats, 16, nullptr, 0, st) != hipSuccess)
        return fail(CCS_ERR_HIP, "zero");
    for (u32 i = 0; i < js->n_jobs; ++i) {
        if (!js->counts[i]) continue;
        const u64 units = (js->counts[i] + js->each_unit[i] - 1) / js->each_unit[i];
        SieveArgs sa;
        DeepArgs da;
        jobset_args(c, js, sa, da, d_out, cap);   // the deep kernel keeps the whole job table
        sa.unit_j = js->each_unit[i];
        sa.jobsrc.jobs = js->d_jobs + i;
        sa.jobsrc.unit_prefix = js->d_each + 2 * (size_t)i;
        sa.jobsrc.n_jobs = 1;
        sa.jobsrc.job_base = i;
        if (!js->group.empty()) {   // one shape per launch: the job's own
            const ccs_shape* sh = js->group[js->job_shape[i]];
            sa.group = nullptr;
            sa.n_group = 1;
            sa.tab = shape_tab(sh);
        }
        rc = run_filter_kernels(c, js->shape, limited(c, js->x), sa, da, units, d_count, st, false, false, js->nb_max);
        if (rc) return rc;
    }
    return CCS_OK;
}

extern "C" int ccs_jobset_status(ccs_ctx* c, ccs_jobset* js, void* stream, ccs_launch_status* out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !out) return fail(CCS_ERR_INVALID, "null argument");
    return read_status(c, js->x, stream ? (hipStream_t)stream : c->stream, out);
}

extern "C" int ccs_jobset_stats(ccs_ctx* c, ccs_jobset* js, uint32_t* d_count, ccs_filter_stats* out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !out) return fail(CCS_ERR_INVALID, "null");
    HIPCHK(hipStreamSynchronize(c->stream));
    u32 n = 0;
    if (d_count) HIPCHK(hipMemcpy(&n, d_count, 4, hipMemcpyDeviceToHost));
    jobset_stats_begin(c, js);
    rc = collect_stats(c, n, limited(c, js->x));
    *out = c->last;
    return rc;
}

struct ccs_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" void ccs_graph_free(ccs_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

// capture the persistent batched launch (counter fills, sieve, deep) as a HIP graph
extern "C" int ccs_graph_create(ccs_ctx* c, ccs_jobset* js, ccs_job_survivor* d_out, uint32_t cap,
                                uint32_t* d_count, ccs_graph** out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !d_out || !d_count || !out) return fail(CCS_ERR_INVALID, "null argument");
    *out = nullptr;
    ccs_graph* g = new ccs_graph();
    g->device = c->device;
    const bool prof = c->profiling, timed = c->timed;
    c->profiling = false;
    c->timed = false;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        c->profiling = prof;
        c->timed = timed;
        delete g;
        return fail(CCS_ERR_HIP, std::string("begin capture: ") + hipGetErrorString(e));
    }
    SieveArgs sa;
    DeepArgs da;
    jobset_args(c, js, sa, da, d_out, cap);
    rc = run_filter_kernels(c, js->shape, limited(c, js->x), sa, da, js->n_units, d_count, c->stream, true, true,
                            js->nb_max);
    e = hipStreamEndCapture(c->stream, &g->graph);
    c->profiling = prof;
    c->timed = timed;
    if (rc || e != hipSuccess) {
        ccs_graph_free(g);
        return rc ? rc : fail(CCS_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        ccs_graph_free(g);
        return fail(CCS_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
    }
    *out = g;
    return CCS_OK;
}

extern "C" int ccs_graph_launch(ccs_ctx* c, ccs_graph* g, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!g) return fail(CCS_ERR_INVALID, "null graph");
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    prof_begin(c, 0, st);
    HIPCHK(hipGraphLaunch(g->exec, st));
    prof_end(c, st);
    return CCS_OK;
}

// convenience forms: upload, launch, synchronise
extern "C" int ccs_batch_launch(ccs_ctx* c, const ccs_shape* s, const ccs_job* jobs, uint32_t n_jobs,
                                ccs_job_survivor* d_out, uint32_t cap, uint32_t* d_count, void* stream) {
    ccs_jobset* js = nullptr;
    int rc = ccs_jobset_create(c, s, jobs, n_jobs, &js);
    if (rc) return rc;
    rc = ccs_jobset_launch(c, js, d_out, cap, d_count, stream);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(CCS_ERR_HIP, "synchronize");
    ccs_jobset_free(js);
    return rc;
}

extern "C" int ccs_batch(ccs_ctx* c, const ccs_shape* s, const ccs_job* jobs, uint32_t n_jobs,
                         ccs_job_survivor* h_out, uint32_t cap, uint32_t* h_count, ccs_filter_stats* stats) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!h_count) return fail(CCS_ERR_INVALID, "null count");
    rc = ensure_buf(&c->out, &c->out_bytes, (u64)(cap ? cap : 1) * sizeof(ccs_job_survivor) + 256);
    if (rc) return rc;
    u32* d_count = (u32*)((char*)c->out + (u64)(cap ? cap : 1) * sizeof(ccs_job_survivor));
    ccs_jobset* js = nullptr;
    rc = ccs_jobset_create(c, s, jobs, n_jobs, &js);
    if (rc) return rc;
    c->timed = true;
    rc = ccs_jobset_launch(c, js, (ccs_job_survivor*)c->out, cap, d_count, c->stream);
    c->timed = false;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(CCS_ERR_HIP, "synchronize");
    u32 n = 0;
    if (!rc && hipMemcpy(&n, d_count, 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(CCS_ERR_HIP, "count");
    const u32 keep = std::min(n, cap);
    if (!rc && keep && h_out &&
        hipMemcpy(h_out, c->out, (size_t)keep * sizeof(ccs_job_survivor), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(CCS_ERR_HIP, "copy");
    *h_count = n;
    if (!rc) rc = collect_stats(c, n, limited(c, js->x));
    if (stats) *stats = c->last;
    ccs_jobset_free(js);
    if (rc) return rc;
    return n > cap ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_prove_jobs(ccs_ctx* c, const ccs_shape* s, const ccs_job* jobs, uint32_t n_jobs,
                              const ccs_job_survivor* h_surv, uint32_t n, const ccs_prove_params* p,
                              ccs_chain_result* h_out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || !jobs || (n && (!h_surv || !h_out))) return fail(CCS_ERR_INVALID, "null argument");
    for (u32 i = 0; i < n; ++i)
        if (h_surv[i].job_idx >= n_jobs) return fail(CCS_ERR_INVALID, "survivor job_idx out of range");
    ccs_jobset* js = nullptr;
    rc = ccs_jobset_create(c, s, jobs, n_jobs, &js);
    if (rc) return rc;
    ProveArgs a;
    rc = fill_prove(s, p, a);
    if (!rc) {
        a.mode = 1;
        a.jobs = js->d_jobs;
        rc = prove_host_common(c, a, h_surv, sizeof(ccs_job_survivor), n, h_out);
    }
    ccs_jobset_free(js);
    return rc;
}

extern "C" int ccs_prove_jobset_launch(ccs_ctx* c, ccs_jobset* js, const ccs_job_survivor* d_surv,
                                       const uint32_t* d_count, uint32_t cap, const ccs_prove_params* p,
                                       ccs_chain_result* d_out, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !d_surv || !d_count || !d_out) return fail(CCS_ERR_INVALID, "null argument");
    ProveArgs a;
    rc = fill_prove(js->shape, p, a);
    if (rc) return rc;
    a.mode = 1;
    a.jobs = js->d_jobs;
    a.surv = d_surv;
    a.surv_count = d_count;
    a.surv_cap = cap;
    a.out = d_out;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    rc = prove_scratch(&js->chain, &js->chain_bytes, cap, a, st);
    if (rc) return rc;
    prof_begin(c, 2, st);
    hipError_t e = launch_prove(a, cap, st);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("prove launch: ") + hipGetErrorString(e));
    prof_end(c, st);
    return CCS_OK;
}

extern "C" int ccs_family_top_gate_launch(ccs_ctx* c, ccs_jobset* js, const ccs_job_survivor* d_in,
                                          const uint32_t* d_in_count, uint32_t in_cap, int32_t top_pos,
                                          int32_t reps, ccs_job_survivor* d_out, uint32_t out_cap,
                                          uint32_t* d_out_count, void* stream) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!js || !d_in || !d_in_count || !d_out || !d_out_count) return fail(CCS_ERR_INVALID, "null argument");
    const ccs_shape* s = js->shape;
    if (!s->family) return fail(CCS_ERR_INVALID, "the top gate needs a family shape");
    for (const ccs_shape* g : js->group)
        if (!g->family) return fail(CCS_ERR_INVALID, "the top gate needs family shapes");
    if (top_pos < 0 || top_pos > 63) return fail(CCS_ERR_INVALID, "top_pos must be in 0..63");
    if (reps != 2 && reps != 3) return fail(CCS_ERR_INVALID, "reps must be 2 ({2,3}) or 3 ({2,3,5})");
    // every job's p_top = (W*Q) << top_pos - 1 must be below 2^127 (Montgomery range)
    for (u32 i = 0; i < js->n_jobs; ++i) {
        const JobRec& r = js->hjobs[i];
        if (!r.count) continue;
        const u128 lim = (((u128)1 << 127) >> top_pos) / (((u128)r.W_hi << 64) | r.W);
        const u128 qlast = (((u128)r.u0_hi << 64) | r.u0_lo) + r.count;   // device u0 = Q0 - 1
        if (qlast > lim) return fail(CCS_ERR_INVALID, "p_top of a job reaches 2^127");
    }
    TopGateArgs a;
    a.jobs = js->d_jobs;
    a.in = d_in;
    a.in_count = d_in_count;
    a.in_cap = in_cap;
    a.top_pos = top_pos;
    a.reps = reps;
    a.out = d_out;
    a.out_cap = out_cap;
    a.out_count = d_out_count;
    hipError_t e = launch_top_gate(a, stream ? (hipStream_t)stream : c->stream);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, std::string("top gate launch: ") + hipGetErrorString(e));
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------

extern "C" int ccs_profile_start(ccs_ctx* c) {
    int rc = ctx_check(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->profiling = true;
    c->ev_used = 0;
    c->marks.clear();
    return CCS_OK;
}

extern "C" int ccs_profile_stop(ccs_ctx* c, ccs_kernel_times* out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!out) return fail(CCS_ERR_INVALID, "null");
    HIPCHK(hipDeviceSynchronize());
    memset(out, 0, sizeof(*out));
    for (auto& m : c->marks) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        if (m.first == 0) { out->n_sieve++; out->ms_sieve += ms; }
        else if (m.first == 1) { out->n_deep++; out->ms_deep += ms; }
        else { out->n_prove++; out->ms_prove += ms; }
    }
    c->profiling = false;
    c->marks.clear();
    c->ev_used = 0;
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// autotuner: coordinate descent over the kernel knobs, each trial timed with HIP events
// on a sample tile range of the shape (filter launch only: sieve + deep)
// ---------------------------------------------------------------------------

static int time_filter(ccs_ctx* c, ccs_shape* s, ccs_u128 ts, uint32_t tile_count, int reps, float* ms_out) {
    int rc = ensure_buf(&c->out, &c->out_bytes, (u64)(1u << 20) * sizeof(ccs_survivor) + 256);
    if (rc) return rc;
    u32* d_count = (u32*)((char*)c->out + (u64)(1u << 20) * sizeof(ccs_survivor));
    float best = 1e30f;
    for (int r = 0; r < reps + 1; ++r) {
        c->timed = true;
        rc = filter_launch_impl(c, s, ts, tile_count, (ccs_survivor*)c->out, 1u << 20, d_count, c->stream);
        c->timed = false;
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->stream));
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIPCHK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        if (r > 0 && a + b < best) best = a + b;   // first run warms up
    }
    *ms_out = best;
    return CCS_OK;
}

// time the batched launch over a sample of jobs (a job set is rebuilt for every tuning,
// since its unit size follows unit_j and the launch shape)
static int time_jobs(ccs_ctx* c, ccs_shape* s, const ccs_job* jobs, u32 n_jobs, int reps, float* ms_out) {
    int rc = ensure_buf(&c->out, &c->out_bytes, (u64)(1u << 20) * sizeof(ccs_job_survivor) + 256);
    if (rc) return rc;
    u32* d_count = (u32*)((char*)c->out + (u64)(1u << 20) * sizeof(ccs_job_survivor));
    ccs_jobset* js = nullptr;
    rc = ccs_jobset_create(c, s, jobs, n_jobs, &js);
    if (rc) return rc;
    float best = 1e30f;
    for (int r = 0; r < reps + 1 && !rc; ++r) {
        c->timed = true;
        rc = ccs_jobset_launch(c, js, (ccs_job_survivor*)c->out, 1u << 20, d_count, c->stream);
        c->timed = false;
        if (rc) break;
        float a = 0, b = 0;
        if (hipStreamSynchronize(c->stream) != hipSuccess || hipEventElapsedTime(&a, c->ev[0], c->ev[1]) != hipSuccess ||
            hipEventElapsedTime(&b, c->ev[1], c->ev[2]) != hipSuccess) {
            rc = fail(CCS_ERR_HIP, "timing");
            break;
        }
        if (r > 0 && a + b < best) best = a + b;
    }
    ccs_jobset_free(js);
    *ms_out = best;
    return rc;
}

static int autotune_impl(ccs_ctx* c, ccs_shape* s, const std::function<int(float*)>& timer, uint32_t max_trials,
                         ccs_tuning* best_out, float* best_ms_out);

extern "C" int ccs_autotune(ccs_ctx* c, ccs_shape* s, ccs_u128 ts, uint32_t tile_count, uint32_t max_trials,
                            ccs_tuning* best_out, float* best_ms_out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || s->classes.empty()) return fail(CCS_ERR_INVALID, "autotune needs a shape with classes");
    return autotune_impl(c, s, [&](float* ms) { return time_filter(c, s, ts, tile_count, 2, ms); }, max_trials,
                         best_out, best_ms_out);
}

extern "C" int ccs_autotune_jobs(ccs_ctx* c, ccs_shape* s, const ccs_job* h_jobs, uint32_t n_jobs,
                                 uint32_t max_trials, ccs_tuning* best_out, float* best_ms_out) {
    int rc = ctx_check(c);
    if (rc) return rc;
    if (!s || !h_jobs || !n_jobs) return fail(CCS_ERR_INVALID, "autotune needs a shape and sample jobs");
    return autotune_impl(c, s, [&](float* ms) { return time_jobs(c, s, h_jobs, n_jobs, 2, ms); }, max_trials,
                         best_out, best_ms_out);
}

static int autotune_impl(ccs_ctx* c, ccs_shape* s, const std::function<int(float*)>& timer, uint32_t max_trials,
                         ccs_tuning* best_out, float* best_ms_out) {
    int rc;
    ccs_tuning best = s->tune;
    float best_ms = 0;
    rc = timer(&best_ms);
    if (rc) return rc;
    u32 trials = 1;
    // candidate values per knob
    std::vector<u32> dense_cut = {48, 64, 80, 100, 128, 160, 200, 256, 320, 400, 512, 640};
    auto count_below = [&](u32 lim) {
        u32 n = 0;
        while (n < s->fp.size() && n < 2u * kMaxDense && s->fp[n] < lim) ++n;
        return n;
    };
    // time one tuning; n_tierb is kept (clamped to the primes below 16384 and to LDS)
    auto try_tuning_nb = [&](ccs_tuning t) -> int {
        if (trials >= max_trials) return CCS_OK;
        u32 avail = 0;
        while (t.n_dense + avail < s->fp.size() && avail < (u32)kMaxTierB && s->fp[t.n_dense + avail] < kTierBMaxQ)
            ++avail;
        if (t.n_tierb > avail) t.n_tierb = avail;
        for (;;) {
            ccs_tuning tt = t;
            if (validate_tuning(s, &tt) == CCS_OK) break;
            if (t.n_tierb == 0) return CCS_OK;
            t.n_tierb = t.n_tierb > 16 ? t.n_tierb - 16 : 0;
        }
        s->tune = t;
        int r2 = upload_tables(s);
        if (r2) return r2;
        float ms = 0;
        r2 = timer(&ms);
        if (r2) return r2;
        ++trials;
        if (ms < best_ms) { best_ms = ms; best = t; }
        return CCS_OK;
    };
    auto try_tuning = [&](ccs_tuning t) -> int {
        if (trials >= max_trials) return CCS_OK;
        // stage B keeps its size but follows the dense cutoff
        const u32 keep = t.n_tierb;
        t.n_tierb = 0;
        while (t.n_dense + t.n_tierb < s->fp.size() && t.n_tierb < keep && s->fp[t.n_dense + t.n_tierb] < kTierBMaxQ)
            ++t.n_tierb;
        for (;;) {
            ccs_tuning tt = t;
            if (validate_tuning(s, &tt) == CCS_OK) break;
            if (t.n_tierb == 0) return CCS_OK;   // does not fit in LDS: skip
            t.n_tierb = t.n_tierb > 16 ? t.n_tierb - 16 : 0;
        }
        s->tune = t;
        int r2 = upload_tables(s);
        if (r2) return r2;
        float ms = 0;
        r2 = timer(&ms);
        if (r2) return r2;
        ++trials;
        if (ms < best_ms) { best_ms = ms; best = t; }
        return CCS_OK;
    };
    for (int pass = 0; pass < 2; ++pass) {
        ccs_tuning base = best;
        for (u32 cut : dense_cut) {
            ccs_tuning t = best;
            t.n_dense = count_below(cut);
            if (t.n_dense == base.n_dense) continue;
            if ((rc = try_tuning(t))) return rc;
        }
        for (u32 pm : {0u, 4096u, 16384u, 65536u, 262144u}) {
            ccs_tuning t = best;
            if (pm == t.pair_max) continue;
            t.pair_max = pm;
            if ((rc = try_tuning(t))) return rc;
        }
        for (u32 rg : {1u, 2u, 4u}) {
            ccs_tuning t = best;
            if (rg == t.rows_per_group) continue;
            t.rows_per_group = rg;
            if ((rc = try_tuning(t))) return rc;
        }
        for (u32 uj : {32768u, 65536u, 131072u, 262144u, 524288u, 1048576u}) {
            ccs_tuning t = best;
            if (uj == t.unit_j) continue;
            t.unit_j = uj;
            if ((rc = try_tuning(t))) return rc;
        }
        for (u32 nb : {64u, 128u, 256u, 384u, 512u, 768u, 1024u}) {
            ccs_tuning t = best;
            t.n_tierb = nb;
            if ((rc = try_tuning_nb(t))) return rc;
        }
        // waves per block and blocks per CU together: the LDS tables are per block, so
        // more waves per block raise occupancy without more LDS
        static const u32 occ[][2] = {{16, 1}, {16, 2}, {8, 2}, {8, 3}, {8, 4}, {4, 4}, {4, 8}};
        for (const auto& o : occ) {
            ccs_tuning t = best;
            if (o[0] == t.waves_per_block && o[1] == t.blocks_per_cu) continue;
            t.waves_per_block = o[0];
            t.blocks_per_cu = o[1];
            if ((rc = try_tuning(t))) return rc;
        }
    }
    s->tune = best;
    rc = upload_tables(s);
    if (rc) return rc;
    if (best_out) *best_out = best;
    if (best_ms_out) *best_ms_out = best_ms;
    return CCS_OK;
}

// ---------------------------------------------------------------------------
// campaign ranges: the tiles whose candidates have `bits` bits, narrowed to a binary
// prefix and split into lanes (v13.cu:450-518 compute_prefix_tile_range, :3474-3573)
// ---------------------------------------------------------------------------

extern "C" int ccs_confirm_host(const ccs_shape* s, ccs_u128 tile_start, const ccs_survivor* h_surv,
                                uint32_t count, const ccs_prove_params* params, uint32_t nthreads,
                                ccs_chain_result* h_out, ccs_confirm_stats* stats) {
    if (!s || (count && (!h_surv || !h_out))) return fail(CCS_ERR_INVALID, "null argument");
    if (s->bases.empty()) return fail(CCS_ERR_INVALID, "shape has no bases");
    std::vector<ccs_u128> n(count);
    const u128 T0 = mk128(tile_start);
    for (uint32_t i = 0; i < count; ++i) {
        if (h_surv[i].base_idx >= s->bases.size()) return fail(CCS_ERR_INVALID, "survivor base_idx out of range");
        const u128 v = ((T0 + h_surv[i].tile_idx) * (u128)s->WP + h_surv[i].k_offset) * (u128)s->M +
                       s->bases[h_surv[i].base_idx];
        n[i].lo = (u64)v;
        n[i].hi = (u64)(v >> 64);
    }
    return ccs_confirm_values(n.data(), count, params, nthreads, h_out, stats);
}

extern "C" int ccs_shape_decompose(const ccs_shape* s, ccs_u128 nv, ccs_decomposition* out) {
    if (!s || !out) return fail(CCS_ERR_INVALID, "null argument");
    if (s->classes.empty()) return fail(CCS_ERR_INVALID, "shape was built without its class table");
    std::memset(out, 0, sizeof(*out));
    const u128 n = mk128(nv);
    out->n_mod6 = (u32)(n % 6);
    const u64 base = (u64)(n % s->M);
    out->base = base;
    const u128 k = (n - base) / s->M;
    const u128 tile = k / s->WP;
    out->k_offset = (u32)(k % s->WP);
    out->tile.lo = (u64)tile;
    out->tile.hi = (u64)(tile >> 64);
    const auto it = std::lower_bound(s->bases.begin(), s->bases.end(), base);
    if (it == s->bases.end() || *it != base) return CCS_OK;
    out->base_found = 1;
    out->base_idx = (u32)(it - s->bases.begin());
    for (const ClassRec& c : s->classes) {
        if (c.base_idx != out->base_idx || c.k != out->k_offset) continue;
        out->class_found = 1;
        out->bucket_ok = (u32)(tile % 6) == c.rho;
        break;
    }
    out->on_lattice = out->base_found && out->class_found && out->bucket_ok && out->n_mod6 == 5;
    return CCS_OK;
}

extern "C" int ccs_tile_range(const ccs_shape* s, int32_t bits, ccs_u128 prefix_value, int32_t prefix_bits,
                              uint32_t lanes, uint32_t lane_id, ccs_u128* tile_min, ccs_u128* tile_max) {
    if (!s || !tile_min || !tile_max) return fail(CCS_ERR_INVALID, "null argument");
    if (bits < 21 || bits > 127) return fail(CCS_ERR_INVALID, "bits must be in 21..127");
    if (s->bases.empty()) return fail(CCS_ERR_INVALID, "shape has no bases");
    const u128 M = s->M, WP = s->WP;
    const u128 min_base = s->bases.front(), max_base = s->bases.back();
    // full bit range: conservative k bounds with the largest / smallest base
    const u128 n_min = (u128)1 << (bits - 1), n_max = ((u128)1 << bits) - 1;
    if (n_min <= max_base) return fail(CCS_ERR_INVALID, "bits too small for the lattice bases");
    u128 tmin = ((n_min - max_base) / M) / WP;
    u128 tmax = ((n_max - min_base) / M) / WP + 1;
    if (prefix_bits > 0) {
        if (prefix_bits >= bits) return fail(CCS_ERR_INVALID, "prefix_bits must be < bits");
        const u128 pv = mk128(prefix_value);
        if (prefix_bits < 127 && (pv >> (prefix_bits - 1)) != 1)
            return fail(CCS_ERR_INVALID, "binary prefix must start with 1 and have prefix_bits bits");
        const int shift = bits - prefix_bits;
        u128 cmin = pv << shift, cmax = ((pv + 1) << shift) - 1;
        if (cmax < n_min || cmin > n_max) return fail(CCS_ERR_INVALID, "prefix does not overlap the bit range");
        if (cmin < n_min) cmin = n_min;
        if (cmax > n_max) cmax = n_max;
        const u128 kmin = cmin > max_base ? (cmin - max_base + M - 1) / M : 0;
        const u128 kmax = (cmax - min_base) / M;
        if (kmin > kmax) return fail(CCS_ERR_INVALID, "prefix produces an empty k range");
        u128 pmin = kmin / WP, pmax = kmax / WP + 1;
        if (pmin < tmin) pmin = tmin;
        if (pmax > tmax) pmax = tmax;
        if (pmin >= pmax) return fail(CCS_ERR_INVALID, "prefix produces an empty tile range");
        tmin = pmin;
        tmax = pmax;
        if (lanes > 0) {
            if (lane_id >= lanes) return fail(CCS_ERR_INVALID, "lane_id must be < lanes");
            const u128 total = tmax - tmin, size = total / lanes, rem = total - size * lanes;
            u128 start = tmin + (u128)lane_id * size + (lane_id < rem ? lane_id : rem);
            u128 end = start + size + (lane_id < rem ? 1 : 0);
            tmin = start;
            tmax = end;
        }
    } else if (lanes > 0) {
        return fail(CCS_ERR_INVALID, "lanes need a prefix");
    }
    tile_min->lo = (u64)tmin;
    tile_min->hi = (u64)(tmin >> 64);
    tile_max->lo = (u64)tmax;
    tile_max->hi = (u64)(tmax >> 64);
    return CCS_OK;
}
