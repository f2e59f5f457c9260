// ccs_search: the campaign command line of the reference GPU engine (v13.cu:3068-4280) on
// libccs.so.  Same options and output files; the GPU work is ccs_filter_launch +
// ccs_prove_launch on two streams (the prove of batch k overlaps the filter of batch k+1),
// the tile walk / stop rules / checkpoint are the library's campaign state.
//
// Host code only: it talks to the GPU through the C ABI in include/ccs.h.
#include <cerrno>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <deque>
#include <string>
#include <vector>

#include <unistd.h>

#include "ccs.h"

namespace {

volatile std::sig_atomic_t g_stop = 0;

void on_signal(int) {   // first signal: finish the batches in flight; second: exit now
    if (!g_stop) {
        g_stop = 1;
        const char m[] = "\n[INTERRUPT] stopping after the batches in flight (Ctrl+C again = force exit)\n";
        (void)!write(STDERR_FILENO, m, sizeof(m) - 1);
        return;
    }
    const char m[] = "\n[INTERRUPT] force exit\n";
    (void)!write(STDERR_FILENO, m, sizeof(m) - 1);
    _exit(130);
}

struct Options {
    int target = 18, bits = 89, report = 5, log = 8, depth = 18, gpu = 0, time_limit = 0;
    uint64_t tiles = 0;
    bool have_start = false, have_seed = false, cpu_prove = false, confirm = false, test = false;
    unsigned __int128 start = 0;
    uint64_t seed = 0;
    uint32_t batch = 16384, prove_threads = 0, filter_max = 863;
    const char* output = nullptr;
    const char* prefix = nullptr;
    const char* prefix_mode = nullptr;
    int lanes = 0, lane_id = 0;
    const char* checkpoint = nullptr;
    const char* resume = nullptr;
    const char* verify_root = nullptr;
    double checkpoint_interval = 60.0;
};

int usage(const char* argv0) {
    std::printf(
        "Cunningham-chain (first kind) search on libccs.so\n\n"
        "Usage: %s [options]\n\n"
        "  --target N         Target chain length (default 18)\n"
        "  --bits N           Bit size of the candidates (default 89)\n"
        "  --report N         Report interval in seconds (default 5)\n"
        "  --log N            Min chain length to log (default 8)\n"
        "  --tiles N          Number of tiles (0 = continuous)\n"
        "  --start N          Starting tile (sequential walk)\n"
        "  --depth N          Filter depth (default 18, clamped to the target)\n"
        "  --filter-max Q     Largest filter prime (default 863, as the reference; up to 65521)\n"
        "  --output FILE      Append results to FILE\n"
        "  --gpu N            GPU index (default 0)\n"
        "  --batch N          Tiles per GPU batch (default 16384)\n"
        "  --time N           Stop after N seconds (graceful)\n"
        "  --seed N           PRNG seed of the random walk (reproducible)\n"
        "  --cpu-prove        Prove on host threads instead of the GPU\n"
        "  --prove-threads N  Host prover threads (implies --cpu-prove)\n"
        "  --confirm          Re-check every logged chain on the host\n"
        "  --coverage full    All CRT bases (the only coverage mode here)\n"
        "  --verify-root HEX  Where a root sits on the lattice, and whether it survives\n"
        "  --prefix 0b...     Binary prefix of the candidates' high bits\n"
        "  --prefix-mode M    sequential or random\n"
        "  --prefix-lanes N   Split the prefix range into N lanes\n"
        "  --prefix-lane-id I This run's lane (0-based)\n"
        "  --checkpoint FILE  Save progress every 60 s and at exit\n"
        "  --resume FILE      Continue from a checkpoint\n"
        "  --test             Self-test (host only, no GPU)\n",
        argv0);
    return 0;
}

bool parse_u128(const char* s, unsigned __int128* out, int base) {
    unsigned __int128 v = 0;
    if (base == 16 && s[0] == '0' && (s[1] == 'x' || s[1] == 'X')) s += 2;
    if (!*s) return false;
    for (; *s; ++s) {
        int d;
        if (*s >= '0' && *s <= '9') d = *s - '0';
        else if (base == 16 && *s >= 'a' && *s <= 'f') d = *s - 'a' + 10;
        else if (base == 16 && *s >= 'A' && *s <= 'F') d = *s - 'A' + 10;
        else return false;
        const unsigned __int128 nv = v * (unsigned)base + (unsigned)d;
        if (nv / (unsigned)base != v) return false;   // overflow
        v = nv;
    }
    *out = v;
    return true;
}

ccs_u128 u(unsigned __int128 v) { return ccs_u128{(uint64_t)v, (uint64_t)(v >> 64)}; }
unsigned __int128 v128(ccs_u128 x) { return ((unsigned __int128)x.hi << 64) | x.lo; }

std::string hex128(unsigned __int128 v) {
    char buf[40];
    const uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    if (hi) std::snprintf(buf, sizeof(buf), "0x%llX%016llX", (unsigned long long)hi, (unsigned long long)lo);
    else std::snprintf(buf, sizeof(buf), "0x%llX", (unsigned long long)lo);
    return buf;
}

int bitlen128(unsigned __int128 v) {
    int n = 0;
    while (v) { ++n; v >>= 1; }
    return n;
}

std::vector<uint32_t> primes_between(uint32_t lo, uint32_t hi) {
    std::vector<uint32_t> out;
    for (uint32_t q = lo; q <= hi; ++q) {
        bool p = q > 1;
        for (uint32_t d = 2; d * d <= q && p; ++d) p = q % d != 0;
        if (p) out.push_back(q);
    }
    return out;
}

// host-only self test: chain lengths of the classic roots (A005602), prefix parsing and the
// campaign walk
int self_test() {
    int fails = 0;
    struct { unsigned long long root; uint32_t len; } chains[] = {
        {1122659ull, 7}, {19099919ull, 8}, {85864769ull, 9}, {26089808579ull, 10},
        {665043081119ull, 11}, {554688278429ull, 12}, {4090932431513069ull, 13}};
    ccs_prove_params pp{0, CCS_TEST_BPSW, 0, {0}};
    for (auto& c : chains) {
        ccs_u128 n = u(c.root);
        ccs_chain_result r;
        const int rc = ccs_confirm_values(&n, 1, &pp, 1, &r, nullptr);
        const bool ok = rc == CCS_OK && r.chain_len == c.len && (r.flags & CCS_F_PRIME) && !r.back_steps;
        std::printf("  chain %-18llu CC%-2u  %s\n", c.root, r.chain_len, ok ? "PASS" : "FAIL");
        fails += !ok;
    }
    ccs_u128 pv;
    int32_t pb;
    const bool p1 = ccs_parse_prefix("0b101", &pv, &pb) == CCS_OK && pv.lo == 5 && pb == 3;
    const bool p2 = ccs_parse_prefix("0b0101", &pv, &pb) != CCS_OK;
    std::printf("  prefix parsing      %s\n", p1 && p2 ? "PASS" : "FAIL");
    fails += !(p1 && p2);
    ccs_campaign_params cp{};
    cp.bits = 89; cp.target = 18; cp.depth = 18; cp.prefix = "0b1"; cp.walk = CCS_WALK_SEQUENTIAL;
    cp.has_seed = 1; cp.seed = 1; cp.batch_tiles = 7;
    ccs_campaign* c = nullptr;
    uint64_t covered = 0;
    bool cok = ccs_campaign_create_range(u(100), u(150), &cp, &c) == CCS_OK;
    for (;;) {
        ccs_u128 ts;
        uint32_t tc;
        const int rc = c ? ccs_campaign_next(c, &ts, &tc) : CCS_ERR_INVALID;
        if (rc == CCS_DONE) break;
        if (rc != CCS_OK || ccs_campaign_record(c, ts, tc, 0, 0, nullptr, 0) != CCS_OK) { cok = false; break; }
        covered += tc;
    }
    cok = cok && covered == 50;
    ccs_campaign_free(c);
    std::printf("  sequential sweep    %s\n", cok ? "PASS" : "FAIL");
    fails += !cok;
    std::printf("\n  self-test: %s\n", fails ? "FAILED" : "all passed");
    return fails ? 1 : 0;
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

#define CHECK(call)                                                                     \
    do {                                                                                \
        const int rc_ = (call);                                                         \
        if (rc_ < 0) {                                                                  \
            std::fprintf(stderr, "ERROR: %s: %s\n", #call, ccs_last_error());           \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

struct Slot {               // one batch's device buffers and its two events
    ccs_survivor* d_surv = nullptr;
    uint32_t* d_count = nullptr;
    ccs_chain_result* d_res = nullptr;
    void* ev_filtered = nullptr;
    void* ev_proved = nullptr;
};

struct InFlight {
    int slot;
    ccs_u128 ts;
    uint32_t tc;
};

}  // namespace

int main(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&](const char* what) -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "ERROR: %s needs a value\n", what); std::exit(1); }
            return argv[++i];
        };
        if (a == "--target") o.target = std::atoi(val("--target"));
        else if (a == "--bits") o.bits = std::atoi(val("--bits"));
        else if (a == "--report") o.report = std::atoi(val("--report"));
        else if (a == "--log") o.log = std::atoi(val("--log"));
        else if (a == "--tiles") o.tiles = std::strtoull(val("--tiles"), nullptr, 10);
        else if (a == "--start") {
            if (!parse_u128(val("--start"), &o.start, 10)) { std::fprintf(stderr, "ERROR: bad --start\n"); return 1; }
            o.have_start = true;
        }
        else if (a == "--depth") o.depth = std::atoi(val("--depth"));
        else if (a == "--filter-max") o.filter_max = (uint32_t)std::atoi(val("--filter-max"));
        else if (a == "--output") o.output = val("--output");
        else if (a == "--gpu") o.gpu = std::atoi(val("--gpu"));
        else if (a == "--batch") {
            const long b = std::atol(val("--batch"));
            if (b < 1 || b > 12500000) { std::fprintf(stderr, "ERROR: --batch %ld out of range [1, 12500000]\n", b); return 1; }
            o.batch = (uint32_t)b;
        }
        else if (a == "--time") o.time_limit = std::atoi(val("--time"));
        else if (a == "--seed") { o.seed = std::strtoull(val("--seed"), nullptr, 0); o.have_seed = true; }
        else if (a == "--cpu-prove") o.cpu_prove = true;
        else if (a == "--prove-threads") { o.prove_threads = (uint32_t)std::atoi(val("--prove-threads")); o.cpu_prove = true; }
        else if (a == "--confirm") o.confirm = true;
        else if (a == "--coverage") {
            const std::string m = val("--coverage");
            if (m != "full") {
                std::fprintf(stderr, "ERROR: --coverage %s is not supported here (only 'full': all CRT bases)\n", m.c_str());
                return 1;
            }
        }
        else if (a == "--verify-root") o.verify_root = val("--verify-root");
        else if (a == "--prefix") o.prefix = val("--prefix");
        else if (a == "--prefix-mode") o.prefix_mode = val("--prefix-mode");
        else if (a == "--prefix-lanes") o.lanes = std::atoi(val("--prefix-lanes"));
        else if (a == "--prefix-lane-id") o.lane_id = std::atoi(val("--prefix-lane-id"));
        else if (a == "--checkpoint") o.checkpoint = val("--checkpoint");
        else if (a == "--resume") o.resume = val("--resume");
        else if (a == "--test") o.test = true;
        else if (a == "--continuous") {}
        else if (a == "--help" || a == "-h") return usage(argv[0]);
        else { std::fprintf(stderr, "ERROR: Unknown argument: %s (use --help)\n", argv[i]); return 1; }
    }
    if (o.test) return self_test();

    // validation that needs no GPU (v13.cu:3176-3264)
    int walk = CCS_WALK_INHERIT;
    ccs_u128 pv{0, 0};
    int32_t pb = 0;
    if (o.prefix) {
        if (ccs_parse_prefix(o.prefix, &pv, &pb) != CCS_OK) { std::fprintf(stderr, "ERROR: %s\n", ccs_last_error()); return 1; }
        if (pb >= o.bits) {
            std::fprintf(stderr, "ERROR: prefix has %d bits but --bits is %d (prefix_bits must be < bits)\n", pb, o.bits);
            return 1;
        }
    }
    if (o.prefix_mode) {
        if (!o.prefix) { std::fprintf(stderr, "ERROR: --prefix-mode requires --prefix\n"); return 1; }
        if (!std::strcmp(o.prefix_mode, "sequential")) walk = CCS_WALK_SEQUENTIAL;
        else if (!std::strcmp(o.prefix_mode, "random")) walk = CCS_WALK_RANDOM;
        else { std::fprintf(stderr, "ERROR: --prefix-mode must be 'sequential' or 'random', got '%s'\n", o.prefix_mode); return 1; }
    }
    if (o.lanes > 0) {
        if (!o.prefix) { std::fprintf(stderr, "ERROR: --prefix-lanes requires --prefix\n"); return 1; }
        if (o.lane_id < 0 || o.lane_id >= o.lanes) {
            std::fprintf(stderr, "ERROR: --prefix-lane-id %d out of range [0, %d)\n", o.lane_id, o.lanes);
            return 1;
        }
    }
    if (o.bits < 21 || o.bits > 127) { std::fprintf(stderr, "ERROR: --bits %d out of supported range [21, 127]\n", o.bits); return 1; }
    if (o.target < 1 || o.target > 64) { std::fprintf(stderr, "ERROR: --target %d out of range [1, 64]\n", o.target); return 1; }
    if (o.filter_max < 37 || o.filter_max > 65521) { std::fprintf(stderr, "ERROR: --filter-max must be in [37, 65521]\n"); return 1; }
    if (o.depth < 1) { std::fprintf(stderr, "WARNING: --depth %d clamped to 1\n", o.depth); o.depth = 1; }
    if (o.depth > o.target) {
        std::fprintf(stderr, "NOTE: --depth %d > --target %d, clamped to %d (filtering beyond target over-rejects)\n",
                     o.depth, o.target, o.target);
        o.depth = o.target;
    }

    std::signal(SIGINT, on_signal);
    std::signal(SIGTERM, on_signal);

    ccs_ctx* ctx = nullptr;
    CHECK(ccs_ctx_create(o.gpu, &ctx));
    // the reference's lattice: CRT 5..19 x wheel 23, 29, 31 x mod 6; filter stages L2 37..61,
    // ext-L2 67..97, line sieve 101..filter_max (v13.cu:155-199)
    static const uint32_t CRT[] = {5, 7, 11, 13, 17, 19}, WHEEL[] = {23, 29, 31};
    const std::vector<uint32_t> fp = primes_between(37, o.filter_max);
    uint32_t stage_end[3] = {7, 14, (uint32_t)fp.size()};
    ccs_shape_params sp{6, CRT, 3, WHEEL, o.target, o.depth, (uint32_t)fp.size(), fp.data(), 3, stage_end, 97, 0};
    ccs_shape* shape = nullptr;
    CHECK(ccs_shape_build(ctx, &sp, nullptr, &shape));
    ccs_shape_info info;
    CHECK(ccs_shape_query(shape, &info));

    if (o.verify_root) {   // --verify-root (v13.cu:2921-3066)
        unsigned __int128 n;
        if (!parse_u128(o.verify_root, &n, 16)) { std::fprintf(stderr, "ERROR: --verify-root: cannot parse '%s' as hex (< 2^128)\n", o.verify_root); return 1; }
        std::printf("\n  === Verify Root: %s ===\n", o.verify_root);
        std::printf("    n = %s (%d bits)\n", hex128(n).c_str(), bitlen128(n));
        ccs_decomposition d;
        CHECK(ccs_shape_decompose(shape, u(n), &d));
        std::printf("    base = %llu, tile = %s, k_offset = %u, n%%6 = %u%s\n", (unsigned long long)d.base,
                    hex128(v128(d.tile)).c_str(), d.k_offset, d.n_mod6, d.n_mod6 == 5 ? " (first kind)" : " (NOT first kind)");
        std::printf("    base %s, class %s, mod-6 bucket %s\n", d.base_found ? "admissible" : "NOT admissible",
                    d.class_found ? "admissible" : "NOT admissible", d.bucket_ok ? "ok" : "NOT ok");
        if (!d.on_lattice) {
            std::printf("    RESULT: NOT REACHABLE on this lattice\n");
        } else {
            std::vector<ccs_survivor> surv(1u << 12);
            uint32_t ns = 0;
            ccs_filter_stats fst;
            CHECK(ccs_filter(ctx, shape, d.tile, 1, surv.data(), (uint32_t)surv.size(), &ns, &fst));
            bool survives = false;
            for (uint32_t i = 0; i < ns && i < surv.size(); ++i)
                survives |= surv[i].base_idx == d.base_idx && surv[i].k_offset == d.k_offset;
            std::printf("    depth-%d filter (%u primes 37..%u): %s\n", o.depth, (unsigned)fp.size(), o.filter_max,
                        survives ? "SURVIVES" : "rejected");
            ccs_prove_params pp{0, CCS_TEST_BPSW, 0, {0}};
            ccs_u128 nv = u(n);
            ccs_chain_result r;
            CHECK(ccs_confirm_values(&nv, 1, &pp, 1, &r, nullptr));
            std::printf("    chain: %s CC%u from root %s (%u steps back)\n", (r.flags & CCS_F_PRIME) ? "prime," : "not prime:",
                        r.chain_len, hex128(((unsigned __int128)r.root_hi << 64) | r.root_lo).c_str(), r.back_steps);
            std::printf("    RESULT: REACHABLE\n");
        }
        ccs_shape_free(shape);
        ccs_ctx_destroy(ctx);
        return 0;
    }

    ccs_campaign_params cp{};
    cp.bits = o.bits; cp.target = o.target; cp.depth = o.depth;
    cp.prefix = o.prefix; cp.walk = walk;
    cp.lanes = (uint32_t)o.lanes; cp.lane_id = (uint32_t)o.lane_id;
    cp.has_start = o.have_start; cp.start = u(o.start);
    cp.has_seed = o.have_seed; cp.seed = o.seed;
    cp.batch_tiles = o.batch; cp.max_tiles = o.tiles;
    ccs_campaign* camp = nullptr;
    CHECK(ccs_campaign_create(shape, &cp, &camp));
    if (o.resume) {
        CHECK(ccs_campaign_resume(camp, o.resume));
        std::printf("[RESUME] Continuing from checkpoint %s (sequential)\n", o.resume);
    }
    ccs_campaign_state st;
    CHECK(ccs_campaign_get_state(camp, &st));

    std::printf("\n================================================================\n");
    std::printf("  Cunningham-chain search (first kind) on libccs.so\n");
    std::printf("  Target: CC%d, %d-bit, depth=%d, filter primes 37..%u, log>=CC%d, gpu=%d\n", o.target, o.bits,
                o.depth, o.filter_max, o.log, o.gpu);
    std::printf("  Coverage: FULL (%u bases, %llu classes)\n", info.n_bases, (unsigned long long)info.n_classes);
    std::printf("  Prove: %s\n", o.cpu_prove ? "host threads (Baillie-PSW)" : "GPU (Baillie-PSW)");
    std::printf("  Tile range: [%s, %s)\n", hex128(v128(st.tile_min)).c_str(), hex128(v128(st.tile_max)).c_str());
    std::printf("  Mode: %s%s, seed 0x%016llX (%s)\n", st.random_mode ? "RANDOM tile selection" : "SEQUENTIAL",
                o.prefix ? " (within prefix subspace)" : "", (unsigned long long)st.seed,
                o.have_seed ? "--seed" : "/dev/urandom");
    if (o.prefix) std::printf("  Prefix: %s (%d bits)%s\n", o.prefix, pb, o.lanes ? "" : "");
    if (o.lanes) std::printf("  Prefix lane: %d/%d\n", o.lane_id, o.lanes);
    std::printf("================================================================\n\n");

    FILE* out = nullptr;
    if (o.output) {
        out = std::fopen(o.output, "a");
        if (!out) std::fprintf(stderr, "WARNING: Cannot open output file '%s', continuing without\n", o.output);
        else {
            std::fprintf(out, "# CCS search — target=%d bits=%d depth=%d log>=%d", o.target, o.bits, o.depth, o.log);
            if (o.prefix) std::fprintf(out, " prefix=%s", o.prefix);
            std::fprintf(out, "\n# Started: %ld\n", (long)std::time(nullptr));
            std::fflush(out);
        }
    }

    const uint32_t surv_cap = 1u << 20, res_cap = 1u << 16;
    void *s_filter = nullptr, *s_prove = nullptr;
    CHECK(ccs_stream_create(ctx, &s_filter));
    CHECK(ccs_stream_create(ctx, &s_prove));
    // three batches in flight: the host collects batch k-3 (its prove long finished) before
    // it launches batch k, so a prove that runs in the next filter's tail never holds the
    // host back from queueing the filter after it
    constexpr int kSlots = 3;
    Slot slots[kSlots];
    for (auto& sl : slots) {
        CHECK(ccs_dmalloc(ctx, (void**)&sl.d_surv, (uint64_t)surv_cap * sizeof(ccs_survivor)));
        CHECK(ccs_dmalloc(ctx, (void**)&sl.d_count, 64));
        CHECK(ccs_dmalloc(ctx, (void**)&sl.d_res, (uint64_t)res_cap * sizeof(ccs_chain_result)));
        CHECK(ccs_event_create(ctx, &sl.ev_filtered));
        CHECK(ccs_event_create(ctx, &sl.ev_proved));
    }
    ccs_prove_params pp{o.bits, CCS_TEST_BPSW, 0, {0}};
    std::vector<ccs_survivor> h_surv;
    std::vector<ccs_chain_result> h_res;
    std::deque<InFlight> flight;
    const double t0 = now_s();
    double last_report = t0, last_ckpt = t0;
    int best = (int)st.chains.best_chain;
    uint64_t mismatches = 0;

    // stage-B overflow of the filter launches since the last check (read at report time and
    // at the end: the read waits for the filter stream, so not per batch); the reference
    // warns and clamps on survivor overflow (v13.cu:3846-3852)
    auto check_stage_b = [&]() -> int {
        ccs_launch_status ls;
        const int rc = ccs_stream_status(ctx, s_filter, &ls);
        if (rc < 0) {
            std::fprintf(stderr, "ccs_stream_status: %s\n", ccs_last_error());
            return 1;
        }
        if (rc == CCS_OVERFLOW)
            std::fprintf(stderr, "WARNING: stage-B buffer overflow in %u batch(es) (worst %u entries, buffer %u): "
                         "%llu candidates dropped before the deep stage\n", ls.overflowed_launches, ls.worst_stage_b,
                         ls.stage_b_cap, (unsigned long long)ls.dropped);
        return 0;
    };

    // collect the oldest batch: counts, results, lines, stats
    auto collect = [&]() -> int {
        const InFlight f = flight.front();
        Slot& sl = slots[f.slot];
        CHECK(ccs_event_synchronize(ctx, o.cpu_prove ? sl.ev_filtered : sl.ev_proved));
        uint32_t n = 0;
        CHECK(ccs_memcpy_d2h(ctx, &n, sl.d_count, 4));
        if (n > surv_cap) std::fprintf(stderr, "WARNING: survivor overflow: %u survivors, %u kept\n", n, surv_cap);
        const uint32_t ns = n < surv_cap ? n : surv_cap;
        const uint32_t nr = o.cpu_prove ? ns : (ns < res_cap ? ns : res_cap);
        if (!o.cpu_prove && ns > res_cap) std::fprintf(stderr, "WARNING: prove overflow: %u survivors, %u proven\n", ns, res_cap);
        h_res.resize(nr);
        if (o.cpu_prove || o.confirm) {
            h_surv.resize(ns);
            if (ns) CHECK(ccs_memcpy_d2h(ctx, h_surv.data(), sl.d_surv, (uint64_t)ns * sizeof(ccs_survivor)));
        }
        if (o.cpu_prove) {
            if (ns) CHECK(ccs_confirm_host(shape, f.ts, h_surv.data(), ns, &pp, o.prove_threads, h_res.data(), nullptr));
        } else if (nr) {
            CHECK(ccs_memcpy_d2h(ctx, h_res.data(), sl.d_res, (uint64_t)nr * sizeof(ccs_chain_result)));
        }
        if (o.confirm && !o.cpu_prove) {    // host re-check of every logged chain
            std::vector<ccs_survivor> sub;
            std::vector<const ccs_chain_result*> gres;
            for (uint32_t i = 0; i < nr; ++i)
                if ((h_res[i].flags & CCS_F_PRIME) && ((int)h_res[i].chain_len >= o.log || (int)h_res[i].chain_len >= o.target)) {
                    sub.push_back(h_surv[h_res[i].survivor_idx]);
                    gres.push_back(&h_res[i]);
                }
            std::vector<ccs_chain_result> hr(sub.size());
            if (!sub.empty()) CHECK(ccs_confirm_host(shape, f.ts, sub.data(), (uint32_t)sub.size(), &pp, o.prove_threads, hr.data(), nullptr));
            for (size_t i = 0; i < sub.size(); ++i) {
                const ccs_chain_result& g = *gres[i];
                if (hr[i].chain_len != g.chain_len || hr[i].flags != g.flags || hr[i].root_lo != g.root_lo ||
                    hr[i].root_hi != g.root_hi) {
                    ++mismatches;
                    std::fprintf(stderr, "ERROR: host confirmation differs for survivor %u (GPU CC%u, host CC%u)\n",
                                 g.survivor_idx, g.chain_len, hr[i].chain_len);
                }
            }
        }
        const uint64_t cand = ccs_count_candidates(shape, f.ts, f.tc);
        CHECK(ccs_campaign_record(camp, f.ts, f.tc, cand, ns, h_res.data(), nr));
        for (uint32_t i = 0; i < nr; ++i) {
            const ccs_chain_result& r = h_res[i];
            if (!(r.flags & CCS_F_PRIME)) continue;
            const int cl = (int)r.chain_len;
            if (cl < o.log && cl < o.target) continue;
            const unsigned __int128 root = ((unsigned __int128)r.root_hi << 64) | r.root_lo;
            const char* nr_tag = (r.flags & CCS_F_NONROOT) ? " [NON-ROOT]" : "";
            if (cl >= o.target) std::printf("\n  !!!! CC%d TARGET HIT: %s (%d bits)%s !!!!\n", cl, hex128(root).c_str(), bitlen128(root), nr_tag);
            else if (cl > best) std::printf("\n  *** CC%d: %s%s (%d bits) ***\n", cl, hex128(root).c_str(), nr_tag, bitlen128(root));
            else std::printf("  CC%d: %s%s (%d bits)\n", cl, hex128(root).c_str(), nr_tag, bitlen128(root));
            if (cl > best) best = cl;
        }
        if (out && nr) {
            uint64_t need = 0;
            ccs_format_results(h_res.data(), nr, o.log, o.target, nullptr, 0, &need);
            std::string buf(need + 1, '\0');
            ccs_format_results(h_res.data(), nr, o.log, o.target, &buf[0], need + 1, &need);
            std::fputs(buf.c_str(), out);
            std::fflush(out);
        }
        flight.pop_front();
        return 0;
    };

    int k = 0;
    bool stop_msg = false;
    for (;;) {
        if (g_stop) break;
        if (o.time_limit > 0 && now_s() - t0 >= o.time_limit) {
            std::printf("[TIME] %d-second limit reached (%.1fs elapsed). Stopping gracefully.\n", o.time_limit, now_s() - t0);
            break;
        }
        ccs_u128 ts;
        uint32_t tc = 0;
        const int rc = ccs_campaign_next(camp, &ts, &tc);
        if (rc == CCS_DONE) {
            CHECK(ccs_campaign_get_state(camp, &st));
            if (!st.random_mode && o.prefix) {
                std::printf("\n[PREFIX] Sequential sweep complete. All prefix tiles visited once. Stopping.\n");
                stop_msg = true;
            }
            break;
        }
        CHECK(rc);
        const int b = k % kSlots;
        if (flight.size() == (size_t)kSlots && collect()) return 1;   // slot b's previous batch
        Slot& sl = slots[b];
        CHECK(ccs_filter_launch(ctx, shape, ts, tc, sl.d_surv, surv_cap, sl.d_count, s_filter));
        CHECK(ccs_event_record(ctx, sl.ev_filtered, s_filter));
        if (!o.cpu_prove) {
            CHECK(ccs_stream_wait_event(ctx, s_prove, sl.ev_filtered));
            CHECK(ccs_prove_launch(ctx, shape, ts, sl.d_surv, sl.d_count, res_cap, &pp, sl.d_res, s_prove));
            CHECK(ccs_event_record(ctx, sl.ev_proved, s_prove));
        }
        // slot b is reused by batch k+3 only after collect() has waited for this batch's
        // prove on the host (at most three batches are in flight)
        flight.push_back({b, ts, tc});
        ++k;
        if (o.cpu_prove && flight.size() == 2 && collect()) return 1;   // host proves while the GPU filters
        const double t = now_s();
        if (t - last_report >= o.report) {
            if (check_stage_b()) return 1;
            CHECK(ccs_campaign_get_state(camp, &st));
            const double el = t - t0;
            std::printf("[%.1fs] tiles=%llu  cand=%.2fB  surv=%llu (%.6f%%)  primes=%llu  best=CC%u  rate=%.2fB/s",
                        el, (unsigned long long)st.total_tiles, st.total_candidates / 1e9,
                        (unsigned long long)st.total_survivors,
                        st.total_candidates ? 100.0 * st.total_survivors / st.total_candidates : 0.0,
                        (unsigned long long)st.chains.primes, st.chains.best_chain, st.total_candidates / el / 1e9);
            if (st.chains.non_roots) std::printf("  nr=%llu", (unsigned long long)st.chains.non_roots);
            bool open = false;
            for (int cl = o.log; cl <= 64; ++cl)
                if (st.chains.chain_dist[cl]) {
                    std::printf("%sCC%d:%llu", open ? " " : "  [", cl, (unsigned long long)st.chains.chain_dist[cl]);
                    open = true;
                }
            std::printf("%s\n", open ? "]" : "");
            std::fflush(stdout);
            last_report = t;
            if (o.checkpoint && t - last_ckpt >= o.checkpoint_interval) {
                if (ccs_campaign_save(camp, o.checkpoint) != CCS_OK)
                    std::fprintf(stderr, "WARNING: checkpoint save failed: %s\n", ccs_last_error());
                last_ckpt = t;
            }
        }
    }
    while (!flight.empty())
        if (collect()) return 1;
    if (check_stage_b()) return 1;
    (void)stop_msg;

    CHECK(ccs_campaign_get_state(camp, &st));
    const double el = now_s() - t0;
    std::printf("\n================================================================\n");
    std::printf("  FINAL RESULTS\n");
    std::printf("================================================================\n");
    std::printf("  Time:         %.2f seconds\n", el);
    std::printf("  Tiles:        %llu\n", (unsigned long long)st.total_tiles);
    std::printf("  Candidates:   %llu (%.2f B)\n", (unsigned long long)st.total_candidates, st.total_candidates / 1e9);
    std::printf("  Survivors:    %llu (%.6f%%)\n", (unsigned long long)st.total_survivors,
                st.total_candidates ? 100.0 * st.total_survivors / st.total_candidates : 0.0);
    std::printf("  Primes:       %llu\n", (unsigned long long)st.chains.primes);
    std::printf("  Chains:       %llu (best: CC%u)\n", (unsigned long long)st.chains.chains, st.chains.best_chain);
    if (st.chains.non_roots)
        std::printf("  Non-roots:    %llu (walked backward to true root)\n", (unsigned long long)st.chains.non_roots);
    std::printf("  Throughput:   %.2f B candidates/sec\n", el > 0 ? st.total_candidates / el / 1e9 : 0.0);
    if (o.confirm) std::printf("  Host confirm: %s\n", mismatches ? "MISMATCHES" : "all logged chains agree");
    std::printf("\n  Chain distribution:\n");
    for (int cl = 1; cl <= 64; ++cl)
        if (st.chains.chain_dist[cl])
            std::printf("    CC%-2d: %llu%s\n", cl, (unsigned long long)st.chains.chain_dist[cl], cl >= o.target ? "  <-- TARGET" : "");
    std::printf("================================================================\n\n");
    if (o.checkpoint) {
        if (ccs_campaign_save(camp, o.checkpoint) == CCS_OK) std::printf("[CHECKPOINT] Saved final state to %s\n", o.checkpoint);
        else std::fprintf(stderr, "WARNING: checkpoint save failed: %s\n", ccs_last_error());
    }
    if (out) {
        std::fprintf(out, "# Finished: %ld  tiles=%llu  primes=%llu  best=CC%u", (long)std::time(nullptr),
                     (unsigned long long)st.total_tiles, (unsigned long long)st.chains.primes, st.chains.best_chain);
        if (o.prefix) std::fprintf(out, "  prefix=%s", o.prefix);
        std::fprintf(out, "\n");
        std::fclose(out);
    }
    for (auto& sl : slots) {
        ccs_dfree(ctx, sl.d_surv);
        ccs_dfree(ctx, sl.d_count);
        ccs_dfree(ctx, sl.d_res);
        ccs_event_destroy(ctx, sl.ev_filtered);
        ccs_event_destroy(ctx, sl.ev_proved);
    }
    ccs_stream_destroy(ctx, s_filter);
    ccs_stream_destroy(ctx, s_prove);
    ccs_campaign_free(camp);
    ccs_shape_free(shape);
    ccs_ctx_destroy(ctx);
    return mismatches ? 2 : 0;
}
