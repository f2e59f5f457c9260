// Campaign driver state: the v13 main loop's range, tile walk, stop rules, statistics and
// checkpoint file, separated from the GPU work so that any host (the ccs_search CLI, a
// service, a test) drives the same semantics.
//
//   range      --bits / --prefix / --prefix-lanes            v13.cu:3474-3573 (ccs_tile_range)
//   walk       random (xorshift64, seeded) or sequential     v13.cu:3584-3601, :3734-3765, :3960-3970
//   stop       --tiles, end of a sequential prefix sweep     v13.cu:3719-3721, :3749-3757
//   resume     sequential from the checkpoint tile           v13.cu:3646-3667
//   checkpoint CC18_CUDA_v10 text file, atomic rename        v13.cu:520-664
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <string>

#include <fcntl.h>
#include <time.h>
#include <unistd.h>

#include "../../include/ccs.h"
#include "ccs_internal.hpp"

typedef unsigned __int128 u128;

struct ccs_campaign {
    ccs_campaign_params p{};
    std::string prefix;        // as given ("0b..."), or "none"
    u128 tmin = 0, tmax = 0, cur = 0;
    bool random_mode = true;
    bool done = false;
    uint64_t rng = 1, seed = 0;
    uint64_t handed = 0;       // batches handed out by next()
    std::deque<std::pair<u128, uint32_t>> outstanding;   // handed out, not yet recorded
    uint64_t batches = 0, total_tiles = 0, total_candidates = 0, total_survivors = 0;
    ccs_confirm_stats chains{};
};

namespace {

int cfail(int code, const std::string& msg) { return ccs::set_error(code, msg.c_str()); }

u128 mk(ccs_u128 v) { return ((u128)v.hi << 64) | v.lo; }
ccs_u128 un(u128 v) {
    ccs_u128 r;
    r.lo = (uint64_t)v;
    r.hi = (uint64_t)(v >> 64);
    return r;
}

uint64_t xorshift64(uint64_t& s) {   // rng_next (v13.cu:714-721)
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x;
}

uint64_t entropy_seed() {            // rng_seed_entropy (v13.cu:725-745)
    uint64_t seed = 0;
    const int fd = open("/dev/urandom", O_RDONLY);
    if (fd >= 0) {
        const ssize_t r = read(fd, &seed, sizeof(seed));
        close(fd);
        if (r == (ssize_t)sizeof(seed) && seed) return seed;
    }
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    seed = (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
    seed ^= (uint64_t)getpid() * 0x517CC1B727220A95ull;
    return seed ? seed : 1;
}

}  // namespace

extern "C" int ccs_parse_prefix(const char* str, ccs_u128* value, int32_t* bits) {
    if (!str || !value || !bits) return cfail(CCS_ERR_INVALID, "null argument");
    if (str[0] != '0' || (str[1] != 'b' && str[1] != 'B'))
        return cfail(CCS_ERR_INVALID, "--prefix must be binary format (0b...)");
    const char* q = str + 2;
    if (!*q) return cfail(CCS_ERR_INVALID, "--prefix has no digits after 0b");
    if (*q != '1') return cfail(CCS_ERR_INVALID, "binary prefix must start with 1");
    u128 v = 0;
    int n = 0;
    for (; *q == '0' || *q == '1'; ++q) {
        if (n >= 127) return cfail(CCS_ERR_INVALID, "prefix too long (max 126 bits)");
        v = (v << 1) | (u128)(*q - '0');
        ++n;
    }
    if (*q) return cfail(CCS_ERR_INVALID, std::string("invalid character '") + *q + "' in binary prefix");
    *value = un(v);
    *bits = n;
    return CCS_OK;
}

static int campaign_init(ccs_campaign* c) {
    const ccs_campaign_params& p = c->p;
    if (p.batch_tiles < 1 || p.batch_tiles > 12500000)
        return cfail(CCS_ERR_INVALID, "batch_tiles out of range [1, 12500000]");
    if (p.walk != CCS_WALK_INHERIT && p.walk != CCS_WALK_SEQUENTIAL && p.walk != CCS_WALK_RANDOM)
        return cfail(CCS_ERR_INVALID, "walk must be CCS_WALK_INHERIT, _SEQUENTIAL or _RANDOM");
    if (p.walk != CCS_WALK_INHERIT && c->prefix == "none")
        return cfail(CCS_ERR_INVALID, "--prefix-mode requires --prefix");
    if (c->tmin >= c->tmax) return cfail(CCS_ERR_INVALID, "empty tile range");
    // mode: --prefix-mode over --start over the default (random)
    if (c->prefix != "none" && p.walk != CCS_WALK_INHERIT) c->random_mode = p.walk == CCS_WALK_RANDOM;
    else c->random_mode = !p.has_start;
    c->cur = p.has_start ? mk(p.start) : c->tmin;
    if (p.has_start && (c->cur < c->tmin || c->cur >= c->tmax) && c->prefix != "none")
        return cfail(CCS_ERR_INVALID, "--start is outside the prefix tile range");
    c->seed = p.has_seed ? p.seed : entropy_seed();
    c->rng = c->seed ? c->seed : 1;   // rng_seed (v13.cu:711-713)
    return CCS_OK;
}

extern "C" int ccs_campaign_create_range(ccs_u128 tile_min, ccs_u128 tile_max, const ccs_campaign_params* p,
                                         ccs_campaign** out) {
    if (!p || !out) return cfail(CCS_ERR_INVALID, "null argument");
    *out = nullptr;
    ccs_campaign* c = new ccs_campaign();
    c->p = *p;
    c->prefix = p->prefix && *p->prefix ? p->prefix : "none";
    c->p.prefix = nullptr;            // the string is owned by the campaign from here
    c->tmin = mk(tile_min);
    c->tmax = mk(tile_max);
    const int rc = campaign_init(c);
    if (rc) { delete c; return rc; }
    *out = c;
    return CCS_OK;
}

extern "C" int ccs_campaign_create(const ccs_shape* shape, const ccs_campaign_params* p, ccs_campaign** out) {
    if (!shape || !p || !out) return cfail(CCS_ERR_INVALID, "null argument");
    ccs_u128 pv{0, 0};
    int32_t pb = 0;
    if (p->prefix && *p->prefix) {
        const int rc = ccs_parse_prefix(p->prefix, &pv, &pb);
        if (rc) return rc;
    } else if (p->lanes) {
        return cfail(CCS_ERR_INVALID, "--prefix-lanes requires --prefix");
    }
    ccs_u128 lo, hi;
    const int rc = ccs_tile_range(shape, p->bits, pv, pb, p->lanes, p->lane_id, &lo, &hi);
    if (rc) return rc;
    return ccs_campaign_create_range(lo, hi, p, out);
}

extern "C" void ccs_campaign_free(ccs_campaign* c) { delete c; }

extern "C" int ccs_campaign_next(ccs_campaign* c, ccs_u128* tile_start, uint32_t* tile_count) {
    if (!c || !tile_start || !tile_count) return cfail(CCS_ERR_INVALID, "null argument");
    *tile_count = 0;
    if (c->done) return CCS_DONE;
    if (c->p.max_tiles && c->handed * (uint64_t)c->p.batch_tiles >= c->p.max_tiles) {   // --tiles
        c->done = true;
        return CCS_DONE;
    }
    const u128 batch = c->p.batch_tiles;
    for (int pass = 0; pass < 2; ++pass) {
        if (c->random_mode) {
            // the reference also folds clock nanoseconds into the state per batch; a given
            // seed here names a reproducible walk
            if (!c->p.has_seed) {
                timespec ts;
                clock_gettime(CLOCK_MONOTONIC, &ts);
                c->rng ^= (uint64_t)ts.tv_nsec;
            }
            const u128 total = c->tmax - c->tmin;
            const u128 range = total > batch ? total - batch : 1;
            const uint64_t hi = xorshift64(c->rng), lo = xorshift64(c->rng);
            c->cur = c->tmin + (((u128)hi << 64) | lo) % range;
        }
        u128 n = batch;
        if (c->cur + n > c->tmax) n = c->cur < c->tmax ? c->tmax - c->cur : 0;
        if (n) {
            *tile_start = un(c->cur);
            *tile_count = (uint32_t)n;
            c->outstanding.push_back({c->cur, (uint32_t)n});
            c->handed++;
            if (!c->random_mode) c->cur += n;           // next sequential tile
            return CCS_OK;
        }
        if (!c->random_mode && c->prefix != "none") {   // sequential prefix sweep complete
            c->done = true;
            return CCS_DONE;
        }
        c->cur = c->tmin;                               // sequential: wrap to the range start
    }
    return cfail(CCS_ERR_INVALID, "empty tile range");
}

extern "C" int ccs_campaign_record(ccs_campaign* c, ccs_u128 tile_start, uint32_t tile_count, uint64_t candidates,
                                   uint64_t survivors, const ccs_chain_result* res, uint32_t n_res) {
    if (!c || (n_res && !res)) return cfail(CCS_ERR_INVALID, "null argument");
    if (c->outstanding.empty() || c->outstanding.front().first != mk(tile_start) ||
        c->outstanding.front().second != tile_count)
        return cfail(CCS_ERR_INVALID, "batches must be recorded in the order next() handed them out");
    c->outstanding.pop_front();
    c->batches++;
    c->total_tiles += tile_count;
    c->total_candidates += candidates;
    c->total_survivors += survivors;
    return ccs_confirm_stats_add(res, n_res, &c->chains);
}

extern "C" int ccs_campaign_get_state(const ccs_campaign* c, ccs_campaign_state* out) {
    if (!c || !out) return cfail(CCS_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->tile_min = un(c->tmin);
    out->tile_max = un(c->tmax);
    out->current = un(c->cur);
    out->seed = c->seed;
    out->batches = c->batches;
    out->total_tiles = c->total_tiles;
    out->total_candidates = c->total_candidates;
    out->total_survivors = c->total_survivors;
    out->random_mode = c->random_mode ? 1 : 0;
    out->done = c->done ? 1 : 0;
    out->chains = c->chains;
    return CCS_OK;
}

extern "C" int ccs_campaign_save(const ccs_campaign* c, const char* path) {
    if (!c || !path) return cfail(CCS_ERR_INVALID, "null argument");
    const std::string tmp = std::string(path) + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "w");
    if (!fp) return cfail(CCS_ERR_INVALID, std::string("cannot write ") + tmp + ": " + std::strerror(errno));
    std::fprintf(fp, "CC18_CUDA_v10\n");
    std::fprintf(fp, "bits=%d\n", c->p.bits);
    std::fprintf(fp, "target=%d\n", c->p.target);
    std::fprintf(fp, "depth=%d\n", c->p.depth);
    std::fprintf(fp, "coverage=full\n");
    std::fprintf(fp, "prefix=%s\n", c->prefix.c_str());
    std::fprintf(fp, "prefix_lanes=%u\n", c->p.lanes);
    std::fprintf(fp, "prefix_lane_id=%u\n", c->p.lanes ? c->p.lane_id : 0u);
    // the first tile whose results are not recorded yet: batches in flight are redone
    const u128 resume_at = c->outstanding.empty() ? c->cur : c->outstanding.front().first;
    std::fprintf(fp, "tile_hi=%016llx\n", (unsigned long long)(uint64_t)(resume_at >> 64));
    std::fprintf(fp, "tile_lo=%016llx\n", (unsigned long long)(uint64_t)resume_at);
    std::fprintf(fp, "total_tiles=%llu\n", (unsigned long long)c->total_tiles);
    std::fprintf(fp, "total_primes=%llu\n", (unsigned long long)c->chains.primes);
    std::fprintf(fp, "total_chains=%llu\n", (unsigned long long)c->chains.chains);
    std::fprintf(fp, "best_chain=%u\n", c->chains.best_chain);
    const bool ok = std::fflush(fp) == 0 && fsync(fileno(fp)) == 0;
    std::fclose(fp);
    if (!ok || std::rename(tmp.c_str(), path) != 0)    // atomic replace
        return cfail(CCS_ERR_INVALID, std::string("cannot replace ") + path + ": " + std::strerror(errno));
    return CCS_OK;
}

extern "C" int ccs_campaign_resume(ccs_campaign* c, const char* path) {
    if (!c || !path) return cfail(CCS_ERR_INVALID, "null argument");
    // the file: a format line, then key=value lines (unknown keys are ignored)
    std::ifstream in(path);
    if (!in) return cfail(CCS_ERR_INVALID, std::string("checkpoint ") + path + ": cannot be read");
    std::string line;
    if (!std::getline(in, line) || line.compare(0, 13, "CC18_CUDA_v10") != 0)
        return cfail(CCS_ERR_INVALID, std::string("checkpoint ") + path + ": not a CC18_CUDA_v10 file");
    std::map<std::string, std::string> kv;
    while (std::getline(in, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string val = line.substr(eq + 1);
        while (!val.empty() && (val.back() == '\r' || val.back() == ' ')) val.pop_back();
        kv[line.substr(0, eq)] = val;
    }
    auto num = [&](const char* key, int base) -> unsigned long long {
        auto it = kv.find(key);
        return it == kv.end() ? 0ull : std::strtoull(it->second.c_str(), nullptr, base);
    };
    const int bits = (int)num("bits", 10), target = (int)num("target", 10), depth = (int)num("depth", 10);
    const unsigned lanes = (unsigned)num("prefix_lanes", 10), lane_id = (unsigned)num("prefix_lane_id", 10);
    const unsigned best = (unsigned)num("best_chain", 10);
    const unsigned long long thi = num("tile_hi", 16), tlo = num("tile_lo", 16), tiles = num("total_tiles", 10);
    const unsigned long long primes = num("total_primes", 10), chains = num("total_chains", 10);
    const std::string prefix = kv.count("prefix") ? kv["prefix"] : std::string();
    auto mismatch = [&](const char* what, const std::string& a, const std::string& b) {
        return cfail(CCS_ERR_INVALID, std::string("checkpoint was written with ") + what + "=" + a + ", this run has " + b);
    };
    if (bits != c->p.bits) return mismatch("bits", std::to_string(bits), std::to_string(c->p.bits));
    if (target != c->p.target) return mismatch("target", std::to_string(target), std::to_string(c->p.target));
    if (depth != c->p.depth) return mismatch("depth", std::to_string(depth), std::to_string(c->p.depth));
    if (c->prefix != prefix) return mismatch("prefix", prefix, c->prefix);
    if (lanes != c->p.lanes) return mismatch("prefix_lanes", std::to_string(lanes), std::to_string(c->p.lanes));
    const unsigned cur_lane = c->p.lanes ? c->p.lane_id : 0u;
    if (lane_id != cur_lane) return mismatch("prefix_lane_id", std::to_string(lane_id), std::to_string(cur_lane));
    const u128 tile = ((u128)thi << 64) | tlo;
    if (tile < c->tmin || tile >= c->tmax)
        return cfail(CCS_ERR_INVALID, "checkpoint tile is outside the current tile range");
    if (!c->outstanding.empty()) return cfail(CCS_ERR_INVALID, "resume before the first batch only");
    c->cur = tile;
    c->total_tiles = tiles;
    c->chains.primes = primes;
    c->chains.chains = chains;
    c->chains.best_chain = best;
    c->random_mode = false;   // resume is always sequential from the checkpoint position
    c->done = false;
    return CCS_OK;
}
