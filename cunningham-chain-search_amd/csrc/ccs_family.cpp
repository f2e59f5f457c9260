// Fingerprint families: the host side of the v16 Q-iteration front end.
//
// A family is p_0 = Q*D*2^E - 1 for an immune-factor product D (a pool entry).  This file
// parses pool lists and pool files into entries (v16...cu:556-738, :828-880), computes a
// family's Q window for a bit band (v16...cu:4405-4468) and flags the first occurrence of
// each p_0 (the dedup of v16...cu:2478-2541).  The family's GPU sieve is an ordinary job
// set on a family shape (ccs_family_shape_build).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/ccs.h"
#include "ccs_internal.hpp"

namespace {

typedef unsigned __int128 u128;

bool is_prime32(uint32_t q) {
    if (q < 2) return false;
    for (uint32_t d = 2; (uint64_t)d * d <= q; ++d)
        if (q % d == 0) return false;
    return true;
}

// the primes up to 503, the table v16 draws primorials and auto variants from
const std::vector<uint32_t>& small_primes() {
    static const std::vector<uint32_t> t = [] {
        std::vector<uint32_t> v;
        for (uint32_t q = 2; q <= 503; ++q)
            if (is_prime32(q)) v.push_back(q);
        return v;
    }();
    return t;
}

int bad(const std::string& msg) { return ccs::set_error(CCS_ERR_INVALID, msg.c_str()); }

void set_name(ccs_pool_entry& e, const std::string& name) {
    std::memset(e.name, 0, sizeof(e.name));
    std::strncpy(e.name, name.c_str(), sizeof(e.name) - 1);
}

// D from the entry's primes, with the u128 overflow check
int finish_entry(ccs_pool_entry& e, const std::string& spec) {
    u128 D = 1;
    for (uint32_t i = 0; i < e.n_primes; ++i) {
        const uint32_t q = e.primes[i];
        if (!is_prime32(q)) return bad(std::to_string(q) + " in '" + spec + "' is not prime");
        const u128 nd = D * (u128)q;
        if (nd / q != D) return bad("D overflows u128 in '" + spec + "'");
        D = nd;
    }
    e.D.lo = (uint64_t)D;
    e.D.hi = (uint64_t)(D >> 64);
    return CCS_OK;
}

// one spec: "Z#[/X/Y...]" or "P[^k]*Q[^k]*..."
int parse_spec(const std::string& spec, ccs_pool_entry& e) {
    std::memset(&e, 0, sizeof(e));
    set_name(e, spec);
    const char* s = spec.c_str();
    const char* hash = std::strchr(s, '#');
    if (hash) {
        const long Z = std::strtol(s, nullptr, 10);
        if (Z < 2 || Z > 500) return bad("bad primorial '" + spec + "'");
        std::vector<uint32_t> ps;
        for (uint32_t q : small_primes()) {
            if ((long)q > Z) break;
            if (ps.size() < 64) ps.push_back(q);
        }
        const char* p = hash + 1;
        while (*p) {
            while (*p == '/' || *p == ' ') ++p;
            if (!*p) break;
            char* end;
            const unsigned long omit = std::strtoul(p, &end, 10);
            if (end == p) return bad("malformed omit in '" + spec + "'");
            auto it = std::find(ps.begin(), ps.end(), (uint32_t)omit);
            if (it == ps.end()) return bad("prime " + std::to_string(omit) + " not in " + std::to_string(Z) + "# ('" + spec + "')");
            ps.erase(it);
            p = end;
        }
        for (uint32_t q : ps) e.primes[e.n_primes++] = q;
    } else {
        const char* p = s;
        while (*p) {
            while (*p == ' ' || *p == '*') ++p;
            if (!*p) break;
            char* end;
            const unsigned long pr = std::strtoul(p, &end, 10);
            if (end == p) return bad("malformed prime in '" + spec + "'");
            long k = 1;
            if (*end == '^') {
                const char* q = end + 1;
                k = std::strtol(q, &end, 10);
                if (end == q || k < 1 || k > 16) return bad("bad exponent in '" + spec + "'");
            }
            for (long i = 0; i < k; ++i) {
                if (e.n_primes >= 64) return bad("too many primes in '" + spec + "'");
                e.primes[e.n_primes++] = (uint32_t)pr;
            }
            p = end;
        }
    }
    return finish_entry(e, spec);
}

struct Sink {
    ccs_pool_entry* out;
    uint32_t cap, n;
    void push(const ccs_pool_entry& e) {
        if (n < cap) out[n] = e;
        ++n;
    }
};

// "auto:Z#:K" / "auto:Z#:K1-K2": every way of dropping K of the primes <= Z outside keep[],
// in lexicographic order of the dropped primes (expand_auto_pool_one + emit_auto_combo)
int expand_auto(const std::string& spec, const uint32_t* keep, uint32_t n_keep, Sink& sink) {
    const char* p = spec.c_str() + 5;
    char* end;
    const long Z = std::strtol(p, &end, 10);
    if (Z < 2 || Z > 500 || end == p) return bad("bad auto spec '" + spec + "' (expected auto:Z#:K)");
    if (*end != '#') return bad("bad auto spec '" + spec + "' (expected # after Z)");
    ++end;
    if (*end != ':') return bad("bad auto spec '" + spec + "' (expected : after Z#)");
    ++end;
    const long kmin = std::strtol(end, &end, 10);
    long kmax = kmin;
    if (*end == '-') {
        ++end;
        kmax = std::strtol(end, &end, 10);
    }
    if (kmin < 0 || kmax < kmin) return bad("bad K range in '" + spec + "'");
    std::vector<uint32_t> all, drop;
    for (uint32_t q : small_primes()) {
        if ((long)q > Z) break;
        all.push_back(q);
        bool must = false;
        for (uint32_t i = 0; i < n_keep; ++i) must |= keep[i] == q;
        if (!must && drop.size() < 64) drop.push_back(q);
    }
    if (kmax > (long)drop.size()) kmax = (long)drop.size();
    std::vector<uint32_t> chosen;
    int rc = CCS_OK;
    // recursive combination walk, smallest dropped primes first
    std::function<void(size_t, long)> walk = [&](size_t start, long k_left) {
        if (rc) return;
        if (k_left == 0) {
            ccs_pool_entry e;
            std::memset(&e, 0, sizeof(e));
            std::string name = std::to_string(Z) + "#";
            for (uint32_t q : chosen) name += "/" + std::to_string(q);
            set_name(e, name);
            for (uint32_t q : all) {
                if (std::find(chosen.begin(), chosen.end(), q) != chosen.end()) continue;
                if (e.n_primes < 64) e.primes[e.n_primes++] = q;
            }
            rc = finish_entry(e, name);
            if (!rc) sink.push(e);
            return;
        }
        for (size_t i = start; i + (size_t)k_left <= drop.size(); ++i) {
            chosen.push_back(drop[i]);
            walk(i + 1, k_left - 1);
            chosen.pop_back();
            if (rc) return;
        }
    };
    for (long k = kmin; k <= kmax && !rc; ++k) walk(0, k);
    return rc;
}

int parse_list(const std::string& list, const uint32_t* keep, uint32_t n_keep, Sink& sink) {
    size_t pos = 0;
    while (pos <= list.size()) {
        size_t comma = list.find(',', pos);
        if (comma == std::string::npos) comma = list.size();
        std::string tok = list.substr(pos, comma - pos);
        pos = comma + 1;
        size_t b = 0;
        while (b < tok.size() && tok[b] == ' ') ++b;
        tok = tok.substr(b);
        while (!tok.empty() && (tok.back() == ' ' || tok.back() == '\n')) tok.pop_back();
        if (tok.empty()) continue;
        int rc;
        if (tok.compare(0, 5, "auto:") == 0) {
            rc = expand_auto(tok, keep, n_keep, sink);
        } else {
            ccs_pool_entry e;
            rc = parse_spec(tok, e);
            if (!rc) sink.push(e);
        }
        if (rc) return rc;
    }
    return CCS_OK;
}

}  // namespace

extern "C" int ccs_pool_parse(const char* list, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                              uint32_t cap, uint32_t* n_out) {
    if (!list || !n_out || (cap && !out) || (n_keep && !keep)) return bad("null argument");
    Sink sink{out, cap, 0};
    const int rc = parse_list(list, keep, n_keep, sink);
    *n_out = sink.n;
    if (rc) return rc;
    return sink.n > cap ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_pool_load(const char* path, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                             uint32_t cap, uint32_t* n_out) {
    if (!path || !n_out || (cap && !out) || (n_keep && !keep)) return bad("null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return bad(std::string("cannot open pool file '") + path + "'");
    Sink sink{out, cap, 0};
    char line[1024];
    int lineno = 0, rc = CCS_OK;
    while (!rc && std::fgets(line, sizeof(line), fp)) {
        ++lineno;
        char* hash = std::strchr(line, '#');
        if (hash) *hash = '\0';
        size_t L = std::strlen(line);
        while (L > 0 && (line[L - 1] == '\n' || line[L - 1] == '\r' || line[L - 1] == ' ' || line[L - 1] == '\t'))
            line[--L] = '\0';
        char* tok = line;
        while (*tok == ' ' || *tok == '\t') ++tok;
        if (!*tok) continue;
        rc = parse_list(tok, keep, n_keep, sink);
        if (rc) {
            const std::string msg = std::string(ccs_last_error()) + " (" + path + " line " + std::to_string(lineno) + ")";
            rc = bad(msg);
        }
    }
    std::fclose(fp);
    *n_out = sink.n;
    if (rc) return rc;
    return sink.n > cap ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_family_window(ccs_u128 Dv, int32_t E, int32_t bits_lo, int32_t bits_hi, ccs_u128* q_min,
                                 ccs_u128* q_max) {
    if (!q_min || !q_max) return bad("null argument");
    if (bits_lo < 2 || bits_hi < bits_lo || bits_hi > 127) return bad("bits must satisfy 2 <= bits_lo <= bits_hi <= 127");
    if (E < 0 || E > 126) return bad("exp_start must be in 0..126");
    const u128 D = ((u128)Dv.hi << 64) | Dv.lo;
    if (D == 0) return bad("D must be >= 1");
    if (D >> (127 - E)) return bad("D*2^E must be below 2^127");
    const u128 DE = D << E;
    const u128 lower = (u128)1 << (bits_lo - 1), upper_m1 = ((u128)1 << bits_hi) - 1;
    const u128 qmin = (lower + DE - 1) / DE, qmax = upper_m1 / DE;
    q_min->lo = (uint64_t)qmin;
    q_min->hi = (uint64_t)(qmin >> 64);
    q_max->lo = (uint64_t)qmax;
    q_max->hi = (uint64_t)(qmax >> 64);
    return qmin > qmax ? CCS_DONE : CCS_OK;
}

extern "C" int ccs_dedup_u128(const ccs_u128* values, uint32_t count, uint8_t* first) {
    if (count && (!values || !first)) return bad("null argument");
    std::vector<uint32_t> idx(count);
    for (uint32_t i = 0; i < count; ++i) idx[i] = i;
    auto key = [&](uint32_t i) { return ((u128)values[i].hi << 64) | values[i].lo; };
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    for (uint32_t k = 0; k < count; ++k)
        first[idx[k]] = (k == 0 || key(idx[k]) != key(idx[k - 1])) ? 1 : 0;
    return CCS_OK;
}
