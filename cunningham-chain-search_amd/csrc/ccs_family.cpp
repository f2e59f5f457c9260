// Fingerprint families: the host side of the v16 Q-iteration front end.
//
// A family is p_0 = Q*D*2^E - 1 for an immune-factor product D (a pool entry).  This file
// reads pool lists and pool files into entries (the input format of v16's --seed-pool and
// pool files, v16...cu:556-738, :828-880, read here by its own grammar), computes a
// family's Q window for a bit band (v16...cu:4405-4468) and flags the first occurrence of
// each p_0 (the dedup of v16...cu:2478-2541).  The family's GPU sieve is an ordinary job
// set on a family shape (ccs_family_shape_build).
#include <algorithm>
#include <cstring>
#include <fstream>
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

constexpr uint64_t kMaxPrimorial = 500;   // largest Z of "Z#" and "auto:Z#:K"
constexpr uint64_t kMaxPower = 16;        // largest exponent of "P^k"

int reject(const std::string& msg) { return ccs::set_error(CCS_ERR_INVALID, msg.c_str()); }

// Pool-list grammar (blanks = spaces and tabs, allowed around every token):
//   list      = item { "," item }               (blank items are skipped)
//   item      = auto | primorial | product
//   primorial = Z "#" { "/" P }                 (primes <= Z, each P left out; 2 <= Z <= 500)
//   product   = factor { "*" factor }           factor = P [ "^" k ]   (1 <= k <= 16)
//   auto      = "auto:" Z "#:" k1 [ "-" k2 ]    (every way of leaving out k1..k2 primes <= Z
//                                                that are not in keep[], k2 clamped to their count)
// An item's D is the product of its primes (with multiplicity) and must fit in 128 bits.
class Reader {
public:
    explicit Reader(const std::string& text) : t_(text) {}
    void blanks() {
        while (i_ < t_.size() && (t_[i_] == ' ' || t_[i_] == '\t')) ++i_;
    }
    bool at_end() {
        blanks();
        return i_ >= t_.size();
    }
    bool accept(char c) {
        blanks();
        if (i_ < t_.size() && t_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    bool accept_word(const char* w) {
        blanks();
        const size_t n = std::strlen(w);
        if (t_.compare(i_, n, w) != 0) return false;
        i_ += n;
        return true;
    }
    // a decimal number (at most 18 digits); false when none starts here
    bool number(uint64_t& v) {
        blanks();
        size_t j = i_;
        v = 0;
        while (j < t_.size() && t_[j] >= '0' && t_[j] <= '9' && j - i_ < 18) v = v * 10 + (uint64_t)(t_[j++] - '0');
        if (j == i_ || (j < t_.size() && t_[j] >= '0' && t_[j] <= '9')) return false;
        i_ = j;
        return true;
    }
    bool starts_primorial() const {   // digits, blanks, then '#'
        size_t j = i_;
        while (j < t_.size() && t_[j] >= '0' && t_[j] <= '9') ++j;
        while (j < t_.size() && (t_[j] == ' ' || t_[j] == '\t')) ++j;
        return j < t_.size() && t_[j] == '#';
    }
    const std::string& text() const { return t_; }

private:
    const std::string& t_;
    size_t i_ = 0;
};

// An entry under construction: prime factors with multiplicity and their product.
struct Factors {
    uint32_t p[64];
    uint32_t n = 0;
    u128 D = 1;
    int add(uint64_t q, const std::string& item) {
        if (q > 0xffffffffull || !is_prime32((uint32_t)q))
            return reject("pool item '" + item + "': factor " + std::to_string(q) + " is not a prime");
        if (n == 64) return reject("pool item '" + item + "': more than 64 prime factors");
        if (D > ~(u128)0 / q) return reject("pool item '" + item + "': D does not fit in 128 bits");
        p[n++] = (uint32_t)q;
        D *= q;
        return CCS_OK;
    }
    void store(const std::string& name, ccs_pool_entry& e) const {
        std::memset(&e, 0, sizeof(e));
        std::strncpy(e.name, name.c_str(), sizeof(e.name) - 1);
        std::memcpy(e.primes, p, n * sizeof(uint32_t));
        e.n_primes = n;
        e.D.lo = (uint64_t)D;
        e.D.hi = (uint64_t)(D >> 64);
    }
};

std::vector<uint32_t> primes_upto(uint64_t z) {
    std::vector<uint32_t> v;
    for (uint32_t q = 2; q <= z; ++q)
        if (is_prime32(q)) v.push_back(q);
    return v;
}

struct Out {
    ccs_pool_entry* dst;
    uint32_t cap, n;
    void put(const ccs_pool_entry& e) {
        if (n < cap) dst[n] = e;
        ++n;
    }
};

int read_primorial(Reader& r, const std::string& item, Out& out) {
    uint64_t z = 0;
    if (!r.number(z) || !r.accept('#') || z < 2 || z > kMaxPrimorial)
        return reject("pool item '" + item + "': a primorial needs 2 <= Z <= 500 before '#'");
    std::vector<uint32_t> keep = primes_upto(z);
    while (r.accept('/')) {
        uint64_t q = 0;
        if (!r.number(q)) return reject("pool item '" + item + "': expected a prime after '/'");
        auto it = std::find(keep.begin(), keep.end(), q);
        if (it == keep.end())
            return reject("pool item '" + item + "': " + std::to_string(q) + " is not one of the primes <= " +
                          std::to_string(z) + " still in the product");
        keep.erase(it);
    }
    if (!r.at_end()) return reject("pool item '" + item + "': unexpected text after the primorial");
    Factors f;
    for (uint32_t q : keep)
        if (int rc = f.add(q, item)) return rc;
    ccs_pool_entry e;
    f.store(item, e);
    out.put(e);
    return CCS_OK;
}

int read_product(Reader& r, const std::string& item, Out& out) {
    Factors f;
    do {
        uint64_t q = 0, k = 1;
        if (!r.number(q)) return reject("pool item '" + item + "': expected a prime factor");
        if (r.accept('^') && (!r.number(k) || k < 1 || k > kMaxPower))
            return reject("pool item '" + item + "': an exponent must be 1..16");
        for (uint64_t m = 0; m < k; ++m)
            if (int rc = f.add(q, item)) return rc;
    } while (r.accept('*'));
    if (!r.at_end()) return reject("pool item '" + item + "': expected '*' between factors");
    ccs_pool_entry e;
    f.store(item, e);
    out.put(e);
    return CCS_OK;
}

// every subset of `pool` of size k, in lexicographic order of the positions, left out of
// `all`; the entries are named "Z#/q1/q2..."
int emit_subsets(const std::vector<uint32_t>& all, const std::vector<uint32_t>& pool, size_t k, uint64_t z,
                 Out& out) {
    std::vector<size_t> pick(k);
    for (size_t i = 0; i < k; ++i) pick[i] = i;
    for (;;) {
        std::string name = std::to_string(z) + "#";
        std::vector<uint32_t> left;
        for (size_t i : pick) {
            left.push_back(pool[i]);
            name += "/" + std::to_string(pool[i]);
        }
        Factors f;
        for (uint32_t q : all)
            if (std::find(left.begin(), left.end(), q) == left.end())
                if (int rc = f.add(q, name)) return rc;
        ccs_pool_entry e;
        f.store(name, e);
        out.put(e);
        // next subset: advance the last position that can still move right
        size_t i = k;
        while (i > 0 && pick[i - 1] == pool.size() - k + (i - 1)) --i;
        if (i == 0) return CCS_OK;
        ++pick[i - 1];
        for (size_t m = i; m < k; ++m) pick[m] = pick[m - 1] + 1;
    }
}

int read_auto(Reader& r, const std::string& item, const uint32_t* keep, uint32_t n_keep, Out& out) {
    uint64_t z = 0, k1 = 0, k2 = 0;
    if (!r.number(z) || z < 2 || z > kMaxPrimorial || !r.accept('#') || !r.accept(':'))
        return reject("pool item '" + item + "': the form is auto:Z#:k or auto:Z#:k1-k2 with 2 <= Z <= 500");
    if (!r.number(k1)) return reject("pool item '" + item + "': expected the number of primes to leave out");
    k2 = k1;
    if (r.accept('-') && !r.number(k2)) return reject("pool item '" + item + "': expected k2 after '-'");
    if (k2 < k1) return reject("pool item '" + item + "': k2 is below k1");
    if (!r.at_end()) return reject("pool item '" + item + "': unexpected text after the auto range");
    const std::vector<uint32_t> all = primes_upto(z);
    std::vector<uint32_t> pool;   // the primes that may be left out
    for (uint32_t q : all)
        if (std::find(keep, keep + n_keep, q) == keep + n_keep) pool.push_back(q);
    k2 = std::min<uint64_t>(k2, pool.size());
    for (uint64_t k = k1; k <= k2; ++k)
        if (int rc = emit_subsets(all, pool, (size_t)k, z, out)) return rc;
    return CCS_OK;
}

int read_list(const std::string& text, const uint32_t* keep, uint32_t n_keep, Out& out) {
    size_t from = 0;
    for (;;) {
        const size_t comma = std::min(text.find(',', from), text.size());
        size_t a = from, b = comma;
        while (a < b && (text[a] == ' ' || text[a] == '\t')) ++a;
        while (b > a && (text[b - 1] == ' ' || text[b - 1] == '\t' || text[b - 1] == '\r' || text[b - 1] == '\n')) --b;
        if (b > a) {
            const std::string item = text.substr(a, b - a);
            Reader r(item);
            int rc;
            if (r.accept_word("auto:")) rc = read_auto(r, item, keep, n_keep, out);
            else if (r.starts_primorial()) rc = read_primorial(r, item, out);
            else rc = read_product(r, item, out);
            if (rc) return rc;
        }
        if (comma >= text.size()) return CCS_OK;
        from = comma + 1;
    }
}

}  // namespace

extern "C" int ccs_pool_parse(const char* list, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                              uint32_t cap, uint32_t* n_out) {
    if (!list || !n_out || (cap && !out) || (n_keep && !keep)) return reject("null argument");
    Out o{out, cap, 0};
    const int rc = read_list(list, keep, n_keep, o);
    *n_out = o.n;
    if (rc) return rc;
    return o.n > cap ? CCS_OVERFLOW : CCS_OK;
}

// A pool file holds one list per line; from a line's first '#' on is a comment, so a file
// lists products only (the primorial form reads as its number).
extern "C" int ccs_pool_load(const char* path, const uint32_t* keep, uint32_t n_keep, ccs_pool_entry* out,
                             uint32_t cap, uint32_t* n_out) {
    if (!path || !n_out || (cap && !out) || (n_keep && !keep)) return reject("null argument");
    std::ifstream in(path);
    if (!in) return reject(std::string("pool file ") + path + ": cannot be read");
    Out o{out, cap, 0};
    std::string line;
    for (uint64_t no = 1; std::getline(in, line); ++no) {
        const size_t c = line.find('#');
        if (c != std::string::npos) line.resize(c);
        if (read_list(line, keep, n_keep, o) != CCS_OK) {
            *n_out = o.n;
            return reject(std::string(path) + ":" + std::to_string(no) + ": " + ccs_last_error());
        }
    }
    *n_out = o.n;
    return o.n > cap ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_family_window(ccs_u128 Dv, int32_t E, int32_t bits_lo, int32_t bits_hi, ccs_u128* q_min,
                                 ccs_u128* q_max) {
    if (!q_min || !q_max) return reject("null argument");
    if (bits_lo < 2 || bits_hi < bits_lo || bits_hi > 127) return reject("bits must satisfy 2 <= bits_lo <= bits_hi <= 127");
    if (E < 0 || E > 126) return reject("exp_start must be in 0..126");
    const u128 D = ((u128)Dv.hi << 64) | Dv.lo;
    if (D == 0) return reject("D must be >= 1");
    if (D >> (127 - E)) return reject("D*2^E must be below 2^127");
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
    if (count && (!values || !first)) return reject("null argument");
    std::vector<uint32_t> idx(count);
    for (uint32_t i = 0; i < count; ++i) idx[i] = i;
    auto key = [&](uint32_t i) { return ((u128)values[i].hi << 64) | values[i].lo; };
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    for (uint32_t k = 0; k < count; ++k)
        first[idx[k]] = (k == 0 || key(idx[k]) != key(idx[k - 1])) ? 1 : 0;
    return CCS_OK;
}
