// Host confirmation engine: the semantics of the reference's CPU prover pool
// (prover_worker, v13.cu:1794-1924, merged by prove_survivors_mt, :1967-2068) on host
// threads, with the library's own u128 Baillie-PSW (the arithmetic the GPU prove kernels
// use, ccs_arith.hpp) in place of GMP's mpz_probab_prime_p(n, 1).
//
// Per value: bit-size gate, trial division by the 62 primes 67..409, primality of n, the
// backward walk to the true root while (r-1)/2 >= 2 is prime, the forward follow from the
// root while 2c+1 is prime (length <= 64).  Values at or above 2^127 are flagged
// CCS_F_OVERFLOW and not followed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/ccs.h"
#include "ccs_arith.hpp"

using namespace ccs;

namespace {

constexpr int kMaxChain = 64;   // MAX_CHAIN (v13.cu:174)

struct Tables {
    PrimeMod small[18];   // 2..61
    PrimeMod trial[62];   // 67..409 (CHAIN_SCREEN_PRIMES, v13.cu:180-199)
};

PrimeMod small_mod(u32 q) {
    PrimeMod pm;
    std::memset(&pm, 0, sizeof(pm));
    pm.q = q;
    pm.c32 = (u32)(((unsigned __int128)1 << 32) % q);
    pm.c64 = (u32)(((unsigned __int128)1 << 64) % q);
    pm.c96 = (u32)(((unsigned __int128)1 << 96) % q);
    pm.rcp = 1.0 / (double)q;
    return pm;
}

const Tables& tables() {
    static const Tables t = [] {
        Tables x;
        int ns = 0, nt = 0;
        for (u32 q = 2; q <= 409; ++q) {
            bool prime = true;
            for (u32 d = 2; d * d <= q; ++d)
                if (q % d == 0) { prime = false; break; }
            if (!prime) continue;
            if (q <= 61) x.small[ns++] = small_mod(q);
            else if (q >= 67) x.trial[nt++] = small_mod(q);
        }
        return x;
    }();
    return t;
}

struct Tester {
    int kind;
    int nb;
    u32 bases[16];
    bool ok = true;
    bool operator()(U2 n) {
        return kind == CCS_TEST_BPSW ? is_prime_bpsw(n, tables().small, ok) : is_prime_mr(n, bases, nb);
    }
};

ccs_chain_result confirm_one(U2 n, uint32_t idx, int bits, Tester& t) {
    ccs_chain_result r;
    std::memset(&r, 0, sizeof(r));
    r.survivor_idx = idx;
    t.ok = true;
    do {
        if (bits && bitlen(n) != bits) break;
        r.flags |= CCS_F_BITS_OK;
        bool div = false;
        for (int i = 0; i < 62 && !div; ++i) div = mod128(n.lo, n.hi, tables().trial[i]) == 0;
        if (div) break;
        r.flags |= CCS_F_TRIAL_OK;
        if (n.hi >> 63) { r.flags |= CCS_F_OVERFLOW; break; }
        if (!t(n)) break;
        r.flags |= CCS_F_PRIME;
        U2 root = n;
        uint32_t back = 0;
        U2 pred = shr1(sub(n, mk(1, 0)));
        if (ge(pred, mk(2, 0)) && t(pred)) {          // n is not a root: walk back
            r.flags |= CCS_F_NONROOT;
            root = pred;
            back = 1;
            for (int steps = 0; steps < kMaxChain; ++steps) {
                pred = shr1(sub(root, mk(1, 0)));
                if (!ge(pred, mk(2, 0)) || !t(pred)) break;
                root = pred;
                ++back;
            }
        }
        uint32_t len = 1;
        U2 cur = root;
        while (len < (uint32_t)kMaxChain) {            // forward follow from the true root
            if (cur.hi >> 62) { r.flags |= CCS_F_OVERFLOW; break; }
            const U2 nx = add(shl1(cur), mk(1, 0));
            if (!t(nx)) break;
            ++len;
            cur = nx;
        }
        r.chain_len = len;
        r.back_steps = back;
        r.root_lo = root.lo;
        r.root_hi = root.hi;
    } while (0);
    if (!t.ok) r.flags |= CCS_F_UNSUPPORTED;
    return r;
}

int confirm_fail(int rc, const char* what) { return ccs::set_error(rc, what); }

}  // namespace

extern "C" int ccs_confirm_stats_add(const ccs_chain_result* res, uint32_t count, ccs_confirm_stats* st) {
    if (!st || (count && !res)) return CCS_ERR_INVALID;
    for (uint32_t i = 0; i < count; ++i) {
        const ccs_chain_result& r = res[i];
        if (!(r.flags & CCS_F_PRIME)) continue;
        st->primes++;
        st->chains++;
        if (r.flags & CCS_F_NONROOT) st->non_roots++;
        if (r.chain_len <= 64) st->chain_dist[r.chain_len]++;
        if (r.chain_len > st->best_chain) st->best_chain = r.chain_len;
    }
    return CCS_OK;
}

extern "C" int ccs_confirm_values(const ccs_u128* h_n, uint32_t count, const ccs_prove_params* params,
                                  uint32_t nthreads, ccs_chain_result* h_out, ccs_confirm_stats* stats) {
    if (!params || (count && (!h_n || !h_out))) return confirm_fail(CCS_ERR_INVALID, "null argument");
    if (params->test_kind != CCS_TEST_BPSW && params->test_kind != CCS_TEST_MR)
        return confirm_fail(CCS_ERR_INVALID, "unknown test kind");
    if (params->test_kind == CCS_TEST_MR && (params->n_bases < 1 || params->n_bases > 16))
        return confirm_fail(CCS_ERR_INVALID, "MR needs 1..16 bases");
    if (params->bits < 0 || params->bits > 128) return confirm_fail(CCS_ERR_INVALID, "bits");
    unsigned hw = std::thread::hardware_concurrency();
    if (nthreads == 0) nthreads = hw ? std::min(hw, 16u) : 1u;
    if (nthreads > 256) nthreads = 256;
    if (count < nthreads) nthreads = count ? count : 1;
    (void)tables();
    auto work = [&](uint32_t t) {
        Tester tst;
        tst.kind = params->test_kind;
        tst.nb = params->n_bases;
        for (int i = 0; i < 16; ++i) tst.bases[i] = params->bases[i];
        const uint64_t beg = (uint64_t)count * t / nthreads, end = (uint64_t)count * (t + 1) / nthreads;
        for (uint64_t i = beg; i < end; ++i)
            h_out[i] = confirm_one(mk(h_n[i].lo, h_n[i].hi), (uint32_t)i, params->bits, tst);
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < nthreads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        ccs_confirm_stats_add(h_out, count, stats);
    }
    return CCS_OK;
}
