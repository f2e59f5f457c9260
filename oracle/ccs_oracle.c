/*
 * ccs_oracle.c — CPU restatement of the reference's Cunningham-chain (first kind)
 * candidate pipeline.  TEST INFRASTRUCTURE ONLY: this file is the CPU model that the
 * tests, the smoke check and the CPU-baseline leg of the benchmark compare the HIP
 * product against.  Nothing in the product links, loads or calls it.
 *
 * Every function names the reference file:line whose behaviour it restates.
 * Reference = nmicic/cunningham-chain-search:
 *   v13  = src/cuda/cc18_filter_cuda_CpC_v13.cu
 *   v33  = src/cpu/cc_gmp_v33_03.c
 *   v34  = src/cpu/cc_gmp_v34_bit-vector_10.c
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   forbidden residues ........ closed form (q | p_i with p_i = 2^i(n+1)-1, brute force),
 *                               v33 unit tests (immune residue q-1, residue 0 at pos 0)
 *   bases / wheel / classes ... reference counts (36 bases, 3432 wheel/base, 123552 classes
 *                               mod 31# at target 18), brute-force admissibility on small shapes,
 *                               v13 self-tests 15-19 (mod-6 bucket, CC17 base, q7 roots)
 *   filter predicate .......... brute-force p_i mod q over every position, and the reference
 *                               root sample (data/sample/cc10plus_roots_sample_1000.txt)
 *   class-form filter ......... orc_filter_class against orc_first_kill_brute over windows of
 *                               classes mod 6*31# and 6*37# (u and n crossing 2^64)
 *   primality / chains ........ sympy.isprime (independent library), known chains CC5..CC18,
 *                               Carmichael numbers and strong pseudoprimes to bases 2..13
 *   BPSW vs GMP on BPSW pseudoprimes: parity unpinned (none is known to exist).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t u8;

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* small number theory                                                        */
/* ------------------------------------------------------------------------- */

/* extended Euclid inverse, restating v13.cu:754-769 (mod_inverse) */
static u64 orc_inv(u64 a, u64 m) {
    if (m == 1) return 0;
    int64_t t = 0, nt = 1;
    int64_t r = (int64_t)m, nr = (int64_t)(a % m);
    while (nr != 0) {
        int64_t qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += (int64_t)m;
    return (u64)t;
}

static u32 umod128(u128 n, u32 q) { return (u32)(n % q); }

/* F_q for positions i in [pos_lo, min(pos_hi, q-1)): residue (2^-i - 1) mod q,
 * deduplicated, in position order.  Restates v13.cu:771-787
 * (compute_forbidden_residues) and the position ranges of v13.cu:797-847
 * (L2/ext: 0..depth-1, line sieve: 1..depth-1). */
ORC_EXPORT int orc_forbidden(u32 q, int pos_lo, int pos_hi, u32 *out) {
    int cnt = 0;
    int hi = pos_hi;
    if (hi > (int)q - 1) hi = (int)q - 1;
    u64 p2 = 1;
    for (int i = 0; i < pos_lo; i++) p2 = (p2 * 2) % q;
    for (int i = pos_lo; i < hi; i++) {
        u32 r = (u32)((orc_inv(p2, q) + q - 1) % q);
        int dup = 0;
        for (int k = 0; k < cnt; k++) if (out[k] == r) { dup = 1; break; }
        if (!dup) out[cnt++] = r;
        p2 = (p2 * 2) % q;
    }
    return cnt;
}

ORC_EXPORT int orc_primes_between(u32 lo, u32 hi, u32 *out, int cap) {
    /* simple sieve, [lo, hi] inclusive */
    if (hi < 2) return 0;
    u8 *comp = (u8 *)calloc(hi + 1, 1);
    int n = 0;
    for (u64 i = 2; i <= hi; i++) {
        if (comp[i]) continue;
        if (i >= lo) { if (n < cap && out) out[n] = (u32)i; n++; }
        for (u64 j = i * i; j <= hi; j += i) comp[j] = 1;
    }
    free(comp);
    return n;
}

/* ------------------------------------------------------------------------- */
/* shape: CRT bases x wheel x mod-6 rule, and the filter prime list           */
/* ------------------------------------------------------------------------- */

typedef struct {
    int ncrt, nwheel, target, depth;
    u32 crt[16], wheel[8];
    u64 M, WP;
    int nbases;
    u64 *bases;          /* sorted admissible residues mod M            */
    u32 *wstart, *wcount;/* per base: range in wk[]                     */
    u32 *wk;             /* wheel offsets k in [0, WP), ascending       */
    u8  *wrho;           /* tile residue mod 6 that makes n = 5 (mod 6) */
    u64 nclasses;
    /* per (base, rho): list of k (ascending) — the reference's mod-6 bucket */
    u32 *bstart, *bcount, *bk;
    /* filter */
    int nfp;
    u32 *fp;             /* ascending filter primes                     */
    u32 pos0_max_q;      /* primes <= this test position 0               */
    u64 **fbits;         /* per filter prime: bitset of forbidden n mod q */
    u32 *WPMq;           /* (WP*M) mod q                                 */
    u32 *Mq;             /* M mod q                                      */
    u32 *basemod;        /* [nbases][nfp] base mod q                     */
    int nlut;            /* primes with a kM LUT                         */
    u16 *kMlut;          /* [WP][nlut] (k*M) mod q                       */
    void *bv;            /* cached bit-vector tables (orc_filter_bv)     */
    int bv_n;
} OrcShape;

static int is_forbidden_small(u32 q, int target, u64 r) {
    u32 f[64];
    int nf = orc_forbidden(q, 0, target, f);
    for (int j = 0; j < nf; j++) if ((u64)f[j] == r % q) return 1;
    return 0;
}

/* mod-6 tile residue rho for (base, k): n = (tile*WP + k)*M + base = 5 (mod 6).
 * The reference derives the same bucket as (target_k_mod6 - tile%6) (v13.cu:64-72,
 * :963-975, :1586-1592) using WP = 1 and M = 5 (mod 6); stated here for any shape. */
static u8 rho_of(u64 WP, u64 M, u64 base, u64 k) {
    for (u32 t = 0; t < 6; t++) {
        u64 n6 = ((t * (WP % 6)) % 6 * (M % 6) + (k % 6) * (M % 6) + base % 6) % 6;
        if (n6 == 5) return (u8)t;
    }
    return 255;
}

ORC_EXPORT void orc_shape_free(void *hp);
static void orc_bv_release(void *v);

/* Restates v13.cu:907-1059 (generate_bases): bases = residues mod M admissible for the
 * CRT primes at target_len; per base the wheel offsets k in [0, WP) whose
 * (k*M + base) is admissible for every wheel prime; partition by mod-6 bucket.
 * And v13.cu:793-901 (init_filter_tables): forbidden-residue tables for the
 * filter primes at depth (sieve_len), position 0 excluded above pos0_max_q. */
ORC_EXPORT void *orc_shape_new(const u32 *crt, int ncrt, const u32 *wheel, int nwheel,
                               int target, int depth, const u32 *fp, int nfp, u32 pos0_max_q) {
    OrcShape *s = (OrcShape *)calloc(1, sizeof(OrcShape));
    s->ncrt = ncrt; s->nwheel = nwheel; s->target = target; s->depth = depth;
    s->M = 1; s->WP = 1;
    for (int i = 0; i < ncrt; i++) { s->crt[i] = crt[i]; s->M *= crt[i]; }
    for (int i = 0; i < nwheel; i++) { s->wheel[i] = wheel[i]; s->WP *= wheel[i]; }

    /* bases */
    u64 cap = 1024;
    s->bases = (u64 *)malloc(cap * sizeof(u64));
    u8 *cforb[16];
    for (int i = 0; i < ncrt; i++) {
        cforb[i] = (u8 *)calloc(crt[i], 1);
        for (u32 r = 0; r < crt[i]; r++) cforb[i][r] = (u8)is_forbidden_small(crt[i], target, r);
    }
    for (u64 r = 0; r < s->M; r++) {
        int ok = 1;
        for (int i = 0; i < ncrt && ok; i++) if (cforb[i][r % crt[i]]) ok = 0;
        if (!ok) continue;
        if ((u64)s->nbases == cap) { cap *= 2; s->bases = (u64 *)realloc(s->bases, cap * sizeof(u64)); }
        s->bases[s->nbases++] = r;
    }
    for (int i = 0; i < ncrt; i++) free(cforb[i]);
    /* wheel per base */
    u32 wforb[8][64]; int wnf[8];
    for (int i = 0; i < nwheel; i++) wnf[i] = orc_forbidden(wheel[i], 0, target, wforb[i]);
    s->wstart = (u32 *)calloc(s->nbases, sizeof(u32));
    s->wcount = (u32 *)calloc(s->nbases, sizeof(u32));
    u64 wcap = 1 << 16, wn = 0;
    s->wk = (u32 *)malloc(wcap * sizeof(u32));
    s->wrho = (u8 *)malloc(wcap);
    for (int b = 0; b < s->nbases; b++) {
        s->wstart[b] = (u32)wn;
        u64 base = s->bases[b];
        for (u64 k = 0; k < s->WP; k++) {
            int ok = 1;
            for (int i = 0; i < nwheel && ok; i++) {
                u32 q = wheel[i];
                u32 r = (u32)((base % q + (k % q) * (s->M % q)) % q);
                for (int j = 0; j < wnf[i]; j++) if (r == wforb[i][j]) { ok = 0; break; }
            }
            if (!ok) continue;
            if (wn == wcap) { wcap *= 2; s->wk = (u32 *)realloc(s->wk, wcap * 4); s->wrho = (u8 *)realloc(s->wrho, wcap); }
            s->wk[wn] = (u32)k;
            s->wrho[wn] = rho_of(s->WP, s->M, base, k);
            wn++;
        }
        s->wcount[b] = (u32)(wn - s->wstart[b]);
    }
    s->nclasses = wn;
    /* mod-6 buckets */
    s->bstart = (u32 *)calloc((size_t)s->nbases * 6, 4);
    s->bcount = (u32 *)calloc((size_t)s->nbases * 6, 4);
    s->bk = (u32 *)malloc((wn ? wn : 1) * 4);
    u64 pos = 0;
    for (int b = 0; b < s->nbases; b++)
        for (int m = 0; m < 6; m++) {
            s->bstart[b * 6 + m] = (u32)pos;
            for (u32 w = 0; w < s->wcount[b]; w++) {
                u64 idx = s->wstart[b] + w;
                if (s->wrho[idx] == m) s->bk[pos++] = s->wk[idx];
            }
            s->bcount[b * 6 + m] = (u32)(pos - s->bstart[b * 6 + m]);
        }

    /* filter tables */
    s->nfp = nfp;
    s->pos0_max_q = pos0_max_q;
    s->fp = (u32 *)malloc((nfp ? nfp : 1) * 4);
    s->fbits = (u64 **)calloc(nfp ? nfp : 1, sizeof(u64 *));
    s->WPMq = (u32 *)malloc((nfp ? nfp : 1) * 4);
    s->Mq = (u32 *)malloc((nfp ? nfp : 1) * 4);
    for (int i = 0; i < nfp; i++) {
        u32 q = fp[i];
        s->fp[i] = q;
        s->fbits[i] = (u64 *)calloc((q + 63) / 64, 8);
        u32 f[64];
        int lo = (q <= pos0_max_q) ? 0 : 1;
        int nf = orc_forbidden(q, lo, depth, f);
        for (int j = 0; j < nf; j++) s->fbits[i][f[j] >> 6] |= 1ull << (f[j] & 63);
        s->WPMq[i] = (u32)((s->WP % q) * (s->M % q) % q);
        s->Mq[i] = (u32)(s->M % q);
    }
    s->basemod = (u32 *)malloc((size_t)(s->nbases ? s->nbases : 1) * (nfp ? nfp : 1) * 4);
    for (int b = 0; b < s->nbases; b++)
        for (int i = 0; i < nfp; i++) s->basemod[(size_t)b * nfp + i] = (u32)(s->bases[b] % fp[i]);
    /* k*M mod q LUT for the leading primes (the reference's kmod / kM_mod tables,
     * v13.cu:849-881), bounded to ~64 MB */
    int nlut = nfp < 256 ? nfp : 256;
    while (nlut > 0 && (u64)nlut * s->WP * 2 > (64ull << 20)) nlut /= 2;
    s->nlut = nlut;
    s->kMlut = (u16 *)malloc((size_t)(s->WP * (nlut ? nlut : 1)) * 2);
    for (u64 k = 0; k < s->WP; k++)
        for (int i = 0; i < nlut; i++)
            s->kMlut[k * nlut + i] = (u16)((k % fp[i]) * s->Mq[i] % fp[i]);
    return s;
}

ORC_EXPORT void orc_shape_free(void *hp) {
    OrcShape *s = (OrcShape *)hp;
    if (!s) return;
    free(s->bases); free(s->wstart); free(s->wcount); free(s->wk); free(s->wrho);
    free(s->bstart); free(s->bcount); free(s->bk);
    for (int i = 0; i < s->nfp; i++) free(s->fbits[i]);
    free(s->fbits); free(s->fp); free(s->WPMq); free(s->Mq); free(s->basemod); free(s->kMlut);
    orc_bv_release(s->bv);
    free(s);
}

ORC_EXPORT int orc_shape_nbases(void *hp) { return ((OrcShape *)hp)->nbases; }
ORC_EXPORT u64 orc_shape_nclasses(void *hp) { return ((OrcShape *)hp)->nclasses; }
ORC_EXPORT u64 orc_shape_M(void *hp) { return ((OrcShape *)hp)->M; }
ORC_EXPORT u64 orc_shape_WP(void *hp) { return ((OrcShape *)hp)->WP; }
ORC_EXPORT void orc_shape_bases(void *hp, u64 *out) {
    OrcShape *s = (OrcShape *)hp;
    memcpy(out, s->bases, (size_t)s->nbases * 8);
}
ORC_EXPORT u32 orc_shape_wheel_count(void *hp, int b) { return ((OrcShape *)hp)->wcount[b]; }
/* bucket m of base b (k ascending): the k offsets tested when tile = m' with rho = m */
ORC_EXPORT u32 orc_shape_bucket(void *hp, int b, int m, u32 *out) {
    OrcShape *s = (OrcShape *)hp;
    u32 n = s->bcount[b * 6 + m];
    if (out) memcpy(out, s->bk + s->bstart[b * 6 + m], n * 4);
    return n;
}
/* all classes in (base, k) order: (base_idx, k, rho) */
ORC_EXPORT void orc_shape_classes(void *hp, u32 *out_b, u32 *out_k, u8 *out_rho) {
    OrcShape *s = (OrcShape *)hp;
    for (int b = 0; b < s->nbases; b++)
        for (u32 w = 0; w < s->wcount[b]; w++) {
            u64 idx = s->wstart[b] + w;
            out_b[idx] = (u32)b; out_k[idx] = s->wk[idx]; out_rho[idx] = s->wrho[idx];
        }
}

/* ------------------------------------------------------------------------- */
/* acceptance predicate                                                       */
/* ------------------------------------------------------------------------- */

/* table form: n passes iff for every filter prime q, n mod q is not forbidden.
 * Restates the predicate of v13.cu:1630-1686 (L2 mask, ext byte table, line kill). */
static inline int killed_at(const OrcShape *s, int i, u32 r) {
    return (int)((s->fbits[i][r >> 6] >> (r & 63)) & 1);
}

/* returns the index of the first filter prime that kills n, or nfp if n passes */
ORC_EXPORT int orc_first_kill(void *hp, u64 lo, u64 hi) {
    OrcShape *s = (OrcShape *)hp;
    u128 n = ((u128)hi << 64) | lo;
    for (int i = 0; i < s->nfp; i++)
        if (killed_at(s, i, umod128(n, s->fp[i]))) return i;
    return s->nfp;
}

/* direct form (independent of the forbidden tables): p_i = 2^i (n+1) - 1, and n
 * passes iff no filter prime q divides p_i for i in its position range —
 * [0, depth) for q <= pos0_max_q, [1, depth) otherwise (v13.cu:797-847). */
ORC_EXPORT int orc_first_kill_brute(void *hp, u64 lo, u64 hi) {
    OrcShape *s = (OrcShape *)hp;
    u128 n = ((u128)hi << 64) | lo;
    for (int i = 0; i < s->nfp; i++) {
        u32 q = s->fp[i];
        int plo = (q <= s->pos0_max_q) ? 0 : 1;
        u64 np1 = (umod128(n, q) + 1) % q;
        u64 t = np1;
        for (int j = 0; j < plo; j++) t = (t * 2) % q;
        for (int pos = plo; pos < s->depth; pos++) {
            if ((t + q - 1) % q == 0) return i;
            t = (t * 2) % q;
        }
    }
    return s->nfp;
}

/* orc_first_kill_brute for every n = a + W*(u0 + j), j in [0, count): the direct
 * predicate over a class window, the model orc_filter_class is checked against */
ORC_EXPORT void orc_first_kill_brute_class(void *hp, u64 a_lo, u64 a_hi, u64 W_lo, u64 W_hi, u64 u0_lo, u64 u0_hi,
                                           u64 count, int *out) {
    u128 a = ((u128)a_hi << 64) | a_lo, W = ((u128)W_hi << 64) | W_lo, u0 = ((u128)u0_hi << 64) | u0_lo;
    for (u64 j = 0; j < count; j++) {
        u128 n = a + W * (u0 + j);
        out[j] = orc_first_kill_brute(hp, (u64)n, (u64)(n >> 64));
    }
}

/* CRT/wheel admissibility of n itself (all p_i, i < target, coprime to every
 * crt/wheel prime) and n = 5 (mod 6): the lattice property of v13.cu:907-1059. */
ORC_EXPORT int orc_on_lattice(void *hp, u64 lo, u64 hi) {
    OrcShape *s = (OrcShape *)hp;
    u128 n = ((u128)hi << 64) | lo;
    if (n % 6 != 5) return 0;
    for (int i = 0; i < s->ncrt; i++) if (is_forbidden_small(s->crt[i], s->target, (u64)umod128(n, s->crt[i]))) return 0;
    for (int i = 0; i < s->nwheel; i++) if (is_forbidden_small(s->wheel[i], s->target, (u64)umod128(n, s->wheel[i]))) return 0;
    return 1;
}

/* decompose n into (base_idx, k_offset, tile) like v13.cu:2921-3062 (--verify-root).
 * returns base_idx or -1 when n is not on the lattice */
ORC_EXPORT int orc_decompose(void *hp, u64 lo, u64 hi, u32 *k_out, u64 *tile_lo, u64 *tile_hi) {
    OrcShape *s = (OrcShape *)hp;
    u128 n = ((u128)hi << 64) | lo;
    u64 base = (u64)(n % s->M);
    int b = -1, L = 0, R = s->nbases - 1;
    while (L <= R) { int m = (L + R) / 2; if (s->bases[m] == base) { b = m; break; } if (s->bases[m] < base) L = m + 1; else R = m - 1; }
    if (b < 0) return -1;
    u128 k = (n - base) / s->M;
    u128 tile = k / s->WP;
    u32 koff = (u32)(k % s->WP);
    u32 rho = (u32)(tile % 6);
    u32 *bk = s->bk + s->bstart[b * 6 + rho];
    u32 nb = s->bcount[b * 6 + rho];
    int found = 0;
    for (u32 w = 0; w < nb; w++) if (bk[w] == koff) { found = 1; break; }
    if (!found) return -1;
    *k_out = koff; *tile_lo = (u64)tile; *tile_hi = (u64)(tile >> 64);
    return b;
}

/* ------------------------------------------------------------------------- */
/* range filter: every lattice point with tile in [t0, t0 + tile_count)        */
/* ------------------------------------------------------------------------- */

typedef struct {
    const OrcShape *s;
    u128 t0;
    u64 tbeg, tend;      /* relative tile range for this worker */
    u32 *out; int64_t cap;
    int64_t n;           /* survivors found (may exceed cap)  */
    int ncut; const int *cut;
    u64 counts[65];      /* counts[0] = candidates, counts[c+1] = alive after cut[c] primes */
} OrcWork;

/* Restates the v33 worker hot loop (v33:1795-1960): per tile incremental residues,
 * per base the mod-6 bucket, per wheel offset ascending-prime early exit.  Survivors
 * are written as the reference's Survivor record {base_idx, k_offset, tile_idx, 0}
 * (v13.cu:207-212) in (tile, base, k) order. */
static void *orc_filter_worker(void *arg) {
    OrcWork *w = (OrcWork *)arg;
    const OrcShape *s = w->s;
    int nfp = s->nfp;
    u32 *T = (u32 *)malloc((nfp ? nfp : 1) * 4);
    u32 *BT = (u32 *)malloc((nfp ? nfp : 1) * 4);
    u128 tile = w->t0 + w->tbeg;
    for (int i = 0; i < nfp; i++) T[i] = (u32)((u128)umod128(tile, s->fp[i]) * s->WPMq[i] % s->fp[i]);
    int nlut = s->nlut;
    for (u64 tr = w->tbeg; tr < w->tend; tr++, tile++) {
        u32 rho = (u32)(tile % 6);
        if (tr != w->tbeg)
            for (int i = 0; i < nfp; i++) { u32 v = T[i] + s->WPMq[i]; if (v >= s->fp[i]) v -= s->fp[i]; T[i] = v; }
        for (int b = 0; b < s->nbases; b++) {
            const u32 *bm = s->basemod + (size_t)b * nfp;
            for (int i = 0; i < nlut; i++) { u32 v = bm[i] + T[i]; if (v >= s->fp[i]) v -= s->fp[i]; BT[i] = v; }
            const u32 *bk = s->bk + s->bstart[b * 6 + rho];
            u32 nb = s->bcount[b * 6 + rho];
            for (u32 wi = 0; wi < nb; wi++) {
                u32 k = bk[wi];
                w->counts[0]++;
                const u16 *lut = s->kMlut + (size_t)k * nlut;
                int kill = nfp;
                for (int i = 0; i < nfp; i++) {
                    u32 q = s->fp[i], r;
                    if (i < nlut) { r = BT[i] + lut[i]; if (r >= q) r -= q; }
                    else {
                        u64 v = (u64)bm[i] + T[i] + (u64)(k % q) * s->Mq[i] % q;
                        r = (u32)(v % q);
                    }
                    if (killed_at(s, i, r)) { kill = i; break; }
                }
                for (int c = 0; c < w->ncut; c++) if (kill >= w->cut[c]) w->counts[c + 1]++;
                if (kill < nfp) continue;
                if (w->n < w->cap) {
                    u32 *o = w->out + 4 * w->n;
                    o[0] = (u32)b; o[1] = k; o[2] = (u32)tr; o[3] = 0;
                }
                w->n++;
            }
        }
    }
    free(T); free(BT);
    return NULL;
}

/* returns the survivor count (which can exceed cap; only cap are written).
 * stage_counts[0] = candidates, stage_counts[c+1] = candidates alive after the first
 * cut[c] filter primes. */
ORC_EXPORT int64_t orc_filter(void *hp, u64 t0_lo, u64 t0_hi, u64 tile_count, u32 *out, int64_t cap,
                              const int *cut, int ncut, u64 *stage_counts, int nthreads) {
    OrcShape *s = (OrcShape *)hp;
    if (nthreads < 1) nthreads = 1;
    if ((u64)nthreads > tile_count) nthreads = tile_count ? (int)tile_count : 1;
    if (ncut > 64) ncut = 64;
    OrcWork *w = (OrcWork *)calloc(nthreads, sizeof(OrcWork));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    u128 t0 = ((u128)t0_hi << 64) | t0_lo;
    /* per-thread private output, merged in tile order */
    for (int t = 0; t < nthreads; t++) {
        w[t].s = s; w[t].t0 = t0;
        w[t].tbeg = tile_count * t / nthreads;
        w[t].tend = tile_count * (t + 1) / nthreads;
        w[t].cap = cap;
        w[t].out = (u32 *)malloc((size_t)(cap > 0 ? cap : 1) * 16);
        w[t].ncut = ncut; w[t].cut = cut;
        pthread_create(&th[t], NULL, orc_filter_worker, &w[t]);
    }
    int64_t total = 0;
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    if (stage_counts) memset(stage_counts, 0, (size_t)(ncut + 1) * 8);
    for (int t = 0; t < nthreads; t++) {
        int64_t keep = w[t].n < w[t].cap ? w[t].n : w[t].cap;
        for (int64_t i = 0; i < keep && total + i < cap; i++) memcpy(out + 4 * (total + i), w[t].out + 4 * i, 16);
        total += w[t].n;
        if (stage_counts) for (int c = 0; c <= ncut; c++) stage_counts[c] += w[t].counts[c];
        free(w[t].out);
    }
    free(w); free(th);
    return total;
}

/* number of lattice points (candidates) with tile in [t0, t0 + tile_count):
 * the v33 counting convention local_wheel++ (v33:1900). */
ORC_EXPORT u64 orc_count_candidates(void *hp, u64 t0_lo, u64 t0_hi, u64 tile_count) {
    OrcShape *s = (OrcShape *)hp;
    u128 t0 = ((u128)t0_hi << 64) | t0_lo;
    u64 per_rho[6] = {0};
    for (int b = 0; b < s->nbases; b++)
        for (int m = 0; m < 6; m++) per_rho[m] += s->bcount[b * 6 + m];
    u64 total = 0;
    u32 r0 = (u32)(t0 % 6);
    for (u32 m = 0; m < 6; m++) {
        /* tiles in range with tile mod 6 == m */
        u64 first = (m + 6 - r0) % 6;
        u64 cnt = tile_count > first ? (tile_count - first + 5) / 6 : 0;
        total += cnt * per_rho[m];
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* u128 Montgomery arithmetic and primality                                   */
/* ------------------------------------------------------------------------- */

typedef struct { u128 n, ninv, r1, r2; } Mont;   /* ninv = -n^-1 mod 2^128, r1 = R mod n, r2 = R^2 mod n */

static void mul256(u128 a, u128 b, u128 *lo, u128 *hi) {
    u64 a0 = (u64)a, a1 = (u64)(a >> 64), b0 = (u64)b, b1 = (u64)(b >> 64);
    u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
    u128 mid = (p00 >> 64) + (u64)p01 + (u64)p10;
    *lo = (u128)(u64)p00 | (mid << 64);
    *hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
}

/* n odd, n < 2^127 */
static void mont_init(Mont *m, u128 n) {
    m->n = n;
    u128 x = n;                               /* n*n = 1 mod 8 */
    for (int i = 0; i < 7; i++) x = x * (2 - n * x);
    m->ninv = (u128)0 - x;
    m->r1 = ((u128)0 - n) % n;                /* 2^128 mod n */
    u128 r = m->r1;
    for (int i = 0; i < 128; i++) { r = r << 1; if (r >= n) r -= n; }   /* n < 2^127: no overflow */
    m->r2 = r;
}

static u128 mont_mul(const Mont *m, u128 a, u128 b) {
    u128 lo, hi, mlo, mhi;
    mul256(a, b, &lo, &hi);
    u128 q = lo * m->ninv;
    mul256(q, m->n, &mlo, &mhi);
    u128 carry = (lo != 0) ? 1 : 0;   /* lo + mlo = 0 mod 2^128 with a carry iff lo != 0 */
    u128 r = hi + mhi + carry;        /* < 2n < 2^128 */
    if (r >= m->n) r -= m->n;
    return r;
}
static u128 to_mont(const Mont *m, u128 a) { return mont_mul(m, a % m->n, m->r2); }
static u128 from_mont(const Mont *m, u128 a) { return mont_mul(m, a, 1); }
static u128 mont_add(const Mont *m, u128 a, u128 b) { u128 r = a + b; if (r >= m->n) r -= m->n; return r; }
static u128 mont_sub(const Mont *m, u128 a, u128 b) { return a >= b ? a - b : a + m->n - b; }

static u128 mont_pow(const Mont *m, u128 base_m, u128 e) {
    u128 r = m->r1;
    int top = 127;
    while (top >= 0 && !((e >> top) & 1)) top--;
    for (int i = top; i >= 0; i--) {
        r = mont_mul(m, r, r);
        if ((e >> i) & 1) r = mont_mul(m, r, base_m);
    }
    return r;
}

/* strong probable-prime test to base a (n odd > 3) */
static int strong_mr(const Mont *m, u64 a) {
    u128 n = m->n, d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    u128 one = m->r1, mone = n - m->r1;
    u128 x = mont_pow(m, to_mont(m, a), d);
    if (x == one || x == mone) return 1;
    for (int i = 1; i < s; i++) {
        x = mont_mul(m, x, x);
        if (x == mone) return 1;
        if (x == one) return 0;
    }
    return 0;
}

/* Jacobi symbol (a/n), n odd positive */
static int jacobi(int64_t a, u128 n) {
    int res = 1;
    u128 aa;
    if (a < 0) { aa = (u128)(-a); if ((n & 3) == 3) res = -res; }
    else aa = (u128)a;
    u128 nn = n;
    aa %= nn;
    while (aa != 0) {
        while (!(aa & 1)) { aa >>= 1; u32 r8 = (u32)(nn & 7); if (r8 == 3 || r8 == 5) res = -res; }
        u128 t = aa; aa = nn; nn = t;
        if ((aa & 3) == 3 && (nn & 3) == 3) res = -res;
        aa %= nn;
    }
    return nn == 1 ? res : 0;
}

static int bitlen128(u128 n);

static int is_square(u128 n) {
    if (n < 2) return 1;
    u128 x = (u128)1 << ((bitlen128(n) + 1) / 2);   /* x >= sqrt(n) */
    for (;;) { u128 y = (x + n / x) >> 1; if (y >= x) break; x = y; }
    return x * x == n;
}

/* strong Lucas probable-prime test with Selfridge's method A parameters
 * (D = 5, -7, 9, -11, ... first with (D/n) = -1; P = 1; Q = (1-D)/4), the Lucas
 * half of GMP's Baillie-PSW (GMP 6.2.1 mpz/stronglucas.c, reached from
 * mpz_probab_prime_p at v13.cu:1841). n odd, not a square, > 13. */
static int strong_lucas(const Mont *m) {
    u128 n = m->n;
    int64_t D = 5;
    for (int tries = 0;; tries++) {
        int j = jacobi(D, n);
        if (j == -1) break;
        if (j == 0) { u128 ad = (u128)(D < 0 ? -D : D); if (ad != n) return 0; }
        if (tries == 20 && is_square(n)) return 0;
        D = D > 0 ? -(D + 2) : -(D - 2);
    }
    int64_t Qs = (1 - D) / 4;
    /* constants in Montgomery form */
    u128 Dm = D >= 0 ? to_mont(m, (u128)D) : mont_sub(m, 0, to_mont(m, (u128)(-D)));
    u128 Qm = Qs >= 0 ? to_mont(m, (u128)Qs) : mont_sub(m, 0, to_mont(m, (u128)(-Qs)));
    u128 d = n + 1; int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    /* binary ladder over the bits of d, left to right: U_1 = 1, V_1 = P = 1, Q^1 */
    u128 U = m->r1, V = m->r1, Qk = Qm;
    int top = 127; while (!((d >> top) & 1)) top--;
    u128 inv2n = (n + 1) >> 1;             /* 2^-1 mod n (plain) */
    u128 half = to_mont(m, inv2n);
    for (int i = top - 1; i >= 0; i--) {
        /* double: U_2k = U V, V_2k = V^2 - 2 Q^k, Q^2k */
        U = mont_mul(m, U, V);
        V = mont_sub(m, mont_mul(m, V, V), mont_add(m, Qk, Qk));
        Qk = mont_mul(m, Qk, Qk);
        if ((d >> i) & 1) {
            /* add one: U_{k+1} = (U + V)/2, V_{k+1} = (D U + V)/2, Q^{k+1} */
            u128 nu = mont_mul(m, mont_add(m, U, V), half);
            u128 nv = mont_mul(m, mont_add(m, mont_mul(m, Dm, U), V), half);
            U = nu; V = nv;
            Qk = mont_mul(m, Qk, Qm);
        }
    }
    if (U == 0 || V == 0) return 1;
    for (int r = 1; r < s; r++) {
        V = mont_sub(m, mont_mul(m, V, V), mont_add(m, Qk, Qk));
        if (V == 0) return 1;
        Qk = mont_mul(m, Qk, Qk);
    }
    return 0;
}

static const u32 SMALLP[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};

/* Baillie-PSW: the decision of mpz_probab_prime_p(n, 1) in GMP 6.2.1 — small
 * divisors, strong base-2 Miller-Rabin, strong Lucas-Selfridge (v13.cu:1841,
 * :1853, :1863, :1878).  Exact for n < 2^64; n must be < 2^127. */
ORC_EXPORT int orc_is_prime(u64 lo, u64 hi) {
    u128 n = ((u128)hi << 64) | lo;
    if (n < 2) return 0;
    for (unsigned i = 0; i < sizeof(SMALLP) / 4; i++) {
        if (n == SMALLP[i]) return 1;
        if (n % SMALLP[i] == 0) return 0;
    }
    if (n < 61 * 61) return 1;
    Mont m; mont_init(&m, n);
    if (!strong_mr(&m, 2)) return 0;
    return strong_lucas(&m);
}

/* strong Miller-Rabin over an explicit base list (config-1 "base-2 MR confirmation") */
ORC_EXPORT int orc_mr(u64 lo, u64 hi, const u32 *bases, int nb) {
    u128 n = ((u128)hi << 64) | lo;
    if (n < 2) return 0;
    if (n < 4) return 1;
    if (!(n & 1)) return 0;
    Mont m; mont_init(&m, n);
    for (int i = 0; i < nb; i++) {
        if (bases[i] % n == 0) continue;
        if (!strong_mr(&m, bases[i])) return 0;
    }
    return 1;
}

ORC_EXPORT void orc_powmod(u64 blo, u64 bhi, u64 elo, u64 ehi, u64 nlo, u64 nhi, u64 *rlo, u64 *rhi) {
    u128 n = ((u128)nhi << 64) | nlo;
    Mont m; mont_init(&m, n);
    u128 r = from_mont(&m, mont_pow(&m, to_mont(&m, ((u128)bhi << 64) | blo), ((u128)ehi << 64) | elo));
    *rlo = (u64)r; *rhi = (u64)(r >> 64);
}

/* ------------------------------------------------------------------------- */
/* GMP confirmation engine (per survivor)                                     */
/* ------------------------------------------------------------------------- */

/* 62 screening primes 67..409 trial-divided at position 0 by the prover
 * (v13.cu:180-199 CHAIN_SCREEN_PRIMES[0..61], used at :1832-1838) */
static const u32 TRIAL[62] = {
    67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163,
    167, 173, 179, 181, 191, 193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251, 257, 263, 269,
    271, 277, 281, 283, 293, 307, 311, 313, 317, 331, 337, 347, 349, 353, 359, 367, 373, 379, 383,
    389, 397, 401, 409};

#define ORC_MAX_CHAIN 64
enum { ORC_TEST_BPSW = 0, ORC_TEST_MR = 1 };

typedef struct { int kind; int nb; u32 bases[16]; } OrcTest;

static int orc_test(const OrcTest *t, u128 n) {
    if (t->kind == ORC_TEST_BPSW) return orc_is_prime((u64)n, (u64)(n >> 64));
    return orc_mr((u64)n, (u64)(n >> 64), t->bases, t->nb);
}

static int bitlen128(u128 n) {
    u64 h = (u64)(n >> 64);
    if (h) return 128 - __builtin_clzll(h);
    u64 l = (u64)n;
    return l ? 64 - __builtin_clzll(l) : 0;
}

/* flags */
#define ORC_F_BITS_OK   1u
#define ORC_F_TRIAL_OK  2u
#define ORC_F_PRIME     4u
#define ORC_F_NONROOT   8u
#define ORC_F_OVERFLOW 16u

/* Restates prover_worker (v13.cu:1794-1924) for one candidate n:
 *   bit-size gate (bits = 0 disables) -> trial division by 67..409 -> primality of n
 *   -> backward walk while (r-1)/2 >= 2 is prime (at most 64 further steps)
 *   -> forward follow from the true root while 2c+1 is prime (length <= 64).
 * Returns flags; writes the root and chain length (0 when n is gated or composite). */
ORC_EXPORT u32 orc_prove_one(u64 lo, u64 hi, int bits, int test_kind, const u32 *bases, int nb,
                             u64 *root_lo, u64 *root_hi, u32 *len_out, u32 *back_out) {
    OrcTest t; t.kind = test_kind; t.nb = nb > 16 ? 16 : nb;
    for (int i = 0; i < t.nb; i++) t.bases[i] = bases[i];
    u128 n = ((u128)hi << 64) | lo;
    u32 flags = 0;
    *root_lo = 0; *root_hi = 0; *len_out = 0; *back_out = 0;
    if (bits && bitlen128(n) != bits) return flags;
    flags |= ORC_F_BITS_OK;
    for (int i = 0; i < 62; i++) if (n % TRIAL[i] == 0) return flags;
    flags |= ORC_F_TRIAL_OK;
    if (n >> 127) { flags |= ORC_F_OVERFLOW; return flags; }   /* beyond the u128 prover */
    if (!orc_test(&t, n)) return flags;
    flags |= ORC_F_PRIME;
    u128 root = n;
    u32 back = 0;
    u128 pred = (n - 1) >> 1;
    if (pred >= 2 && orc_test(&t, pred)) {
        flags |= ORC_F_NONROOT;
        root = pred; back = 1;
        int steps = 0;
        while (steps < ORC_MAX_CHAIN) {
            pred = (root - 1) >> 1;
            if (pred < 2) break;
            if (!orc_test(&t, pred)) break;
            root = pred; steps++; back++;
        }
    }
    u32 len = 1;
    u128 cur = root;
    while (len < ORC_MAX_CHAIN) {
        if (cur >> 126) { flags |= ORC_F_OVERFLOW; break; }
        u128 nx = (cur << 1) + 1;
        if (!orc_test(&t, nx)) break;
        len++; cur = nx;
    }
    *root_lo = (u64)root; *root_hi = (u64)(root >> 64);
    *len_out = len; *back_out = back;
    return flags;
}

/* Batch form over Survivor records {base_idx, k_offset, tile_idx, pad} relative to
 * tile_start, reconstructing n = ((T0 + tile_idx) * WP + k) * M + base (v13.cu:1819-1826).
 * out: per survivor {flags, len, back, pad, root_lo, root_hi} (6 x u64 slots = 48 B) */
typedef struct {
    const OrcShape *s; u128 t0; const u32 *surv; int64_t beg, end; int bits; int kind;
    const u32 *bases; int nb; u64 *out;
} OrcProveWork;

static void *orc_prove_worker(void *arg) {
    OrcProveWork *w = (OrcProveWork *)arg;
    for (int64_t i = w->beg; i < w->end; i++) {
        const u32 *sv = w->surv + 4 * i;
        u128 tile = w->t0 + sv[2];
        u128 n = (tile * w->s->WP + sv[1]) * w->s->M + w->s->bases[sv[0]];
        u64 rl, rh; u32 len, back;
        u32 f = orc_prove_one((u64)n, (u64)(n >> 64), w->bits, w->kind, w->bases, w->nb, &rl, &rh, &len, &back);
        u64 *o = w->out + 6 * i;
        o[0] = f; o[1] = len; o[2] = back; o[3] = 0; o[4] = rl; o[5] = rh;
    }
    return NULL;
}

ORC_EXPORT void orc_prove(void *hp, u64 t0_lo, u64 t0_hi, const u32 *surv, int64_t n, int bits,
                          int test_kind, const u32 *bases, int nb, u64 *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (n < nthreads) nthreads = n > 0 ? (int)n : 1;
    OrcProveWork *w = (OrcProveWork *)calloc(nthreads, sizeof(OrcProveWork));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        w[t].s = (OrcShape *)hp; w[t].t0 = ((u128)t0_hi << 64) | t0_lo; w[t].surv = surv;
        w[t].beg = n * t / nthreads; w[t].end = n * (t + 1) / nthreads;
        w[t].bits = bits; w[t].kind = test_kind; w[t].bases = bases; w[t].nb = nb; w[t].out = out;
        pthread_create(&th[t], NULL, orc_prove_worker, &w[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(w); free(th);
}

/* ------------------------------------------------------------------------- */
/* class-form filter: n = a + W*u for explicit classes (batched / served shapes) */
/* ------------------------------------------------------------------------- */

/* Survivors among u in [u0, u0 + count) of the class a mod W, tested with the
 * shape's filter primes; out receives u - u0 (u32).  The same predicate as
 * orc_first_kill, enumerated per class (the v16 q-iter form, v16...cu:2038-2164). */
ORC_EXPORT int64_t orc_filter_class(void *hp, u64 a, u64 W, u64 u0_lo, u64 u0_hi, u64 count,
                                    u32 *out, int64_t cap, u64 *alive_after, int ncut, const int *cut) {
    OrcShape *s = (OrcShape *)hp;
    int nfp = s->nfp;
    u32 *r0 = (u32 *)malloc((nfp ? nfp : 1) * 4), *step = (u32 *)malloc((nfp ? nfp : 1) * 4);
    u128 u0 = ((u128)u0_hi << 64) | u0_lo;
    u128 n0 = (u128)a + (u128)W * u0;
    for (int i = 0; i < nfp; i++) { r0[i] = umod128(n0, s->fp[i]); step[i] = (u32)(W % s->fp[i]); }
    int64_t ns = 0;
    if (alive_after) memset(alive_after, 0, (size_t)(ncut + 1) * 8);
    for (u64 j = 0; j < count; j++) {
        int kill = nfp;
        for (int i = 0; i < nfp; i++) {
            u32 q = s->fp[i];
            u32 r = (u32)((r0[i] + (u64)(j % q) * step[i]) % q);   /* n_j = n_0 + W*j */
            if (killed_at(s, i, r)) { kill = i; break; }
        }
        if (alive_after) { alive_after[0]++; for (int c = 0; c < ncut; c++) if (kill >= cut[c]) alive_after[c + 1]++; }
        if (kill == nfp) { if (ns < cap) out[ns] = (u32)j; ns++; }
    }
    free(r0); free(step);
    return ns;
}

/* ------------------------------------------------------------------------- */
/* fingerprint family: p_j = Q*D*2^(E+j) - 1 (the v16 Q-iteration sieve)       */
/* ------------------------------------------------------------------------- */

/* Survivors Q in [q0, q0 + count) of the family D, E: Q is killed when a prime q of the
 * list with q not dividing D divides some p_j, j in [0, depth), i.e. when
 * (Q*D*2^E mod q) * 2^j = 1 (mod q) (v16_forbidden_residues v16...cu:1244-1262 and the
 * per-Q mask probe of v16_q_iter_kernel :2038-2164).  out receives Q - q0. */
ORC_EXPORT int64_t orc_filter_family(u64 d_lo, u64 d_hi, int E, int depth, const u32 *primes, int np,
                                     u64 q0_lo, u64 q0_hi, u64 count, u32 *out, int64_t cap) {
    u128 D = ((u128)d_hi << 64) | d_lo, q0 = ((u128)q0_hi << 64) | q0_lo;
    u32 *wq = (u32 *)malloc((np ? np : 1) * 4), *qm = (u32 *)malloc((np ? np : 1) * 4);
    for (int i = 0; i < np; i++) {
        u32 q = primes[i];
        u64 w = (u64)(D % q);
        for (int e = 0; e < E; e++) w = (w * 2) % q;
        wq[i] = (u32)w;                       /* D*2^E mod q (0: immune) */
        qm[i] = umod128(q0, q);
    }
    int64_t ns = 0;
    for (u64 k = 0; k < count; k++) {
        int dead = 0;
        for (int i = 0; i < np && !dead; i++) {
            u32 q = primes[i];
            if (!wq[i]) continue;
            u64 t = (u64)wq[i] * ((qm[i] + k % q) % q) % q;   /* p_0 + 1 mod q */
            for (int j = 0; j < depth && !dead; j++) {
                if (t == 1) dead = 1;             /* q | 2^j (p_0 + 1) - 1 */
                t = t * 2 % q;
            }
        }
        if (!dead) { if (ns < cap) out[ns] = (u32)k; ns++; }
    }
    free(wq); free(qm);
    return ns;
}

/* ------------------------------------------------------------------------- */
/* 64-candidate bit-vector filter (the v34 CPU baseline)                      */
/* ------------------------------------------------------------------------- */

/* Restates the bit-vector stage of v34 (cc_gmp_v34_bit-vector_10.c:642-780 precompute,
 * :2270-2300 search): per (base, mod-6 bucket, block of 64 wheel offsets) and per leading
 * filter prime q, a table kill[j] (j = (base + tile*WP*M) mod q) of the 64 candidates q
 * kills; a block's survivors are ~OR_q kill[q][j_q].  The remaining primes are tested per
 * candidate as in orc_filter.  Same survivors, same (tile, base, k) order. */
typedef struct {
    int nbv;            /* primes covered by kill tables */
    u32 *qoff;          /* [nbv] offset of prime i's q entries in a block record */
    u32 sumq;
    u32 *bm_blk0;       /* [(b*6+m)] first block record of bucket (b, m) */
    u64 *kill;          /* [block record][sumq] */
} OrcBV;

static OrcBV *orc_bv_build(const OrcShape *s, int nbv) {
    OrcBV *v = (OrcBV *)calloc(1, sizeof(OrcBV));
    if (nbv > s->nfp) nbv = s->nfp;
    if (nbv > s->nlut) nbv = s->nlut;
    v->nbv = nbv;
    v->qoff = (u32 *)malloc((nbv ? nbv : 1) * 4);
    for (int i = 0; i < nbv; i++) { v->qoff[i] = v->sumq; v->sumq += s->fp[i]; }
    v->bm_blk0 = (u32 *)malloc((size_t)s->nbases * 6 * 4 + 4);
    u64 nrec = 0;
    for (int bm = 0; bm < s->nbases * 6; bm++) { v->bm_blk0[bm] = (u32)nrec; nrec += (s->bcount[bm] + 63) / 64; }
    v->kill = (u64 *)calloc((size_t)(nrec ? nrec : 1) * (v->sumq ? v->sumq : 1), 8);
    for (int bm = 0; bm < s->nbases * 6; bm++) {
        const u32 *bk = s->bk + s->bstart[bm];
        u32 nb = s->bcount[bm];
        for (u32 blk = 0; blk * 64 < nb; blk++) {
            u64 *rec = v->kill + (size_t)(v->bm_blk0[bm] + blk) * v->sumq;
            u32 cnt = nb - blk * 64 < 64 ? nb - blk * 64 : 64;
            u64 pad = cnt < 64 ? ~((1ull << cnt) - 1) : 0;
            for (int i = 0; i < nbv; i++) {
                u32 q = s->fp[i];
                u64 *kv = rec + v->qoff[i];
                for (u32 t = 0; t < cnt; t++) {
                    u32 km = s->kMlut[(size_t)bk[blk * 64 + t] * s->nlut + i];
                    for (u32 j = 0; j < q; j++) {
                        u32 r = j + km; if (r >= q) r -= q;
                        if (killed_at(s, i, r)) kv[j] |= 1ull << t;
                    }
                }
                for (u32 j = 0; j < q; j++) kv[j] |= pad;
            }
        }
    }
    return v;
}

static void orc_bv_free(OrcBV *v) { if (!v) return; free(v->qoff); free(v->bm_blk0); free(v->kill); free(v); }
static void orc_bv_release(void *v) { orc_bv_free((OrcBV *)v); }

/* build the tables once per shape and prime count (startup work, like v34's precompute) */
ORC_EXPORT void orc_bv_prepare(void *hp, int nbv) {
    OrcShape *s = (OrcShape *)hp;
    if (s->bv && s->bv_n == nbv) return;
    orc_bv_free((OrcBV *)s->bv);
    s->bv = orc_bv_build(s, nbv);
    s->bv_n = nbv;
}

typedef struct {
    const OrcShape *s; const OrcBV *v;
    u128 t0; u64 tbeg, tend;
    u32 *out; int64_t cap; int64_t n; u64 cand;
} OrcBVWork;

static void *orc_bv_worker(void *arg) {
    OrcBVWork *w = (OrcBVWork *)arg;
    const OrcShape *s = w->s;
    const OrcBV *v = w->v;
    int nfp = s->nfp, nlut = s->nlut, nbv = v->nbv;
    u32 *T = (u32 *)malloc((nfp ? nfp : 1) * 4);
    u32 *BT = (u32 *)malloc((nfp ? nfp : 1) * 4);
    u128 tile = w->t0 + w->tbeg;
    for (int i = 0; i < nfp; i++) T[i] = (u32)((u128)umod128(tile, s->fp[i]) * s->WPMq[i] % s->fp[i]);
    for (u64 tr = w->tbeg; tr < w->tend; tr++, tile++) {
        u32 rho = (u32)(tile % 6);
        if (tr != w->tbeg)
            for (int i = 0; i < nfp; i++) { u32 x = T[i] + s->WPMq[i]; if (x >= s->fp[i]) x -= s->fp[i]; T[i] = x; }
        for (int b = 0; b < s->nbases; b++) {
            const u32 *bm = s->basemod + (size_t)b * nfp;
            for (int i = 0; i < nlut; i++) { u32 x = bm[i] + T[i]; if (x >= s->fp[i]) x -= s->fp[i]; BT[i] = x; }
            int bmi = b * 6 + (int)rho;
            const u32 *bk = s->bk + s->bstart[bmi];
            u32 nb = s->bcount[bmi];
            w->cand += nb;
            for (u32 blk = 0; blk * 64 < nb; blk++) {
                const u64 *rec = v->kill + (size_t)(v->bm_blk0[bmi] + blk) * v->sumq;
                u64 c = 0;
                for (int i = 0; i < nbv; i++) c |= rec[v->qoff[i] + BT[i]];
                u64 alive = ~c;
                while (alive) {
                    int t = __builtin_ctzll(alive);
                    alive &= alive - 1;
                    u32 k = bk[blk * 64 + t];
                    const u16 *lut = s->kMlut + (size_t)k * nlut;
                    int kill = 0;
                    for (int i = nbv; i < nfp; i++) {
                        u32 q = s->fp[i], r;
                        if (i < nlut) { r = BT[i] + lut[i]; if (r >= q) r -= q; }
                        else r = (u32)(((u64)bm[i] + T[i] + (u64)(k % q) * s->Mq[i] % q) % q);
                        if (killed_at(s, i, r)) { kill = 1; break; }
                    }
                    if (kill) continue;
                    if (w->n < w->cap) { u32 *o = w->out + 4 * w->n; o[0] = (u32)b; o[1] = k; o[2] = (u32)tr; o[3] = 0; }
                    w->n++;
                }
            }
        }
    }
    free(T); free(BT);
    return NULL;
}

/* survivors (count returned, cap written) and the candidate count, like orc_filter */
ORC_EXPORT int64_t orc_filter_bv(void *hp, u64 t0_lo, u64 t0_hi, u64 tile_count, u32 *out, int64_t cap,
                                 int nbv, u64 *candidates, int nthreads) {
    OrcShape *s = (OrcShape *)hp;
    orc_bv_prepare(hp, nbv);
    OrcBV *v = (OrcBV *)s->bv;
    if (nthreads < 1) nthreads = 1;
    if ((u64)nthreads > tile_count) nthreads = tile_count ? (int)tile_count : 1;
    OrcBVWork *w = (OrcBVWork *)calloc(nthreads, sizeof(OrcBVWork));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    u128 t0 = ((u128)t0_hi << 64) | t0_lo;
    for (int t = 0; t < nthreads; t++) {
        w[t].s = s; w[t].v = v; w[t].t0 = t0;
        w[t].tbeg = tile_count * t / nthreads; w[t].tend = tile_count * (t + 1) / nthreads;
        w[t].cap = cap; w[t].out = (u32 *)malloc((size_t)(cap > 0 ? cap : 1) * 16);
        pthread_create(&th[t], NULL, orc_bv_worker, &w[t]);
    }
    int64_t total = 0; u64 cand = 0;
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    for (int t = 0; t < nthreads; t++) {
        int64_t keep = w[t].n < w[t].cap ? w[t].n : w[t].cap;
        for (int64_t i = 0; i < keep && total + i < cap; i++) memcpy(out + 4 * (total + i), w[t].out + 4 * i, 16);
        total += w[t].n; cand += w[t].cand;
        free(w[t].out);
    }
    if (candidates) *candidates = cand;
    free(w); free(th);
    return total;
}
