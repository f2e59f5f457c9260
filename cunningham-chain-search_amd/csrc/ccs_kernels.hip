// gfx950 kernels of the Cunningham-chain (first kind) candidate pipeline.
//
//   sieve_kernel  (K1)  enumerate candidates n = a + W*u in registers, dense bit-pattern
//                        stage over the small filter primes (LDS patterns, 32 candidates per
//                        lane per row), wave64 ballot compaction into an LDS queue, then
//                        per-candidate stage-B tests against LDS bitsets.
//                        Replaces cc18_filter_kernel (v13.cu:1551-1714).
//   deep_kernel   (K1c) remaining (large) filter primes on the rare stage-B survivors and
//                        emission in the reference Survivor format (v13.cu:207-212, :1688-1710).
//   prove_kernel  (K2)  the GMP confirmation engine semantics of prover_worker
//                        (v13.cu:1794-1924) with u128 Montgomery Baillie-PSW, replacing
//                        cc18_prove_kernel (v13.cu:1462-1538).
#include <hip/hip_runtime.h>
#include "ccs_arith.hpp"
#include "../../include/ccs.h"

namespace ccs {

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ u32 lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// number of set bits of m below this lane
__device__ __forceinline__ u32 lanes_below(u64 m) {
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

__device__ __forceinline__ u64 ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ u32 bcast32(u32 v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ u64 bcast64(u64 v) {
    u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
    u32 hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
    return ((u64)hi << 32) | lo;
}

// order this wave's LDS writes before its later LDS reads
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// exact residues
// ---------------------------------------------------------------------------

// phase of u_start for a prime: s = (u_start + a * W^-1) mod q, so that a candidate
// j units after u_start is killed iff (s + j) mod q lies in the prime's pattern.
__device__ __forceinline__ u32 phase(const PrimeMod& pm, u64 a, u64 a_hi, u64 us_lo, u64 us_hi) {
    u32 am = mod128(a, a_hi, pm);
    u32 um = mod128(us_lo, us_hi, pm);
    u32 t = mod52((u64)am * pm.winv, pm.q, pm.rcp);
    u32 s = um + t;
    if (s >= pm.q) s -= pm.q;
    return s;
}

// ---------------------------------------------------------------------------
// work units
// ---------------------------------------------------------------------------

struct Unit {
    u64 a, a_hi;
    u64 us_lo, us_hi;
    u32 n;       // candidates in the unit
    u32 owner;   // class or job index
    u32 j0;      // j of the first candidate
    u32 ph_slot; // resident phases of the job (kNoSlot: none)
    u64 ph_off;  // u_start minus the u of the resident phases
};

// (r + x) mod q for a resident phase r < q and an offset x < 2^40 (exact below 2^52)
__device__ __forceinline__ u32 advance_phase(u32 r, u64 x, const PrimeMod& pm) { return mod52(x + r, pm.q, pm.rcp); }

__device__ __forceinline__ void add128(u64& lo, u64& hi, u64 v) {
    u64 l = lo + v;
    hi += (l < lo);
    lo = l;
}

// job owning a unit: 64-ary search over the unit prefix array, all lanes loading
__device__ __forceinline__ u32 find_job(const JobSrc& J, u64 unit, u32 lane) {
    u32 lo = 0, hi = J.n_jobs;   // invariant: prefix[lo] <= unit < prefix[hi]
    while (hi - lo > 1) {
        u32 span = hi - lo;
        u32 step = (span + 63) / 64;
        u32 idx = lo + lane * step;
        bool le = (idx < hi) && (J.unit_prefix[idx] <= unit);
        u64 m = ballot(le);
        u32 last = 63 - __builtin_clzll(m);     // lane 0 is always <= (prefix[lo] <= unit)
        u32 nlo = lo + last * step;
        u32 nhi = nlo + step < hi ? nlo + step : hi;
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

__device__ __forceinline__ bool resolve(const SieveArgs& A, u64 unit, u32 lane, Unit& U) {
    if (A.mode == 0) {
        const PlainSrc& P = A.plain;
        u64 c = unit / P.chunks_per_class;
        u64 ch = unit - c * P.chunks_per_class;
        ClassRec cr = P.cls[c];
        u64 jlo = cr.rho < P.off ? 1 : 0;
        u64 jhi = ((u64)P.tile_count + P.off + 5 - cr.rho) / 6;
        u64 beg = ch * A.unit_j, end = beg + A.unit_j;
        if (beg < jlo) beg = jlo;
        if (end > jhi) end = jhi;
        if (end <= beg) return false;
        U.a = cr.a;
        U.a_hi = 0;
        U.us_lo = P.U0_lo;
        U.us_hi = P.U0_hi;
        add128(U.us_lo, U.us_hi, beg);
        U.n = (u32)(end - beg);
        U.owner = (u32)c;
        U.j0 = (u32)beg;
        U.ph_slot = kNoSlot;
        U.ph_off = 0;
        return true;
    }
    const JobSrc& J = A.jobsrc;
    u32 job = find_job(J, unit, lane);
    JobRec jr = J.jobs[job];
    const u32 uj = A.unit_j ? A.unit_j : jr.unit_j;
    u64 beg = (unit - J.unit_prefix[job]) * uj;
    if (beg >= jr.count) return false;
    u64 n = jr.count - beg;
    if (n > uj) n = uj;
    U.a = jr.a;
    U.a_hi = jr.a_hi;
    U.us_lo = jr.u0_lo;
    U.us_hi = jr.u0_hi;
    add128(U.us_lo, U.us_hi, beg);
    U.n = (u32)n;
    U.owner = job + J.job_base;
    U.j0 = (u32)beg;
    U.ph_slot = A.phases ? jr.ph_slot : kNoSlot;   // (read by the resident variant only)
    U.ph_off = jr.ph_off + beg;
    return true;
}

// ---------------------------------------------------------------------------
// K1: sieve kernel
// ---------------------------------------------------------------------------

// x mod q for x < 2^21, q < 2^14 from the f32 reciprocal of the hardware (about 1 ulp): the
// truncated quotient is within one of floor(x / q) (the product's error is below 2^-11),
// and one unsigned min each way puts the remainder back in [0, q)
__device__ __forceinline__ u32 mod_small(u32 x, u32 q) {
    const int qt = (int)((float)x * __builtin_amdgcn_rcpf((float)q));
    u32 r = x - (u32)__mul24(qt, (int)q);
    r = min(r, r + q);   // quotient one too large: r wrapped below zero
    r = min(r, r - q);   // quotient one too small: r >= q
    return r;
}

// x mod q for x < 2^21 with the host's reciprocal of q, rounded up so that the truncated
// f32 product is exactly floor(x / q)
__device__ __forceinline__ u32 mod_exact(u32 x, int nq, u32 rcp_bits) {
    const int qt = (int)((float)x * __uint_as_float(rcp_bits));
    return (u32)(__mul24(qt, nq) + (int)x);
}

// Per-lane copies of the stage-B prime constants: lane l holds prime 64*i + l in group i.
// Group 0, probed for every dense survivor, keeps its phase, -q, bitset offset and
// reciprocal in four registers; the rarely probed later groups keep offset | q << 18
// (q < 2^14) in one register and, up to 8 groups, the phase and the reciprocal in two more.
// With 12 or 16 groups that made the instances spill, so there a later group's phase is a
// 16-bit half of a register shared with the next group, the reciprocal is the hardware's
// approximate one (mod_small), and the packed values reach each probe through an opaque
// copy (else the compiler hoists every group's unpacked values out of the row loops and
// the packing is undone).
template <int NBV>
struct TierBRegs {
    static constexpr bool kPack = NBV > 8;
    static constexpr int kLater = NBV > 1 ? NBV - 1 : 1;
    static constexpr int kPhaseRegs = kPack ? NBV / 2 : kLater;
    u32 s0;                        // group 0: phase of the current unit
    int nq0;                       // group 0: -q
    u32 off0;                      // group 0: dword offset of the bitset (LDS)
    u32 rcp0;                      // group 0: f32 bits of the rounded-up 1/q
    u32 qo[NBV];                   // groups >= 1: offset | q << 18; [0] unused
    u32 sp[kPhaseRegs];            // groups >= 1: phases (packed: group i in half (i-1)&1 of sp[(i-1)/2])
    u32 rc[kPack ? 1 : kLater];    // groups >= 1, unpacked form: f32 bits of the rounded-up 1/q

    __device__ __forceinline__ u32 qo_at(int i) const {
        u32 v = qo[i];
        if constexpr (kPack) asm volatile("" : "+v"(v));
        return v;
    }
    __device__ __forceinline__ u32 s(int i) const {
        if (i == 0) return s0;
        if constexpr (kPack) {
            u32 w = sp[(i - 1) >> 1];
            asm volatile("" : "+v"(w));
            return __builtin_amdgcn_ubfe(w, 16u * (u32)((i - 1) & 1), 16u);
        } else {
            return sp[i - 1];
        }
    }
    __device__ __forceinline__ void set_s(int i, u32 v) {
        if (i == 0) {
            s0 = v;
        } else if constexpr (kPack) {
            u32& w = sp[(i - 1) >> 1];
            w = ((i - 1) & 1) ? ((w & 0xffffu) | (v << 16)) : ((w & 0xffff0000u) | v);
        } else {
            sp[i - 1] = v;
        }
    }
    __device__ __forceinline__ void set_rcp(int i, float r) {
        if (i == 0) rcp0 = __float_as_uint(r);
        else if constexpr (!kPack) rc[i - 1] = __float_as_uint(r);
    }
    // x mod q_i for x < 2^21
    __device__ __forceinline__ u32 residue(int i, u32 x) const {
        if (i == 0) return mod_exact(x, nq0, rcp0);
        const u32 q = qo_at(i) >> kTierBOffBits;
        if constexpr (kPack) return mod_small(x, q);
        else return mod_exact(x, -(int)q, rc[i - 1]);
    }
    __device__ __forceinline__ u32 off(int i) const {
        return i == 0 ? off0 : (qo_at(i) & ((1u << kTierBOffBits) - 1u));
    }
};

// Stage B for one candidate j (uniform across the wave): the stage-B primes are held one
// per lane (prime 64*i + lane in group i), so one probe of every lane tests 64 primes at
// once; x = j + s < 2^21, and the probe reads bit (x mod q) of the prime's forbidden-residue
// bitset.  Most candidates die in the first group of 64 primes, so the later groups are
// rarely read; their bitsets sit in LDS as far as it holds them, the rest in device memory
// (L2-resident).
// the probe of stage-B group i for candidate j: this lane's prime kills j
template <int NBV>
__device__ __forceinline__ bool stage_b_bit(const u32* __restrict__ mem, const TierBRegs<NBV>& R, int i, u32 j) {
    const u32 r = R.residue(i, j + R.s(i));
    const u32 w = mem[R.off(i) + __builtin_amdgcn_ubfe(r, 5u, 11u)];
    return __builtin_amdgcn_ubfe(w, r, 1u) != 0u;
}

// Group-0 probes of several candidates with their bitset reads all in flight: the words
// are read first, and a scheduling barrier keeps every read ahead of the first use (left
// to itself the scheduler reuses the first word's register for the next residue, which
// waits out each read in turn).
template <int NBV, int N>
__device__ __forceinline__ void stage_b_bits0(const u32* __restrict__ lds, const TierBRegs<NBV>& R, const u32 (&j)[N],
                                              bool (&kill)[N]) {
    u32 r[N], w[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        r[k] = R.residue(0, j[k] + R.s(0));
        w[k] = lds[R.off(0) + __builtin_amdgcn_ubfe(r[k], 5u, 11u)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < N; ++k) kill[k] = __builtin_amdgcn_ubfe(w[k], r[k], 1u) != 0u;
}

// groups 1..NBV-1 for one candidate (groups past nb_lds live in device memory)
template <int NBV>
__device__ __forceinline__ bool stage_b_rest(const u32* __restrict__ lds, const u32* __restrict__ far,
                                             u32 nb_lds, u32 nb, const TierBRegs<NBV>& R, u32 j) {
#pragma unroll
    for (int i = 1; i < NBV; ++i) {
        if ((u32)i * 64u >= nb) break;   // a shape with fewer groups than its instance holds
        if (ballot(stage_b_bit<NBV>((u32)i * 64u < nb_lds ? lds : far, R, i, j))) return false;
    }
    return true;
}

// Two candidates at once: their first-group probes are independent, so both LDS reads are
// in flight together; the rare candidates that pass the first group continue alone.
template <int NBV>
__device__ __forceinline__ void stage_b_pair(const u32* __restrict__ lds, const u32* __restrict__ far,
                                             u32 nb_lds, u32 nb, const TierBRegs<NBV>& R, u32 ja, u32 jb,
                                             bool& pa, bool& pb) {
    const u32 j[2] = {ja, jb};
    bool b[2];
    stage_b_bits0<NBV, 2>(lds, R, j, b);
    pa = !ballot(b[0]) && stage_b_rest<NBV>(lds, far, nb_lds, nb, R, ja);
    pb = !ballot(b[1]) && stage_b_rest<NBV>(lds, far, nb_lds, nb, R, jb);
}

template <int NBV>
__device__ __forceinline__ bool stage_b_pass(const u32* __restrict__ lds, const u32* __restrict__ far,
                                             u32 nb_lds, u32 nb, const TierBRegs<NBV>& R, u32 j) {
    return !ballot(stage_b_bit<NBV>(lds, R, 0, j)) && stage_b_rest<NBV>(lds, far, nb_lds, nb, R, j);
}

typedef u32 u32x4 __attribute__((ext_vector_type(4)));   // one 16-byte LDS block

// Stage-B group 0 for the n queued candidates at q (a wave's LDS queue, 16-byte aligned; n a
// multiple of four, at most 64): the probes run back to back, four candidates per 16-byte
// read of the queue (all lanes read the same block), with no branch per candidate.  Returns
// the mask of the candidates that no group-0 prime kills.
template <int NBV>
__device__ __forceinline__ u64 stage_b_group0(const u32* __restrict__ lds, const TierBRegs<NBV>& R, const u32* q,
                                              u32 n) {
    u64 pass = 0;
#pragma unroll 1
    for (u32 t = 0; t < n; t += 4) {
        const u32x4 jq = *(const u32x4*)(q + t);
        const u32 j[4] = {jq.x, jq.y, jq.z, jq.w};
        bool b[4];
        stage_b_bits0<NBV, 4>(lds, R, j, b);
        const u64 live = (u64)(ballot(b[0]) == 0) | (u64)(ballot(b[1]) == 0) << 1 | (u64)(ballot(b[2]) == 0) << 2 |
                         (u64)(ballot(b[3]) == 0) << 3;
        pass |= live << t;
    }
    return pass;
}

// Marks the four words of a 16-byte LDS block as used together at this point (an empty asm
// that reads and writes them: no instruction).  Without it the compiler may split a block
// read whose words are not all used into 32-bit reads, which run at the narrow reads' LDS
// rate.
__device__ __forceinline__ void keep_block(u32x4& v) { __asm__ volatile("" : "+v"(v)); }

// The 128 pattern bits of a lane's window, for RG rows: bits [sh mod 32, +128) of the words
// C, C+1, ..., C+4 of the eight words lo.xyzw, hi.xyzw read at the window's aligned 16-byte
// block (v_alignbit takes the low five bits of sh); C is the pattern's word offset within the
// block, the same for every lane and fixed at compile time (the dense loop runs the patterns
// of each offset in a loop of its own).
template <int C, int RG>
__device__ __forceinline__ void shift_window(const u32x4 (&lo)[RG], const u32x4 (&hi)[RG], u32 sh,
                                             u32 (&x)[RG][kRowWords]) {
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        const u32 w[8] = {lo[r].x, lo[r].y, lo[r].z, lo[r].w, hi[r].x, hi[r].y, hi[r].z, hi[r].w};
#pragma unroll
        for (int k = 0; k < kRowWords; ++k) x[r][k] = __builtin_amdgcn_alignbit(w[C + k + 1], w[C + k], sh);
    }
}

// One dense pattern's kill bits for RG rows, OR-ed into acc: the lane reads the two aligned
// 16-byte blocks at its window (ds_read_b128: the 64 lanes read 1 KiB apiece, conflict-free).
template <int C, int RG>
__device__ __forceinline__ void dense_read(const u32* lds, u32 lane, u32 ba, u32x4 (&lo)[RG], u32x4 (&hi)[RG]) {
    const u32x4* src = (const u32x4*)(lds + ((ba >> 5) & ~3u)) + lane;
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        lo[r] = src[64 * r];
        hi[r] = src[64 * r + 1];
    }
}

// The dense stage of one row group over the patterns whose word offset is C (mask: their
// slot indices), two patterns per step so that their four block reads per row are in flight
// together.  sd: lane p holds pattern p's bit address in LDS.
template <int C, int RG>
__device__ __forceinline__ void dense_class(const u32* lds, u32 lane, u32 sd, u32 mask, u32 first,
                                            u32 (&acc)[RG][kRowWords]) {
    while (mask & (mask - 1u)) {   // two patterns or more left
        const u32 p0 = (u32)__builtin_ctz(mask);
        mask &= mask - 1u;
        const u32 p1 = (u32)__builtin_ctz(mask);
        mask &= mask - 1u;
        const u32 b0 = __builtin_amdgcn_readlane(sd, first + p0), b1 = __builtin_amdgcn_readlane(sd, first + p1);
        u32x4 lo0[RG], hi0[RG], lo1[RG], hi1[RG];
        dense_read<C, RG>(lds, lane, b0, lo0, hi0);
        dense_read<C, RG>(lds, lane, b1, lo1, hi1);
#pragma unroll
        for (int r = 0; r < RG; ++r) {
            keep_block(lo0[r]);
            keep_block(hi0[r]);
            keep_block(lo1[r]);
            keep_block(hi1[r]);
        }
        u32 x0[RG][kRowWords], x1[RG][kRowWords];
        shift_window<C, RG>(lo0, hi0, b0, x0);
        shift_window<C, RG>(lo1, hi1, b1, x1);
#pragma unroll
        for (int r = 0; r < RG; ++r)
#pragma unroll
            for (int k = 0; k < kRowWords; ++k) acc[r][k] |= x0[r][k] | x1[r][k];
    }
    if (mask) {   // the last one
        const u32 b0 = __builtin_amdgcn_readlane(sd, first + (u32)__builtin_ctz(mask));
        u32x4 lo0[RG], hi0[RG];
        dense_read<C, RG>(lds, lane, b0, lo0, hi0);
#pragma unroll
        for (int r = 0; r < RG; ++r) {
            keep_block(lo0[r]);
            keep_block(hi0[r]);
        }
        u32 x0[RG][kRowWords];
        shift_window<C, RG>(lo0, hi0, b0, x0);
#pragma unroll
        for (int r = 0; r < RG; ++r)
#pragma unroll
            for (int k = 0; k < kRowWords; ++k) acc[r][k] |= x0[r][k];
    }
}

// the patterns of word offset C: slots 0..31, then 32..63 (rarely any)
template <int C, int RG>
__device__ __forceinline__ void dense_offset(const u32* lds, u32 lane, u32 sd, u64 mask, u32 (&acc)[RG][kRowWords]) {
    dense_class<C, RG>(lds, lane, sd, (u32)mask, 0u, acc);
    if (mask >> 32) dense_class<C, RG>(lds, lane, sd, (u32)(mask >> 32), 32u, acc);
}

// GROUP: the launch sieves the jobs of several shapes (tables loaded per shape);
// RESIDENT: its jobs may carry resident phases (a group-launch variant of its own, so the
// launches without them keep their register allocation)
template <int RG, int NBV, bool GROUP, bool RESIDENT>
__global__ __launch_bounds__(1024) void sieve_kernel(SieveArgs A) {
    constexpr bool kQueue = survivor_queue(RG);
    extern __shared__ uint4 lds_blocks[];   // 16-byte aligned: the dense patterns are read as b128
    u32* const lds = (u32*)lds_blocks;
    const u32 tid = threadIdx.x;
    u32* const ctl = lds + A.lds_ctl;   // scheduler words and queues past the tables (16-byte aligned)
    const u32 lane = lane_id();

    // the tables of the shape whose patterns are in LDS (a one-shape launch reads them from
    // its arguments); per lane: dense slot `lane` (modulus, pattern offset, per-row-group
    // advance) and stage-B prime 64*i + lane
    u32 g_nd = 0, g_nb = 0, g_nb_lds = 0;
    const u32* g_far = nullptr;
    const PrimeMod* g_pmod = nullptr;
    auto t_nd = [&]() -> u32 { if constexpr (GROUP) return g_nd; else return A.tab.nd; };
    auto t_nb = [&]() -> u32 { if constexpr (GROUP) return g_nb; else return A.tab.nb; };
    auto t_nb_lds = [&]() -> u32 { if constexpr (GROUP) return g_nb_lds; else return A.tab.nb_lds; };
    auto t_far = [&]() -> const u32* { if constexpr (GROUP) return g_far; else return A.tab.tierb_far; };
    auto t_pmod = [&]() -> const PrimeMod* { if constexpr (GROUP) return g_pmod; else return A.tab.pmod; };
    u32 dq = 1, doff = 0, dstep = 0;
    // lanes past the last stage-B prime hold q = 1 and point at a zero word of the group's
    // memory, so their probe never kills and no lane mask is needed
    TierBRegs<NBV> R;
    auto set_regs = [&](const ShapeTab T) {
        g_nd = T.nd;
        g_nb = T.nb;
        g_nb_lds = T.nb_lds;
        g_far = T.tierb_far;
        g_pmod = T.pmod;
        dq = 1;
        doff = 0;
        dstep = 0;
        if (lane < T.nd) {
            const DensePrime d = T.dense[lane];
            dq = d.q;
            doff = d.lds_off;
            dstep = d.step;
        }
#pragma unroll
        for (int i = 0; i < NBV; ++i) {
            const u32 p = (u32)i * 64u + lane;
            TierBPrime tp;
            tp.q = 1;
            tp.lds_off = (i == 0 || (u32)i * 64u < T.nb_lds) ? T.zero_lds : T.zero_far;   // group 0 is in LDS
            tp.rcp = 1.0f;
            tp.pad = 0;
            if (p < T.nb) tp = T.tierb[p];
            R.set_s(i, 0);
            R.set_rcp(i, tp.rcp);
            if (i == 0) {
                R.nq0 = -(int)tp.q;
                R.off0 = tp.lds_off;
            }
            R.qo[i] = tp.lds_off | (tp.q << kTierBOffBits);
        }
    };
    u64 cnt_dense = 0;   // candidates alive after the dense stage (= tested in stage B)

    // Stage-B survivors wait in registers (lane k holds the k-th) and go out 64 at a time
    // with one atomic, so the shared output cursor is not a point of contention when
    // survivors are dense (low depths).
    u32 q_n = 0, q_owner = 0, q_j = 0;
    auto flush = [&]() {
        if (!q_n) return;
        u32 base = 0;
        if (lane == 0) base = atomicAdd(A.mid_count, q_n);
        base = bcast32(base);
        if (lane < q_n && base + lane < A.mid_cap) {
            MidRec rec;
            rec.owner = q_owner;
            rec.j = q_j;
            A.mid[base + lane] = rec;
        }
        q_n = 0;
    };

    // Work distribution with few global atomics, in two phases.
    //  A: the first n_a units are split into equal contiguous ranges, one per block, and in
    //     a launch of several shapes each shape's share of them is split so (a unit's cost
    //     differs between shapes several times over, so equal unit counts per block would
    //     leave the blocks with the cheap shapes idle); the block's waves take runs of
    //     consecutive units from a range through an LDS cursor (a run is
    //     1/(8 * waves per block) of the range's rest, 1..64 units).
    //  B: the last n_pool units (a sixteenth, at least one per wave of the grid) form a
    //     pool shared by all blocks, drawn in block-wide rounds: thread 0 claims one unit
    //     per wave for the whole block with one global atomic (the next round's claim is in
    //     flight while this round is sieved).
    // The work comes in steps: a phase-A range, or a round, cut where the units change
    // shape.  Every step starts with two barriers (and a table load between them when the
    // shape changes), every round with two more; the step sequence depends only on
    // block-uniform values (the ranges, the claims read back from LDS, the shape bounds),
    // so every wave passes the same barriers, and all leave in the same round.
    const u32 wpb = blockDim.x >> 6, wib = tid >> 6;
    u32* const cqb = ctl + kSchedWords + wib * kQueueWords;   // this wave's candidate queue
    const u32 n_blocks = gridDim.x;
    const u32 n_units = (u32)A.n_units;   // < 2^32: the launchers reject larger launches
    u32 n_pool = n_units / 16;
    if (n_pool < n_blocks * wpb) n_pool = n_blocks * wpb;
    if (n_pool > n_units) n_pool = n_units;
    const u32 n_a = n_units - n_pool;
    // shape of a unit, and the end of that shape's units (one shape outside group launches)
    auto shape_of = [&](u32 u) -> u32 {
        if constexpr (!GROUP) return 0;
        u32 lo = 0, hi = A.n_group;   // shape_units[lo] <= u < shape_units[hi]
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if ((u32)A.shape_units[mid] <= u) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    auto shape_end = [&](u32 sh) -> u32 {
        if constexpr (GROUP) return (u32)A.shape_units[sh + 1];
        return n_units;
    };

    if constexpr (!GROUP) set_regs(A.tab);   // one shape: its lane constants are loop-invariant
    u32 next_g = 0;   // thread 0: the claim for the next round (one unit per wave)
    if (tid == 0) next_g = (u32)atomicAdd(A.unit_counter, (u64)wpb);
    u32 cur_shape = 0xffffffffu;
    u32 a_shape = 0;                           // phase A: the next shape's range
    const u32 n_shapes = GROUP ? A.n_group : 1u;
    bool in_pool = false;
    u32 r_base = 0, r_next = 0, r_end = 0;     // phase B: the round, the next step's start
    u32 sd = 0;                                             // dense phase (lane < nd)
    u32 prev_owner = 0xffffffffu, prev_j0 = 0, prev_n = 0;  // the wave's previous unit
    for (;;) {
        u32 st_lo, st_hi, sh;
        if (!in_pool && a_shape < n_shapes) {
            // this block's share of the shape's units below n_a (block-uniform values: a
            // block skips an empty share as a whole)
            sh = a_shape++;
            u32 lo = 0, hi = n_units;
            if constexpr (GROUP) {
                lo = (u32)A.shape_units[sh];
                hi = (u32)A.shape_units[sh + 1];
            }
            if (hi > n_a) hi = n_a;
            if (hi <= lo) continue;
            const u32 len = hi - lo;
            st_lo = lo + (u32)((u64)len * blockIdx.x / n_blocks);
            st_hi = lo + (u32)((u64)len * (blockIdx.x + 1) / n_blocks);
            if (st_hi == st_lo) continue;
        } else {
            in_pool = true;
            if (r_next >= r_end) {
                __syncthreads();                       // the previous round is finished
                if (tid == 0) ctl[2] = next_g;
                __syncthreads();                       // the round's claim is visible
                const u32 g = bcast32(ctl[2]);            // uniform: keep the schedule in SGPRs
                if (g >= n_pool) break;
                if (tid == 0) next_g = (u32)atomicAdd(A.unit_counter, (u64)wpb);
                r_base = n_a + g;
                r_next = r_base;
                r_end = n_units - r_base > wpb ? r_base + wpb : n_units;
            }
            st_lo = r_next;
            sh = shape_of(st_lo);
            st_hi = shape_end(sh) < r_end ? shape_end(sh) : r_end;
            r_next = st_hi;
        }
        const bool reload = sh != cur_shape;
        __syncthreads();                               // every wave is past the last step
        if (reload) {
            ShapeTab T;
            if constexpr (GROUP) T = A.group[sh];
            else T = A.tab;
            for (u32 i = tid; i < T.lds_words; i += blockDim.x) lds[i] = T.lds_image[i];
        }
        if (tid == 0) ctl[0] = 0;
        __syncthreads();                               // the tables and the cursor are ready
        if (reload) {
            if constexpr (GROUP) set_regs(A.group[sh]);
            cur_shape = sh;
            prev_owner = 0xffffffffu;
        }
        const u32 st_n = st_hi - st_lo;
        u32 seen = 0;                      // phase-A cursor value this wave last drew
        bool took = false;                 // phase B: this wave's unit of the round taken
        u32 cur = 0, cur_end = 0;
        for (;;) {
            if (cur == cur_end) {
                if (!in_pool) {
                    const u32 rem = st_n > seen ? st_n - seen : 0;
                    u32 c = rem / (8 * wpb);
                    c = c < 1 ? 1 : (c > 64 ? 64 : c);
                    u32 got = 0;
                    if (lane == 0) got = atomicAdd(&ctl[0], c);
                    got = bcast32(got);
                    if (got >= st_n) break;
                    cur = st_lo + got;
                    cur_end = st_n - got > c ? st_lo + got + c : st_hi;
                    seen = got + c;
                } else {
                    const u32 mine = r_base + wib;
                    if (took || mine < st_lo || mine >= st_hi) break;
                    took = true;
                    cur = mine;
                    cur_end = mine + 1;
                }
            }
            const u64 unit = (u64)cur + A.unit_base;
            ++cur;
            Unit U;
            if (!resolve(A, unit, lane, U)) continue;
            U.a = bcast64(U.a);
            U.a_hi = bcast64(U.a_hi);
            U.us_lo = bcast64(U.us_lo);
            U.us_hi = bcast64(U.us_hi);
            U.n = bcast32(U.n);
            U.owner = bcast32(U.owner);
            U.j0 = bcast32(U.j0);
            U.ph_slot = bcast32(U.ph_slot);
            U.ph_off = bcast64(U.ph_off);

            // phases: dense prime p held by lane p as an LDS bit address, stage-B prime
            // (64*i + lane) held in group i of R.  A unit that continues the previous one (same
            // owner, next u, previous unit a whole number of row groups) keeps the dense phases
            // where the row loop left them and advances the stage-B phases by the previous
            // unit's length; otherwise the phases are computed from u_start.
            if (U.owner == prev_owner && U.j0 == prev_j0 + prev_n && prev_n % (u32)(RG * kRowBits) == 0u) {
    #pragma unroll
                for (int i = 0; i < NBV; ++i) R.set_s(i, R.residue(i, R.s(i) + prev_n));   // < 2^14 + 2^20
            } else if (RESIDENT && U.ph_slot != kNoSlot) {
                // a served job's phases are resident: advanced from its slot by the unit's offset
                const u32* ph = A.phases + (size_t)U.ph_slot * A.ph_stride;
                sd = 0;
                if (lane < t_nd()) sd = doff * 32u + advance_phase(ph[lane], U.ph_off, t_pmod()[lane]);
                // one group at a time through opaque running pointers (16 precomputed 64-bit
                // addresses, or 16 groups' loads in flight, would spill)
                const u32* php = ph + t_nd() + lane;
                const PrimeMod* pmp = t_pmod() + t_nd() + lane;
    #pragma unroll
                for (int i = 0; i < NBV; ++i) {
                    const u32 p = (u32)i * 64u + lane;
                    R.set_s(i, p < t_nb() ? advance_phase(*php, U.ph_off, *pmp) : 0u);
                    php += 64;
                    pmp += 64;
                    asm volatile("" : "+v"(php), "+v"(pmp) : : "memory");
                }
            } else {
                sd = 0;
                if (lane < t_nd()) sd = doff * 32u + phase(t_pmod()[lane], U.a, U.a_hi, U.us_lo, U.us_hi);
                const PrimeMod* pmp = t_pmod() + t_nd() + lane;
    #pragma unroll
                for (int i = 0; i < NBV; ++i) {
                    const u32 p = (u32)i * 64u + lane;
                    R.set_s(i, p < t_nb() ? phase(*pmp, U.a, U.a_hi, U.us_lo, U.us_hi) : 0u);
                    pmp += 64;
                    asm volatile("" : "+v"(pmp) : : "memory");
                }
            }
            prev_owner = U.owner;
            prev_j0 = U.j0;
            prev_n = U.n;

            auto emit = [&](u32 j) {
                if (lane == q_n) {
                    q_owner = U.owner;
                    q_j = U.j0 + j;
                }
                if (++q_n == 64u) flush();
            };
            // The unit's dense survivors wait in the wave's candidate queue (LDS, offsets in
            // the unit) and meet stage B kQueueTest or more at a time: group 0 for all of them
            // back to back, the later groups for the few that pass it.  The queue is emptied
            // at the end of the unit, whose stage-B phases it was tested against.
            u32 cq_n = 0;
            auto test_queue = [&]() {
                const u32 n4 = (cq_n + 3u) & ~3u;
                if (lane < n4 - cq_n) cqb[cq_n + lane] = 0u;   // padding probes read candidate 0
                wave_sync();                                    // the queue writes before its reads
                for (u32 base = 0; base < n4; base += 64u) {
                    const u32 n = n4 - base < 64u ? n4 - base : 64u;
                    u64 pass = stage_b_group0<NBV>(lds, R, cqb + base, n);
                    const u32 real = cq_n - base;
                    if (real < 64u) pass &= (1ull << real) - 1u;
                    while (pass) {
                        const u32 t = (u32)__builtin_ctzll(pass);
                        pass &= pass - 1u;
                        const u32 j = cqb[base + t];
                        if (stage_b_rest<NBV>(lds, t_far(), t_nb_lds(), t_nb(), R, j)) emit(j);
                    }
                }
                wave_sync();   // the reads before the next writes
                cq_n = 0;
            };

            const u32 rows = (U.n + kRowBits - 1) / kRowBits;
            for (u32 row0 = 0; row0 < rows; row0 += RG) {
                // lane l owns candidates row*8192 + 128*l + 32*k + b (k = 0..3, b = 0..31): per
                // pattern and row it reads the two aligned 16-byte blocks at its 128-bit window
                // (ds_read_b128, the 64 lanes over 1 KiB: conflict-free) and shifts the window by
                // the pattern's word and bit offset, which are the same for the whole wave
                u32 acc[RG][kRowWords];
    #pragma unroll
                for (int r = 0; r < RG; ++r)
    #pragma unroll
                    for (int k = 0; k < kRowWords; ++k) acc[r][k] = 0;
                // the patterns by word offset within their 16-byte block (lane p: pattern p),
                // each offset's patterns in a loop of its own with no branch per pattern
                {
                    const u32 c = (sd >> 5) & 3u;
                    const bool mine = lane < t_nd();
                    dense_offset<0, RG>(lds, lane, sd, ballot(mine && c == 0u), acc);
                    dense_offset<1, RG>(lds, lane, sd, ballot(mine && c == 1u), acc);
                    dense_offset<2, RG>(lds, lane, sd, ballot(mine && c == 2u), acc);
                    dense_offset<3, RG>(lds, lane, sd, ballot(mine && c == 3u), acc);
                }
                if (lane < t_nd()) {
                    u32 s = sd - doff * 32u + dstep;
                    if (s >= dq) s -= dq;
                    sd = doff * 32u + s;
                }
                // acc holds the killed candidates; in the unit's last, partial row (uniform) the
                // candidates past its end count as killed.  A word row is tested as is (!= ~0)
                // and complemented only when some lane has a survivor in it.
    #pragma unroll
                for (int r = 0; r < RG; ++r) {
                    const u32 rowbase = (row0 + (u32)r) * (u32)kRowBits;
                    if (rowbase + (u32)kRowBits > U.n) {
    #pragma unroll
                        for (int k = 0; k < kRowWords; ++k) {
                            const u32 jb = rowbase + lane * 128u + 32u * (u32)k;
                            if (jb >= U.n) acc[r][k] = ~0u;
                            else if (U.n - jb < 32u) acc[r][k] |= ~((1u << (U.n - jb)) - 1u);
                        }
                    }
                }
                u32 cnt_rows = 0;   // this row group's dense survivors (a 32-bit scalar sum)
                if constexpr (kQueue) {
                    // survivors of the dense stage, per word row (32 candidates a lane): every
                    // lane with a live bit writes its lowest one to the queue at its rank among
                    // those lanes (the others to the discard slot), until no lane has one left
        #pragma unroll
                    for (int rk = 0; rk < RG * kRowWords; ++rk) {
                        const int r = rk / kRowWords, k = rk % kRowWords;
                        u64 m = ballot(acc[r][k] != ~0u);
                        if (!m) continue;
                        u32 live = ~acc[r][k];
                        const u32 jb = lane * 128u + ((row0 + (u32)r) * (u32)kRowBits + 32u * (u32)k);
                        do {
                            const u32 rank = __builtin_amdgcn_mbcnt_hi((u32)(m >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((u32)m, cq_n));
                            cqb[live != 0u ? rank : kQueueDiscard] = jb + (u32)__builtin_ctzg(live, 0);
                            const u32 c = (u32)__builtin_popcountll(m);
                            cq_n += c;
                            cnt_rows += c;
                            if (cq_n >= kQueueTest) test_queue();
                            live &= live - 1u;
                            m = ballot(live != 0u);
                        } while (m);
                    }
                } else {
                    // survivors of the dense stage in uniform control flow: the lanes with a
                    // word not all killed (ballot), then that word's live bits (readlane);
                    // stage B takes them two at a time, over every group of the instance
                    // (stopping at the shape's last group, as the queue form does, cost this
                    // instance 2 % at config 2; in a launch of narrower shapes the groups past
                    // a shape's own hold q = 1 lanes on a zero word, which never kill)
                    constexpr u32 kAllGroups = (u32)NBV * 64u;
                    u32 pend = 0xffffffffu;
        #pragma unroll
                    for (int rk = 0; rk < RG * kRowWords; ++rk) {
                        const int r = rk / kRowWords, k = rk % kRowWords;
                        u64 m = ballot(acc[r][k] != ~0u);
                        while (m) {
                            const u32 l = (u32)__builtin_ctzll(m);
                            m &= m - 1u;
                            u32 w = ~(u32)__builtin_amdgcn_readlane(acc[r][k], l);
                            cnt_rows += (u32)__builtin_popcount(w);
                            const u32 jb = (row0 + (u32)r) * (u32)kRowBits + l * 128u + 32u * (u32)k;
                            while (w) {
                                const u32 j = jb + (u32)__builtin_ctz(w);
                                w &= w - 1u;
                                if (pend == 0xffffffffu) {
                                    pend = j;
                                    continue;
                                }
                                bool pa, pb;
                                stage_b_pair<NBV>(lds, t_far(), t_nb_lds(), kAllGroups, R, pend, j, pa, pb);
                                if (pa) emit(pend);
                                if (pb) emit(j);
                                pend = 0xffffffffu;
                            }
                        }
                    }
                    if (pend != 0xffffffffu && stage_b_pass<NBV>(lds, t_far(), t_nb_lds(), kAllGroups, R, pend)) emit(pend);
                }
                cnt_dense += cnt_rows;
            }
            if (kQueue && cq_n) test_queue();
        }
    }
    flush();

    if (A.stats) {   // cnt_dense is uniform (counted from readlane values)
        // one atomic per block: the waves' counts meet in LDS once every wave is past its
        // last table read
        unsigned long long* acc = (unsigned long long*)lds;
        __syncthreads();
        if (tid == 0) *acc = 0;
        __syncthreads();
        if (lane == 0 && cnt_dense) atomicAdd(acc, (unsigned long long)cnt_dense);
        __syncthreads();
        if (tid == 0 && *acc) {
            atomicAdd(&A.stats[0], *acc);
            atomicAdd(&A.stats[1], *acc);
        }
    }
}

__constant__ PrimeMod c_small[18];   // 2..61
__constant__ PrimeMod c_trial[62];   // 67..409 (v13.cu:180-199, CHAIN_SCREEN_PRIMES[0..61])

__device__ __forceinline__ bool test_prime(U2 n, const ProveArgs& A, bool& ok) {
    return A.test_kind == 0 ? is_prime_bpsw(n, c_small, ok) : is_prime_mr(n, A.bases_mr, A.n_bases);
}

// ---------------------------------------------------------------------------
// K1c: deep stage
// ---------------------------------------------------------------------------

// n = a + W*u (mod 2^128) for u128 a, W, u
__device__ __forceinline__ U2 cand_n(u64 W, u64 W_hi, u64 a, u64 a_hi, u64 u_lo, u64 u_hi) {
    U2 n = mul_u64(mk(u_lo, u_hi), W);
    n.hi += u_lo * W_hi;
    return add(n, mk(a, a_hi));
}

// n's four 32-bit limbs as doubles (exact), converted once per candidate
struct Limbs {
    double l0, l1, l2, l3;
};

__device__ __forceinline__ Limbs limbs_of(U2 n) {
    Limbs L;
    L.l0 = (double)(u32)n.lo;
    L.l1 = (double)(u32)(n.lo >> 32);
    L.l2 = (double)(u32)n.hi;
    L.l3 = (double)(u32)(n.hi >> 32);
    return L;
}

// A deep prime's constants as doubles: q, the limb factors c_k = 2^(32k) mod q, 1/q
// (correctly rounded, from the table), and its bucket multiplier.
struct DeepConst {
    double q, rq, c32, c64, c96;
    u32 qi, mult;
};

__device__ __forceinline__ DeepConst deep_const(const DeepPrime& d) {
    DeepConst C;
    C.qi = d.q;
    C.mult = d.mult;
    C.q = (double)d.q;
    C.c32 = (double)d.c32;
    C.c64 = (double)d.c64;
    C.c96 = (double)d.c96;
    C.rq = d.rq;
    return C;
}

// n mod q for a deep prime (q < 2^18), all in f64: the limb fold y = l0 + l1*c32 + l2*c64 +
// l3*c96 is an integer below 2^52, so every FMA of it is exact; with 1/q correctly rounded
// the truncated quotient is floor(y/q) or one off, the FMA remainder y - qt*q is exact, and
// it lies in [-q, 2q) before the two corrections (checked in exact arithmetic for primes
// below 2^18 and y at the ends of its range and next to multiples of q).
__device__ __forceinline__ u32 deep_mod(double l0, double l1, double l2, double l3, const DeepConst& C) {
    const double y = __fma_rn(l3, C.c96, __fma_rn(l2, C.c64, __fma_rn(l1, C.c32, l0)));
    const double qt = __builtin_trunc(y * C.rq);
    int rem = (int)__fma_rn(-qt, C.q, y);
    rem += (rem >> 31) & (int)C.qi;
    const u32 ur = (u32)rem;
    return ur >= C.qi ? ur - C.qi : ur;
}

// bit r mod 128 of a deep prime's 128-bit residue mask (masked tables: only the residues
// whose bit is set can be forbidden)
__device__ __forceinline__ bool mask_bit(const uint4& m, u32 r) {
    const u32 w = (r >> 5) & 3;
    const u32 mw = w == 0 ? m.x : w == 1 ? m.y : w == 2 ? m.z : m.w;
    return (mw >> (r & 31)) & 1u;
}

// whether r is one of a deep prime's forbidden residues: its bucket of the prime's table
// (one 16-byte load), four compares
__device__ __forceinline__ bool deep_hit(const u32* __restrict__ tab, u32 r, u32 mult) {
    const uint4 v = ((const uint4*)tab)[(r * mult) >> 28];
    return (v.x == r) | (v.y == r) | (v.z == r) | (v.w == r);
}

// a stage-B buffer that overflowed: the sieve dropped the records past its capacity, so the
// launch's survivors are incomplete; the sticky status tells the caller
__device__ __forceinline__ void note_stage_b(const DeepArgs& A, u32 all) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (A.stats) atomicAdd(&A.stats[2], (unsigned long long)all);
    if (A.status && all > A.mid_cap) {
        atomicAdd(&A.status[0], 1u);
        atomicMax(&A.status[1], all);
        atomicAdd((unsigned long long*)(A.status + 2), (unsigned long long)(all - A.mid_cap));
    }
}

// The deep stage: every stage-B record against the shape's remaining filter primes, the 64
// lanes of a wave testing 64 primes at a time (lane p: prime p0 + p, its residue of n by the
// f64 limb fold, then its forbidden-residue mask and bucket table).
__global__ __launch_bounds__(256) void deep_kernel(DeepArgs A) {
    const u32 lane = lane_id();
    const u32 all = *A.mid_count;
    u32 total = all;
    if (total > A.mid_cap) total = A.mid_cap;
    note_stage_b(A, all);
    const u32 waves = gridDim.x * (blockDim.x >> 6);
    const u32 wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    u32 q_n = 0, q_owner = 0, q_j = 0;
    auto flush = [&]() {
        if (!q_n) return;
        u32 base = 0;
        if (lane == 0) base = atomicAdd(A.out_count, q_n);
        base = bcast32(base);
        if (lane < q_n && base + lane < A.out_cap) {
            const u32 slot = base + lane;
            if (A.mode == 0) {
                const ClassRec cr = A.plain.cls[q_owner];
                ccs_survivor s;
                s.base_idx = cr.base_idx;
                s.k_offset = cr.k;
                s.tile_idx = 6u * q_j + cr.rho - A.plain.off;
                s._pad = 0;
                ((ccs_survivor*)A.out)[slot] = s;
            } else {
                ccs_job_survivor s;
                s.job_idx = q_owner;
                s.u_off = q_j;
                ((ccs_job_survivor*)A.out)[slot] = s;
            }
        }
        q_n = 0;
    };
    // a record's n (as four f64 limbs) and its shape's deep tables
    auto load_rec = [&](u32 e, u32& owner, u32& j, u32& shape, Limbs& L) {
        const MidRec rec = A.mid[e];
        owner = rec.owner;
        j = rec.j;
        shape = 0;
        u64 ulo, uhi, a, a_hi = 0, W = A.W, W_hi = A.W_hi;
        if (A.mode == 0) {
            const ClassRec cr = A.plain.cls[owner];
            a = cr.a;
            ulo = A.plain.U0_lo;
            uhi = A.plain.U0_hi;
        } else {
            const JobRec jr = A.jobsrc.jobs[owner];
            a = jr.a;
            a_hi = jr.a_hi;
            ulo = jr.u0_lo;
            uhi = jr.u0_hi;
            W = jr.W;
            W_hi = jr.W_hi;
            shape = jr.shape;
        }
        add128(ulo, uhi, j);
        L = limbs_of(cand_n(W, W_hi, a, a_hi, ulo, uhi));
    };
    auto table = [&](u32 shape) -> DeepTab {
        if (A.group) return A.group[shape];
        DeepTab T;
        T.deep = A.deep;
        T.hash = A.hash;
        T.nc = A.nc;
        T.masked = A.masked;
        return T;
    };

    // records per wave batch: as many as keep kDeepFillWaves waves busy, up to kDeepBatch; a
    // launch with fewer than kDeepBatchMin per wave runs one record per wave at a time
    u32 per = A.per_wave;
    if (!per) {
        per = total / kDeepFillWaves;
        if (per < kDeepBatchMin) per = 1;
    }
    if (per > kDeepBatch) per = kDeepBatch;
    if (per < 2) {
        // one record per wave at a time (the deep shapes' few records spread over the chip)
        for (u32 e = wid; e < total; e += waves) {
            u32 owner, j, shape;
            Limbs n;
            load_rec(e, owner, j, shape, n);
            const DeepTab T = table(shape);
            bool dead = false;
            for (u32 p0 = 0; p0 < T.nc; p0 += 64) {
                const u32 p = p0 + lane;
                bool kill = false;
                if (p < T.nc) {
                    const DeepPrime dp = T.deep[p];
                    uint4 m = make_uint4(0u, 0u, 0u, 0u);
                    if (T.masked) m = ((const uint4*)(T.hash + (size_t)T.nc * kDeepTableWords))[p];
                    const DeepConst C = deep_const(dp);
                    const u32 r = deep_mod(n.l0, n.l1, n.l2, n.l3, C);
                    if (!T.masked || mask_bit(m, r)) kill = deep_hit(T.hash + (size_t)p * kDeepTableWords, r, C.mult);
                }
                if (ballot(kill)) {
                    dead = true;
                    break;
                }
            }
            if (dead) continue;
            // survivors wait in registers (lane k holds the k-th) and go out 64 per atomic
            if (lane == q_n) {
                q_owner = owner;
                q_j = j;
            }
            if (++q_n == 64u) flush();
        }
        flush();
        return;
    }

    // Batches (the shallow shapes, 10^5..10^7 records per launch): a
    // wave takes a batch of up to kDeepBatch consecutive records (lane k computes record k's
    // limbs and leaves them in the wave's LDS rows).  Per group of 64 deep primes (lane p:
    // prime p0 + p) the constants, the 1/q rebuild and the masks are done once for the whole
    // batch, and each record still alive is tested against them, two at a time (their limbs
    // read back by every lane, both table loads in flight together), one ballot each.
    // A batch of a group launch may hold records of several shapes; they run shape by shape.
    // The records left alive are final survivors, written with one atomic per wave.
    __shared__ double2 rows[4][kDeepBatch][2];   // per wave of the 256-thread block: limbs of its records
    double2(*const my)[2] = rows[threadIdx.x >> 6];
    for (u32 e0 = wid * per; e0 < total; e0 += waves * per) {
        const bool have = lane < per && e0 + lane < total;
        u32 owner = 0, j = 0, shape = 0;
        if (have) {
            Limbs L;
            load_rec(e0 + lane, owner, j, shape, L);
            my[lane][0] = make_double2(L.l0, L.l1);
            my[lane][1] = make_double2(L.l2, L.l3);
        }
        wave_sync();   // the rows are written before any lane reads them
        u64 pending = ballot(have), alive = 0;
        while (pending) {
            const u32 sh = (u32)__builtin_amdgcn_readlane((int)shape, (int)__builtin_ctzll(pending));
            const u64 batch = A.group ? pending & ballot(shape == sh) : pending;
            const DeepTab T = table(sh);
            pending &= ~batch;
            u64 live = batch;
            for (u32 p0 = 0; p0 < T.nc && live; p0 += 64) {
                const u32 p = p0 + lane;
                const bool mine = p < T.nc;
                DeepPrime dp = {1u, 0u, 0u, 0u, 1.0, 1u, 0u};
                uint4 m = make_uint4(0u, 0u, 0u, 0u);
                if (mine) {
                    dp = T.deep[p];
                    if (T.masked) m = ((const uint4*)(T.hash + (size_t)T.nc * kDeepTableWords))[p];
                }
                const DeepConst C = deep_const(dp);
                const u32* h = T.hash + (size_t)(mine ? p : 0u) * kDeepTableWords;
                for (u64 rem = live; rem;) {
                    const u32 k0 = (u32)__builtin_ctzll(rem);
                    rem &= rem - 1u;
                    const u32 k1 = rem ? (u32)__builtin_ctzll(rem) : k0;   // (a lone record twice)
                    rem &= rem - 1u;
                    const double2 a0 = my[k0][0], a1 = my[k0][1], b0 = my[k1][0], b1 = my[k1][1];
                    const u32 ra = deep_mod(a0.x, a0.y, a1.x, a1.y, C);
                    const u32 rb = deep_mod(b0.x, b0.y, b1.x, b1.y, C);
                    const bool pa = mine && (!T.masked || mask_bit(m, ra));
                    const bool pb = mine && (!T.masked || mask_bit(m, rb));
                    bool ka = false, kb = false;
                    if (pa) ka = deep_hit(h, ra, C.mult);
                    if (pb) kb = deep_hit(h, rb, C.mult);
                    if (ballot(ka)) live &= ~(1ull << k0);
                    if (ballot(kb)) live &= ~(1ull << k1);
                }
            }
            alive |= live;
        }
        wave_sync();   // every lane is past its row reads before the next batch writes them
        if (!alive) continue;
        u32 base = 0;
        if (lane == 0) base = atomicAdd(A.out_count, (u32)__builtin_popcountll(alive));
        base = bcast32(base);
        const u32 slot = base + lanes_below(alive);
        if (((alive >> lane) & 1ull) && slot < A.out_cap) {
            if (A.mode == 0) {
                const ClassRec cr = A.plain.cls[owner];
                ccs_survivor sv;
                sv.base_idx = cr.base_idx;
                sv.k_offset = cr.k;
                sv.tile_idx = 6u * j + cr.rho - A.plain.off;
                sv._pad = 0;
                ((ccs_survivor*)A.out)[slot] = sv;
            } else {
                ccs_job_survivor sv;
                sv.job_idx = owner;
                sv.u_off = j;
                ((ccs_job_survivor*)A.out)[slot] = sv;
            }
        }
    }
}

// A shape without deep primes: its stage-B survivors are final, so one lane per entry
// writes the output record (the deep kernel would spend a wave on each, in turn); each
// wave takes its slots with one atomic.
__global__ __launch_bounds__(256) void emit_kernel(DeepArgs A) {
    const u32 lane = lane_id();
    const u32 all = *A.mid_count;
    u32 total = all;
    if (total > A.mid_cap) total = A.mid_cap;
    note_stage_b(A, all);
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 e0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); e0 < total; e0 += stride) {
        const u32 e = e0 + lane;
        const bool live = e < total;
        const u64 act = ballot(live);
        u32 base = 0;
        if (lane == 0) base = atomicAdd(A.out_count, (u32)__builtin_popcountll(act));
        base = bcast32(base);
        if (!live) continue;
        const u32 slot = base + lanes_below(act);
        if (slot >= A.out_cap) continue;
        const MidRec rec = A.mid[e];
        if (A.mode == 0) {
            const ClassRec cr = A.plain.cls[rec.owner];
            ccs_survivor s;
            s.base_idx = cr.base_idx;
            s.k_offset = cr.k;
            s.tile_idx = 6u * rec.j + cr.rho - A.plain.off;
            s._pad = 0;
            ((ccs_survivor*)A.out)[slot] = s;
        } else {
            ccs_job_survivor s;
            s.job_idx = rec.owner;
            s.u_off = rec.j;
            ((ccs_job_survivor*)A.out)[slot] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// K2: prove kernel — GMP confirmation engine per survivor
// ---------------------------------------------------------------------------

__device__ __forceinline__ U2 survivor_n(const ProveArgs& A, u32 idx) {
    if (A.mode == 0) {
        const ccs_survivor s = ((const ccs_survivor*)A.surv)[idx];
        U2 tile = mk(A.T0_lo, A.T0_hi);
        tile = add(tile, mk(s.tile_idx, 0));
        const U2 k = add(mul_u64(tile, A.WP), mk(s.k_offset, 0));
        return add(mul_u64(k, A.M), mk(A.bases[s.base_idx], 0));
    }
    const ccs_job_survivor s = ((const ccs_job_survivor*)A.surv)[idx];
    const JobRec jr = A.jobs[s.job_idx];
    u64 ulo = jr.u0_lo, uhi = jr.u0_hi;
    add128(ulo, uhi, s.u_off);
    return cand_n(jr.W, jr.W_hi, jr.a, jr.a_hi, ulo, uhi);
}

// private flag of a K2a record whose n passed the strong MR to base 2: K2b runs its Lucas
// test (lanes that all need one) and clears the flag, or the prime flag with it
constexpr u32 kLucasPending = 1u << 31;

// K2a: per survivor the gates of prover_worker (v13.cu:1828-1842): bit size, trial division
// by 67..409, primality of n.  Composites are final here; primes go to the chain queue.
// Under Baillie-PSW this kernel runs the strong MR to base 2 only: a wave of survivors holds
// a few primes (~1 in 7 at 89 bits), and running their Lucas tests here would hold every
// lane of the wave for them; the base-2 probable primes are queued and their Lucas tests
// run in K2b, one per lane.
__global__ __launch_bounds__(128) void prove_kernel(ProveArgs A) {
    u32 total = *A.surv_count;
    if (total > A.surv_cap) total = A.surv_cap;
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const U2 n = survivor_n(A, idx);
        ccs_chain_result res;
        res.survivor_idx = idx;
        res.chain_len = 0;
        res.back_steps = 0;
        res.flags = 0;
        res.root_lo = 0;
        res.root_hi = 0;
        bool ok = true, queue = false;
        do {
            if (A.bits && bitlen(n) != A.bits) break;
            res.flags |= CCS_F_BITS_OK;
            bool div = false;
            for (int i = 0; i < 62; ++i)
                if (mod128(n.lo, n.hi, c_trial[i]) == 0) { div = true; break; }
            if (div) break;
            res.flags |= CCS_F_TRIAL_OK;
            if (n.hi >> 63) { res.flags |= CCS_F_OVERFLOW; break; }
            if (A.test_kind == 0) {
                const int pre = bpsw_pre(n, c_small);
                if (pre == 0) break;
                res.flags |= CCS_F_PRIME | (pre == 2 ? kLucasPending : 0u);
            } else {
                if (!test_prime(n, A, ok)) break;
                res.flags |= CCS_F_PRIME;
            }
            queue = true;
        } while (0);
        if (!ok) res.flags |= CCS_F_UNSUPPORTED;
        ((ccs_chain_result*)A.out)[idx] = res;
        if (queue) {
            const u32 q = atomicAdd(A.chain_count, 1u);
            if (q < A.surv_cap) A.chain_queue[q] = idx;
        }
    }
}

// forward chain positions tested at once (lanes 16 .. 16 + kFwdWindow - 1): a test's
// latency grows with the value's bit length, so the window stops at position 23 (chains
// longer than that, from n, continue one value at a time)
constexpr int kFwdWindow = 24;

// this lane's chain value: position lane - 16 of the chain through n (lanes 0..15 the
// backward values (v-1)/2, lanes 16..39 the forward values 2v+1, lanes 40..63 none); false
// when the value is undefined (the backward walk reaches 1), not below 2^127 or not tested
__device__ __forceinline__ bool chain_value(u32 lane, U2& v) {
    if (lane < 16) {
        const int k = 16 - (int)lane;          // steps back
        for (int t = 0; t < k; ++t) {
            if (!ge(v, mk(2, 0))) return false;
            v = shr1(sub(v, mk(1, 0)));
        }
        return ge(v, mk(2, 0));
    }
    const int i = (int)lane - 16;              // steps forward
    if (i >= kFwdWindow) return false;
    for (int t = 0; t < i; ++t) {
        if (v.hi >> 62) return false;
        v = add(shl1(v), mk(1, 0));
    }
    return true;
}

// the chain of a prime n from the lane masks of its window (pm: prime, vm: value defined,
// okm: test supported; lane l = position l - 16), continued one value at a time where the
// walk leaves the window
__device__ __forceinline__ void chain_readoff(const ProveArgs& A, U2 n, u64 pm, u64 vm, u64 okm,
                                       ccs_chain_result& res) {
    bool ok = true;
    // backward: consecutive primes at 1, 2, ... steps back (lanes 15, 14, ...)
    u32 back = 0;
    while (back < 16 && ((pm >> (15 - back)) & 1)) ++back;
    U2 root = n;
    for (u32 t = 0; t < back; ++t) root = shr1(sub(root, mk(1, 0)));
    if (back == 16) {   // longer walk: continue one value at a time (cap: 65 steps)
        for (;;) {
            if (back >= 65) break;
            const U2 pred = shr1(sub(root, mk(1, 0)));
            if (!ge(pred, mk(2, 0))) break;
            if (!test_prime(pred, A, ok)) break;
            root = pred;
            ++back;
        }
    }
    // forward from n: consecutive primes at 0, 1, ... steps (lanes 16, 17, ...)
    u32 fwd = 0;
    bool over = false;
    while (fwd < (u32)kFwdWindow && ((pm >> (16 + fwd)) & 1)) ++fwd;
    u32 len = back + fwd;
    if (len > 64) len = 64;
    if (fwd < (u32)kFwdWindow && len < 64 && !((vm >> (16 + fwd)) & 1)) over = true;   // next value >= 2^127
    if (fwd == (u32)kFwdWindow && len < 64) {   // longer chain: continue one value at a time
        U2 cur = n;
        for (u32 t = 0; t + 1 < (u32)kFwdWindow; ++t) cur = add(shl1(cur), mk(1, 0));
        while (len < 64) {
            if (cur.hi >> 62) { over = true; break; }
            const U2 nx = add(shl1(cur), mk(1, 0));
            if (!test_prime(nx, A, ok)) break;
            ++len;
            cur = nx;
        }
    }
    // lanes on the walk: the backward run and its first composite, n's forward run and its
    // first composite; only their tests count for the unsupported flag
    const u32 blo = back < 16 ? 15 - back : 0;
    const u32 fhi = 16 + fwd < 63 ? 16 + fwd : 63;
    const u64 walk = (fhi >= 63 ? ~0ull : ((1ull << (fhi + 1)) - 1)) & ~((1ull << blo) - 1);
    if (back) res.flags |= CCS_F_NONROOT;
    if (over) res.flags |= CCS_F_OVERFLOW;
    if ((~okm & walk) || !ok) res.flags |= CCS_F_UNSUPPORTED;
    res.root_lo = root.lo;
    res.root_hi = root.hi;
    res.chain_len = len;
    res.back_steps = back;
}

// K2b for large batches: one lane per prime survivor walks its chain one value at a time
// (backward walk to the true root, then forward follow from n), stopping at the first
// composite on each side, as prover_worker does (v13.cu:1849-1881).  Most chains end
// within a few values, so this does a few tests per survivor where a 64-lane window does
// 40; the lanes of a wave diverge only in how far their chains go.
__global__ __launch_bounds__(256) void chain_kernel(ProveArgs A) {
    u32 total = *A.chain_count;
    if (total > A.surv_cap) total = A.surv_cap;
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const u32 idx = A.chain_queue[e];
        const U2 n = survivor_n(A, idx);
        ccs_chain_result res = ((ccs_chain_result*)A.out)[idx];
        bool ok = true;
        if (res.flags & kLucasPending) {   // n passed the base-2 MR: its Lucas test decides
            res.flags &= ~kLucasPending;
            bool okn = true;
            const bool prime = bpsw_lucas(n, okn);
            if (!okn) res.flags |= CCS_F_UNSUPPORTED;
            if (!prime) {
                res.flags &= ~CCS_F_PRIME;
                ((ccs_chain_result*)A.out)[idx] = res;
                continue;
            }
        }
        // backward: consecutive primes at 1, 2, ... steps back (at most 65)
        u32 back = 0;
        U2 root = n;
        while (back < 65) {
            const U2 pred = shr1(sub(root, mk(1, 0)));
            if (!ge(pred, mk(2, 0))) break;
            if (!test_prime(pred, A, ok)) break;
            root = pred;
            ++back;
        }
        // forward from n (prime): consecutive primes at 1, 2, ... steps, length <= 64
        u32 len = back + 1 > 64 ? 64 : back + 1;
        bool over = false;
        U2 cur = n;
        while (len < 64) {
            if (cur.hi >> 62) { over = true; break; }   // the next value would reach 2^127
            const U2 nx = add(shl1(cur), mk(1, 0));
            if (!test_prime(nx, A, ok)) break;
            ++len;
            cur = nx;
        }
        if (back) res.flags |= CCS_F_NONROOT;
        if (over) res.flags |= CCS_F_OVERFLOW;
        if (!ok) res.flags |= CCS_F_UNSUPPORTED;
        res.root_lo = root.lo;
        res.root_hi = root.hi;
        res.chain_len = len;
        res.back_steps = back;
        ((ccs_chain_result*)A.out)[idx] = res;
    }
}

// K2 for few survivors, as one kernel: one wave per survivor does the gates of K2a (the
// trial primes spread over the lanes) and, with every lane testing its chain position at
// once (lane 16 testing n itself), the walk of K2b.  Its latency is one primality test,
// and the prove of a batch is a single launch, so it fits beside the next filter launch.
__global__ __launch_bounds__(256) void prove_wave_kernel(ProveArgs A) {
    const u32 lane = lane_id();
    u32 total = *A.surv_count;
    if (total > A.surv_cap) total = A.surv_cap;
    const u32 waves = gridDim.x * (blockDim.x >> 6);
    const u32 wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (u32 idx = wid; idx < total; idx += waves) {
        const U2 n = survivor_n(A, idx);
        ccs_chain_result res;
        res.survivor_idx = idx;
        res.chain_len = 0;
        res.back_steps = 0;
        res.flags = 0;
        res.root_lo = 0;
        res.root_hi = 0;
        bool pass = false;
        if (!A.bits || bitlen(n) == A.bits) {
            res.flags |= CCS_F_BITS_OK;
            const bool div = lane < 62 && mod128(n.lo, n.hi, c_trial[lane]) == 0;
            if (!ballot(div)) {
                res.flags |= CCS_F_TRIAL_OK;
                if (n.hi >> 63) res.flags |= CCS_F_OVERFLOW;
                else pass = true;
            }
        }
        if (!pass) {   // uniform
            if (lane == 0) ((ccs_chain_result*)A.out)[idx] = res;
            continue;
        }
        U2 v = n;
        const bool valid = chain_value(lane, v);
        bool ok = true;
        const bool prime = valid && test_prime(v, A, ok);
        const u64 pm = ballot(prime);
        const u64 vm = ballot(valid);
        const u64 okm = ballot(ok);
        if (lane != 0) continue;
        if (!((pm >> 16) & 1)) {   // n composite: the gate's answer
            if (!((okm >> 16) & 1)) res.flags |= CCS_F_UNSUPPORTED;
            ((ccs_chain_result*)A.out)[idx] = res;
            continue;
        }
        res.flags |= CCS_F_PRIME;
        chain_readoff(A, n, pm, vm, okm, res);
        ((ccs_chain_result*)A.out)[idx] = res;
    }
}


// K1.5: top-position gate of family survivors (v16_k15_topmr_kernel, v16...cu:2185-2296):
// p_top = (W*Q) << top_pos - 1 with W = D*2^E must pass strong MR to the first `reps` bases
// of {2, 3, 5}; the passing survivors are compacted into the output.
__global__ __launch_bounds__(256) void top_gate_kernel(TopGateArgs A) {
    u32 n = *A.in_count;
    if (n > A.in_cap) n = A.in_cap;
    const u32 bases[3] = {2u, 3u, 5u};
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const ccs_job_survivor sv = ((const ccs_job_survivor*)A.in)[i];
        const JobRec jr = A.jobs[sv.job_idx];
        u64 qlo = jr.u0_lo, qhi = jr.u0_hi;
        add128(qlo, qhi, sv.u_off);
        U2 t = add(cand_n(jr.W, jr.W_hi, jr.a, jr.a_hi, qlo, qhi), mk(1, 0));   // p_0 + 1 = W*Q
        if (A.top_pos > 0) {
            t.hi = (t.hi << A.top_pos) | (t.lo >> (64 - A.top_pos));
            t.lo <<= A.top_pos;
        }
        const U2 p_top = sub(t, mk(1, 0));
        if (is_prime_mr(p_top, bases, A.reps)) {
            const u32 slot = atomicAdd(A.out_count, 1u);
            if (slot < A.out_cap) ((ccs_job_survivor*)A.out)[slot] = sv;
        }
    }
}

// The v13 K2 output (cc18_prove_kernel, v13.cu:1462-1538): one ChainResult per prime
// survivor that is a chain root ((n-1)/2 not prime: back_steps 0), with the forward length
// from it; compacted a wave at a time (one atomic per wave); prime survivors counted.
__global__ __launch_bounds__(256) void chain_roots_kernel(RootsArgs A) {
    u32 n = *A.count;
    if (n > A.cap) n = A.cap;
    const u32 lane = lane_id();
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 e0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); e0 < n; e0 += stride) {
        const u32 i = e0 + lane;
        bool prime = false, root = false;
        ccs_chain_result r;
        if (i < n) {
            r = A.res[i];
            prime = (r.flags & CCS_F_PRIME) != 0;
            root = prime && r.back_steps == 0;
        }
        const u64 pm = ballot(prime), rm = ballot(root);
        u32 base = 0;
        if (lane == 0) {
            if (pm && A.primes) atomicAdd(A.primes, (u32)__builtin_popcountll(pm));
            if (rm) base = atomicAdd(A.out_count, (u32)__builtin_popcountll(rm));
        }
        base = bcast32(base);
        if (!root) continue;
        const u32 slot = base + lanes_below(rm);
        if (slot >= A.out_cap) continue;
        const ccs_survivor sv = A.surv[i];
        ccs_chain_root o;
        o.base_idx = sv.base_idx;
        o.k_offset = sv.k_offset;
        o.tile_idx = sv.tile_idx;
        o.chain_len = r.chain_len;
        A.out[slot] = o;
    }
}

// resident phases of newly admitted served jobs: one block per job, a thread per prime
__global__ __launch_bounds__(256) void phase_init_kernel(const PhaseInit* jobs, u32 n_jobs, u32* phases,
                                                         u32 stride) {
    for (u32 j = blockIdx.x; j < n_jobs; j += gridDim.x) {
        const PhaseInit J = jobs[j];
        u32* out = phases + (size_t)J.slot * stride;
        for (u32 p = threadIdx.x; p < J.n; p += blockDim.x) out[p] = phase(J.pmod[p], J.a, J.a_hi, J.u_lo, J.u_hi);
    }
}

// a no-allocation launch whose stage-B buffer overflowed reports it through its count
__global__ void mark_overflow_kernel(const u32* status, u32* count) {
    if (threadIdx.x == 0 && status[0]) *count = 0xffffffffu;
}

__global__ void zero_kernel(u32* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// up to three short word ranges in one launch (a launch's counters, statistics and output
// count: one fill kernel instead of three in front of every sieve)
__global__ void zero3_kernel(u32* a, u32 na, u32* b, u32 nb, u32* c, u32 nc) {
    const u32 i = threadIdx.x;
    if (a && i < na) a[i] = 0;
    if (b && i < nb) b[i] = 0;
    if (c && i < nc) c[i] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <int RG, int NBV>
static hipError_t launch_sieve_t(const SieveArgs& a, int grid, int block, size_t lds, hipStream_t s) {
    (void)hipGetLastError();   // launch errors only: drop any earlier non-launch error
    if (a.group && a.phases)
        hipLaunchKernelGGL((sieve_kernel<RG, NBV, true, true>), dim3(grid), dim3(block), lds, s, a);
    else if (a.group)
        hipLaunchKernelGGL((sieve_kernel<RG, NBV, true, false>), dim3(grid), dim3(block), lds, s, a);
    else
        hipLaunchKernelGGL((sieve_kernel<RG, NBV, false, false>), dim3(grid), dim3(block), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_sieve(const SieveArgs& a, int rg, int nbv, int grid, int block, size_t lds_bytes,
                        hipStream_t s) {
#define CCS_SIEVE_CASE(R, B) \
    if (rg == R && nbv == B) return launch_sieve_t<R, B>(a, grid, block, lds_bytes, s);
#define CCS_SIEVE_ROW(R)                                                                      \
    CCS_SIEVE_CASE(R, 1) CCS_SIEVE_CASE(R, 2) CCS_SIEVE_CASE(R, 3) CCS_SIEVE_CASE(R, 4)      \
    CCS_SIEVE_CASE(R, 5) CCS_SIEVE_CASE(R, 6) CCS_SIEVE_CASE(R, 7) CCS_SIEVE_CASE(R, 8)      \
    CCS_SIEVE_CASE(R, 12) CCS_SIEVE_CASE(R, 16)
    // more than 8 stage-B groups: the next instance up (lanes past the last prime are padding)
    nbv = (int)sieve_stage_b_width((u32)nbv * 64u);
    CCS_SIEVE_ROW(1) CCS_SIEVE_ROW(2) CCS_SIEVE_ROW(4)
#undef CCS_SIEVE_ROW
#undef CCS_SIEVE_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_deep(const DeepArgs& a, u32 max_entries, hipStream_t s) {
    (void)hipGetLastError();
    // up to 65536 waves: with a few entries per wave the dispatcher keeps CUs that are
    // shared with other work (a concurrent prove) from holding back the end of the launch
    if (a.nc == 0 && !a.group) {   // no deep primes: every stage-B survivor is final
        u32 eblocks = (max_entries + 255) / 256;
        if (eblocks < 1) eblocks = 1;
        if (eblocks > 4096) eblocks = 4096;
        hipLaunchKernelGGL(emit_kernel, dim3(eblocks), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    u32 blocks = (max_entries + 3) / 4;
    if (blocks < 1) blocks = 1;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(deep_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_prove(const ProveArgs& a, u32 max_entries, hipStream_t s) {
    (void)hipGetLastError();
    if (max_entries <= kProveWaveMax) {
        // one wave per survivor (grid-stride beyond 1024 waves), blocks of 4 waves: beside
        // a running filter launch, a CU that holds prove waves slows what else runs there
        // (the deep stage), so they are spread thin (16-wave blocks: 5.64 ms per config-2
        // step, 4-wave blocks: 5.46 ms)
        const u32 waves = max_entries < 1024u ? (max_entries ? max_entries : 1u) : 1024u;
        hipLaunchKernelGGL(prove_wave_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    u32 blocks = (max_entries + 127) / 128;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, a.chain_count, (size_t)1);
    hipLaunchKernelGGL(prove_kernel, dim3(blocks), dim3(128), 0, s, a);
    u32 cblocks = (max_entries + 255) / 256;
    if (cblocks < 1) cblocks = 1;
    if (cblocks > 2048) cblocks = 2048;
    hipLaunchKernelGGL(chain_kernel, dim3(cblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_top_gate(const TopGateArgs& a, hipStream_t s) {
    (void)hipGetLastError();
    u32 blocks = (a.in_cap + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, a.out_count, (size_t)1);
    hipLaunchKernelGGL(top_gate_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_phase_init(const PhaseInit* d_jobs, u32 n_jobs, u32* phases, u32 stride, hipStream_t s) {
    (void)hipGetLastError();
    if (!n_jobs) return hipSuccess;
    const u32 blocks = n_jobs < 1024u ? n_jobs : 1024u;
    hipLaunchKernelGGL(phase_init_kernel, dim3(blocks), dim3(256), 0, s, d_jobs, n_jobs, phases, stride);
    return hipGetLastError();
}

hipError_t launch_chain_roots(const RootsArgs& a, hipStream_t s) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(zero3_kernel, dim3(1), dim3(256), 0, s, a.out_count, 1u, a.primes, a.primes ? 1u : 0u,
                       (u32*)nullptr, 0u);
    u32 blocks = (a.cap + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(chain_roots_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mark_overflow(const u32* status, u32* count, hipStream_t s) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(mark_overflow_kernel, dim3(1), dim3(64), 0, s, status, count);
    return hipGetLastError();
}

hipError_t launch_zero(u32* p, size_t n_words, hipStream_t s) {
    (void)hipGetLastError();
    u32 blocks = (u32)((n_words + 255) / 256);
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, p, n_words);
    return hipGetLastError();
}

hipError_t launch_zero3(u32* a, u32 na, u32* b, u32 nb, u32* c, u32 nc, hipStream_t s) {
    if (na > 256 || nb > 256 || nc > 256) return hipErrorInvalidValue;
    (void)hipGetLastError();
    hipLaunchKernelGGL(zero3_kernel, dim3(1), dim3(256), 0, s, a, na, b, nb, c, nc);
    return hipGetLastError();
}

hipError_t upload_prove_constants(const PrimeMod* small18, const PrimeMod* trial62) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_small), small18, sizeof(PrimeMod) * 18);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_trial), trial62, sizeof(PrimeMod) * 62);
}

}  // namespace ccs
