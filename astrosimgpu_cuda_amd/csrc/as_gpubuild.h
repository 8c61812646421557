// as_gpubuild.h -- connectivity generation in parallel, draw for draw equal to
// the sequential setup stream of src/astrosim_fused.cu:631-735.
//
// The reference draws every synapse of the network from ONE sequential
// counter stream: for each presynaptic row a geometric skip (S draw) per
// candidate synapse, then on excitatory rows a third-factor coin (T draw)
// and, on success, a pool pick (P draw). Where a row starts depends on how
// many draws every earlier row took, so the rows cannot be walked
// independently as they stand. This file walks them in parallel anyway:
//
// 1. Every counter's draw is decoded once into a 32-bit entry (table_fill):
//    the skip value, the coin and the pool slot, the three ways a walk may
//    read it. Skips whose fast evaluation is too close to an integer to
//    floor safely are resolved on the host with the C library's log1p, as
//    the reference evaluates them.
// 2. Chains: one lane per anchor counter a_j = r0 + j*W walks rows as if a
//    row started at a_j, recording every row start up to a_{j+1} (chain_a),
//    then keeps walking until it reaches a row start that the next chains
//    recorded (chain_b). Two walks at the same row start continue
//    identically, so once coupled the chains describe one walk; stitching
//    the chains from the true first row start gives the candidate list of
//    row starts (path_write).
// 3. The chains do not know their row index, and an autapse (the row's own
//    neuron drawn as a target: no synapse and no coin draw) depends on it.
//    verify_row flags every row of the candidate list with a drawn target
//    within +-K of its own position; repair_row walks such a row exactly for
//    each possible index shift until it rejoins the list. The host composes
//    the repairs that apply in order (the shift changes only at repairs).
// 4. count_row and emit_row walk every row exactly from its final start with
//    its true index. count_row also checks that each row ends where the next
//    begins: together with the exact first start this proves the whole list
//    equals the sequential walk, whatever the speculation did.
//
// The bodies below are plain host/device functions over raw pointers; the
// same code runs as GPU kernels (as_gpubuild.hip) and in a host emulation.
#pragma once

#include <cstddef>
#include <cstdint>

#include "as_dmath.h"
#include "as_model.h"
#include "as_rng.h"

namespace astrosim {
namespace gb {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i64 = std::int64_t;

// Entry layout: bits 0-27 skip (saturated), bit 28 coin, bits 29-31 slot.
constexpr u32 kGMask = (1u << 28) - 1;
constexpr u32 kCoin = 1u << 28;
constexpr int kSlotShift = 29;
constexpr u64 kNoSelf = ~u64(0);
constexpr u32 kNone = 0xFFFFFFFFu;  // no own target position

AS_HD u32 atomic_add_u32(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
AS_HD void atomic_or_u32(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
#endif
}

// u = uniform() as geometric_skip reads it (src/astrosim_fused.cu:619-620).
AS_HD double skip_uniform(double r) { return r >= 1.0 ? 0x1.fffffffffffffp-1 : r; }

// floor(log1p(-u) / log1m_p) without the C library: u lies on the 2^-53 grid,
// so 1 - u is exact and log1p(-u) = log(1 - u); as_log is within a few ulps
// and the C library's log1p within one, so the two quotients differ by less
// than 2^-46 of their size. When the fast quotient is farther than that from
// an integer both floors agree (returns true, g saturated at 2^32 - 1);
// otherwise only the C library can decide (returns false).
AS_HD bool skip_fast(double u, double log1m_p, u32& g) {
    const double q = as_log(1.0 - u) / log1m_p;
    const double margin = __builtin_fabs(q) * 0x1p-46 + 0x1p-1000;
    const double lo = __builtin_floor(q - margin);
    if (lo != __builtin_floor(q + margin)) return false;
    g = lo >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<u32>(lo);
    return true;
}

// The three readings of one draw r (src/astrosim_fused.cu:617-624 skip,
// :703 coin, :705 pool pick over `take` pool entries).
AS_HD u32 encode_entry(double r, u32 g, u32 take) {
    const u32 gs = g < kGMask ? g : kGMask;
    const u32 coin = r < ScalingConfig::kPThird ? kCoin : 0u;
    const u32 slot = static_cast<u32>(static_cast<u64>(r * static_cast<double>(take)) % take);
    return gs | coin | (slot << kSlotShift);
}

// One generation pass: rows of presynaptic neurons over a block of `pc`
// targets. Passes 0 and 1 are the excitatory rows over the excitatory and
// the inhibitory targets (with third-factor draws), pass 2 the inhibitory
// rows over all neurons (src/astrosim_fused.cu:688-723). A row's own neuron
// sits at target position self_base + row index (kNone: never).
struct PassDesc {
    u32 pc;
    u32 tf;
    u32 n_rows;
    u32 post_offset;
    u32 self_base;  // kNone: no row meets its own neuron
};

// Walk state: next counter, position in the row (0 exactly at a row start
// with phase 0), expected draw (0 skip, 1 coin, 2 pool pick), last target.
struct Walker {
    u64 c;
    u32 pos, phase, last, pad;
};

// Event flags of one draw.
enum : int { kEvEdge = 1, kEvTf = 2, kEvRowEnd = 4 };

// Consumes the draw at w.c (entry e) and returns its events. A skip of at
// least the remaining targets ends the row (src/astrosim_fused.cu:621-623,
// 694); so does a synapse on the last target, once its coin (and pick) are
// drawn, because the row loop's condition then fails (:692); a synapse onto
// the row's own target position `self` (kNone: none) is dropped without a
// coin draw (:697). Written with selects only: lanes of a wave in different
// phases then run one instruction stream.
AS_HD int walker_step(Walker& w, u32 e, const PassDesc& P, u32 self) {
    ++w.c;
    const u32 g = e & kGMask;
    const bool s = w.phase == 0, t = w.phase == 1;
    const bool end_s = s && g >= P.pc - w.pos;
    const u32 last = w.pos + g;
    const bool hit = s && !end_s;
    const bool edge = hit && last != self;
    const u32 phase = edge ? P.tf : ((t && (e & kCoin) != 0) ? 2u : 0u);
    const u32 pos = hit ? last + 1 : w.pos;
    const bool end = end_s || (phase == 0 && pos == P.pc);
    w.pos = end ? 0u : pos;
    w.last = hit ? last : w.last;
    w.phase = phase;
    return (edge ? kEvEdge : 0) | ((!s && !t) ? kEvTf : 0) | (end ? kEvRowEnd : 0);
}

// w = go ? nw : w, field by field (a whole-struct select is kept in memory).
AS_HD void commit(Walker& w, const Walker& nw, bool go) {
    w.c = go ? nw.c : w.c;
    w.pos = go ? nw.pos : w.pos;
    w.phase = go ? nw.phase : w.phase;
    w.last = go ? nw.last : w.last;
}

// Decoded draws for counters [base, base + len).
struct Table {
    const u32* e;
    u64 base, len;
};

AS_HD bool table_get(const Table& t, u64 c, u32& e) {
    const u64 k = c - t.base;
    if (c < t.base || k >= t.len) return false;
    e = t.e[k];
    return true;
}

// Sixteen table entries: a walk reads the table one chunk at a time, the
// next chunk requested while the current one is consumed, so that a lane
// walking alone pays one memory latency per sixteen draws, not per draw.
struct alignas(16) Quad {
    u32 v[4];
};
constexpr u64 kChunk = 16;

AS_HD void load_chunk(const Table& t, u64 chunk, Quad (&q)[4]) {
    const Quad* p = reinterpret_cast<const Quad*>(t.e) + chunk * 4;
    q[0] = p[0];
    q[1] = p[1];
    q[2] = p[2];
    q[3] = p[3];
}

// Calls step(e) with the entries of counters c, c + 1, ... while it returns
// true. Returns false when the table ends first. The table's storage holds
// whole chunks (its length rounded up to kChunk, plus one chunk). The walk
// leaves the unrolled chunk body only at its end (a `live` flag guards each
// step), which keeps the compiled loop free of per-step exit paths.
template <class Step>
AS_HD bool drive(const Table& t, u64 c, Step&& step) {
    if (c < t.base) return false;
    const u64 k = c - t.base;
    if (k >= t.len) return false;
    u64 chunk = k / kChunk;
    u32 first = static_cast<u32>(k % kChunk);
    Quad cur[4];
    load_chunk(t, chunk, cur);
    bool live = true;
    for (;;) {
        // here chunk * kChunk < len, so chunk + 1 lies inside the storage
        Quad nxt[4];
        load_chunk(t, chunk + 1, nxt);
        const u64 k0 = chunk * kChunk;
        bool ran_out = false;
        // written out sixteen times (constant indices keep cur in registers)
#define AS_GB_STEP(j)                                      \
    if (live && (j) >= first) {                            \
        if (k0 + (j) < t.len) {                            \
            live = step(cur[(j) / 4].v[(j) % 4]);          \
        } else {                                           \
            live = false;                                  \
            ran_out = true;                                \
        }                                                  \
    }
        AS_GB_STEP(0) AS_GB_STEP(1) AS_GB_STEP(2) AS_GB_STEP(3)
        AS_GB_STEP(4) AS_GB_STEP(5) AS_GB_STEP(6) AS_GB_STEP(7)
        AS_GB_STEP(8) AS_GB_STEP(9) AS_GB_STEP(10) AS_GB_STEP(11)
        AS_GB_STEP(12) AS_GB_STEP(13) AS_GB_STEP(14) AS_GB_STEP(15)
#undef AS_GB_STEP
        if (!live) return !ran_out;
        ++chunk;
        first = 0;
        if (chunk * kChunk >= t.len) return false;
#pragma unroll
        for (int j = 0; j < 4; ++j) cur[j] = nxt[j];
    }
}

// One bit per counter from `base`, `bits` of them (whole 32-bit words, the
// unused tail zero); base is congruent to the table's base modulo 32, so a
// table chunk's sixteen bits lie in one word.
struct Bitmap {
    const u32* w;
    u64 base, bits;
};

AS_HD u32 chunk_bits(const Table& t, const Bitmap& b, u64 chunk) {
    const u64 c = t.base + chunk * kChunk;
    if (c < b.base || c - b.base >= b.bits) return 0u;
    const u64 rel = c - b.base;
    return (b.w[rel >> 5] >> (rel & 31)) & 0xFFFFu;
}

// As drive, passing each counter's bit of `b` along with its entry.
template <class Step>
AS_HD bool drive_bits(const Table& t, const Bitmap& b, u64 c, Step&& step) {
    if (c < t.base) return false;
    const u64 k = c - t.base;
    if (k >= t.len) return false;
    u64 chunk = k / kChunk;
    u32 first = static_cast<u32>(k % kChunk);
    Quad cur[4];
    load_chunk(t, chunk, cur);
    u32 bcur = chunk_bits(t, b, chunk);
    bool live = true;
    for (;;) {
        Quad nxt[4];
        load_chunk(t, chunk + 1, nxt);
        const u32 bnxt = chunk_bits(t, b, chunk + 1);
        const u64 k0 = chunk * kChunk;
        bool ran_out = false;
#define AS_GB_STEP(j)                                                  \
    if (live && (j) >= first) {                                        \
        if (k0 + (j) < t.len) {                                        \
            live = step(cur[(j) / 4].v[(j) % 4], (bcur >> (j)) & 1u);  \
        } else {                                                       \
            live = false;                                              \
            ran_out = true;                                            \
        }                                                              \
    }
        AS_GB_STEP(0) AS_GB_STEP(1) AS_GB_STEP(2) AS_GB_STEP(3)
        AS_GB_STEP(4) AS_GB_STEP(5) AS_GB_STEP(6) AS_GB_STEP(7)
        AS_GB_STEP(8) AS_GB_STEP(9) AS_GB_STEP(10) AS_GB_STEP(11)
        AS_GB_STEP(12) AS_GB_STEP(13) AS_GB_STEP(14) AS_GB_STEP(15)
#undef AS_GB_STEP
        if (!live) return !ran_out;
        ++chunk;
        first = 0;
        if (chunk * kChunk >= t.len) return false;
#pragma unroll
        for (int j = 0; j < 4; ++j) cur[j] = nxt[j];
        bcur = bnxt;
    }
}

AS_HD u64 lower_bound_u32(const u32* a, u64 n, u32 v) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

AS_HD u64 lower_bound_u64(const u64* a, u64 lo, u64 hi, u64 v) {
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// ---- 1. decode every counter ----

struct FillArgs {
    u32* e;
    u64 base, len;
    u64 seed;
    double log1m_p;
    u32 take, amb_cap;
    u32* n_amb;
    u64* amb;  // counters whose skip needs the C library
};

AS_HD void table_fill(const FillArgs& a, u64 k) {
    const double r = rng_uniform(a.seed, kSetupStream, a.base + k);
    u32 g;
    if (!skip_fast(skip_uniform(r), a.log1m_p, g)) {
        g = 0;
        const u32 s = atomic_add_u32(a.n_amb, 1u);
        if (s < a.amb_cap) a.amb[s] = a.base + k;
    }
    a.e[k] = encode_entry(r, g, a.take);
}

// ---- 2. chains ----

struct ChainPass {
    PassDesc P;
    u64 r0;        // first anchor
    u32 n_chains;  // anchors r0 + j*W, j < n_chains; region end r0 + n_chains*W
    u32 chain0;    // global index of chain 0
};

enum : u32 { kChainOk = 0, kChainTable = 1, kChainCapA = 2, kChainRegion = 3, kChainLast = 4, kChainCapB = 5 };

struct ChainArgs {
    Table T;
    ChainPass pass[3];
    u32* abits[3];  // per pass: the chains' recorded row starts, from pass[q].r0
    u64 W;
    u32 capA;
    u32* A;       // [chain][capA]: row starts in [a_j, a_{j+1}) as offsets from a_j
    u32* nA;
    Walker* state;  // chain_a's final state (first draw at or past a_{j+1})
    u32* status_a;
    u32* status_b;
    u32 capB;
    u32* B;         // [chain][capB]: the first row starts chain_b passed before coupling, offsets from a_{j+1}
    Walker* over;   // where a chain with more than capB of them stood at its (capB+1)-th
    u32* nB;        // all row starts chain_b passed before coupling
    u64* couple_x;  // the coupling row start
    u32* couple_k;  // the chain (index within the pass) that recorded it
    u32* couple_r;  // its rank in that chain's list
};

AS_HD int chain_pass_of(const ChainArgs& a, u64 g) {
    return g < a.pass[1].chain0 ? 0 : (g < a.pass[2].chain0 ? 1 : 2);
}

AS_HD void chain_a(const ChainArgs& a, u64 g) {
    const ChainPass& cp = a.pass[chain_pass_of(a, g)];
    const u64 j = g - cp.chain0;
    const u64 a0 = cp.r0 + j * a.W, a1 = a0 + a.W;
    u32* A = a.A + g * a.capA;
    u32 n = 0, st = kChainOk;
    A[n++] = 0;
    Walker w{a0, 0, 0, 0, 0};
    // the chain owns the bitmap words of [a0, a1) (W is a multiple of 32)
    u32* bm = a.abits[chain_pass_of(a, g)];
    bm[(a0 - cp.r0) >> 5] |= 1u << ((a0 - cp.r0) & 31);
    const bool stopped = drive(a.T, w.c, [&](u32 e) {
        if (w.c >= a1) return false;
        if ((walker_step(w, e, cp.P, kNone) & kEvRowEnd) && w.c < a1) {
            if (n == a.capA) {
                st = kChainCapA;
                return false;
            }
            A[n++] = static_cast<u32>(w.c - a0);
            const u64 rel = w.c - cp.r0;
            bm[rel >> 5] |= 1u << (rel & 31);
        }
        return true;
    });
    if (!stopped) st = kChainTable;
    a.nA[g] = n;
    a.state[g] = w;
    a.status_a[g] = st;
}

AS_HD void chain_b(const ChainArgs& a, u64 g) {
    const ChainPass& cp = a.pass[chain_pass_of(a, g)];
    const u64 j = g - cp.chain0;
    a.nB[g] = 0;
    if (j + 1 >= cp.n_chains) { a.status_b[g] = kChainLast; return; }
    if (a.status_a[g] != kChainOk) { a.status_b[g] = a.status_a[g]; return; }
    const u64 rend = cp.r0 + u64(cp.n_chains) * a.W;
    Walker w = a.state[g];
    u32 n = 0, st = kChainTable;
    const u64 a1 = cp.r0 + (j + 1) * a.W;
    u32* B = a.B + g * a.capB;
    const Bitmap bm{a.abits[chain_pass_of(a, g)], cp.r0, u64(cp.n_chains) * a.W};
    drive_bits(a.T, bm, w.c, [&](u32 e, u32 bit) {
        const bool rs = w.pos == 0 && w.phase == 0;
        const bool done = rs && bit;
        const bool region = rs && !bit && w.c >= rend;
        st = done ? kChainOk : (region ? kChainRegion : st);
        const bool go = !(done || region);
        if (rs && go) {
            if (n < a.capB) B[n] = static_cast<u32>(w.c - a1);
            if (n == a.capB) a.over[g] = w;
            ++n;
        }
        Walker nw = w;
        walker_step(nw, e, cp.P, kNone);
        commit(w, nw, go);
        return go;
    });
    if (st == kChainOk) {
        // the chain that recorded this row start, and its rank there
        const u64 k = (w.c - cp.r0) / a.W;
        const u64 gk = cp.chain0 + k;
        const u32* Ak = a.A + gk * a.capA;
        const u64 r = lower_bound_u32(Ak, a.nA[gk], static_cast<u32>(w.c - (cp.r0 + k * a.W)));
        a.couple_x[g] = w.c;
        a.couple_k[g] = static_cast<u32>(k);
        a.couple_r[g] = static_cast<u32>(r);
    }
    a.nB[g] = st == kChainOk ? n : 0;
    a.status_b[g] = st;
}

// One stretch of the stitched path: chain k's recorded starts from rank r on,
// then (when nb > 0 or more is wanted) the nb starts its chain_b walk passed.
struct PathEntry {
    u32 g;      // global chain index
    u32 rank;   // first A entry taken
    u32 na;     // A entries taken
    u32 nb;     // B starts taken
    u64 out;    // position in the list
};

struct PathArgs {
    Table T;
    ChainArgs C;
    const PathEntry* ent;
    u64* L;
    u32* fail;
};

AS_HD void path_write(const PathArgs& a, u64 idx) {
    const PathEntry pe = a.ent[idx];
    const ChainPass& cp = a.C.pass[chain_pass_of(a.C, pe.g)];
    const u64 j = pe.g - cp.chain0;
    const u64 a0 = cp.r0 + j * a.C.W;
    const u32* A = a.C.A + u64(pe.g) * a.C.capA;
    const u32* B = a.C.B + u64(pe.g) * a.C.capB;
    u64 o = pe.out;
    for (u32 i = 0; i < pe.na; ++i) a.L[o++] = a0 + A[pe.rank + i];
    const u32 stored = pe.nb < a.C.capB ? pe.nb : a.C.capB;
    for (u32 i = 0; i < stored; ++i) a.L[o++] = a0 + a.C.W + B[i];
    if (pe.nb <= a.C.capB) return;
    // the rest again from where the chain stood at its (capB+1)-th start
    Walker w = a.C.over[pe.g];
    u32 n = stored;
    const bool stopped = drive(a.T, w.c, [&](u32 e) {
        if (w.pos == 0 && w.phase == 0) {
            a.L[o++] = w.c;
            if (++n == pe.nb) return false;
        }
        walker_step(w, e, cp.P, kNone);
        return true;
    });
    if (!stopped) atomic_or_u32(a.fail, 1u);
}

// ---- 3. autapse repair ----

struct VerifyArgs {
    Table T;
    PassDesc P;
    const u64* L;
    u64 m;  // rows [m, nL - 1) are checked (rows before m were walked exactly)
    int K;
    u32 cap;
    u32* n_rec;
    u64* rec_i;
    int* rec_off;
    u32* fail;
};

AS_HD void verify_row(const VerifyArgs& a, u64 idx) {
    const u64 i = a.m + idx;
    const i64 target = static_cast<i64>(a.P.self_base + i);
    Walker w{a.L[i], 0, 0, 0, 0};
    const bool stopped = drive(a.T, w.c, [&](u32 e) {
        const int ev = walker_step(w, e, a.P, kNone);
        if (ev & kEvEdge) {
            const i64 off = static_cast<i64>(w.last) - target;
            if (off >= -a.K && off <= a.K) {
                const u32 s = atomic_add_u32(a.n_rec, 1u);
                if (s < a.cap) {
                    a.rec_i[s] = i;
                    a.rec_off[s] = static_cast<int>(off);
                } else {
                    atomic_or_u32(a.fail, 2u);
                }
            }
        }
        return (ev & kEvRowEnd) == 0;
    });
    if (!stopped) {
        atomic_or_u32(a.fail, 1u);
        return;
    }
    if (w.c != a.L[i + 1]) atomic_or_u32(a.fail, 4u);
}

struct RepairArgs {
    Table T;
    PassDesc P;
    const u64* L;
    u64 nL;
    Bitmap lbits;  // the list's row starts
    const u64* rec_i;
    const int* rec_off;
    u32 cap_rows;
    u64* out_j;    // list index where the exact walk rejoins (kNoSelf: none; kPassEnd)
    u32* out_m;    // rows walked up to there
    u64* rows;     // [record][cap_rows]: the row starts it passed (and, at the pass end, the end)
    // copy mode: for record sel[idx], its starts go to F[pos[idx] ...]
    const u32* sel;
    const u64* pos;
    u64* F;
    u32* fail;
};

constexpr u64 kPassEnd = ~u64(0) - 1;

// Exact walk from list row i whose true index is i + off, row after row,
// until a row start is again in the list, or through the pass's last row
// (out_j = kPassEnd; the pass's end is then stored as the last start).
AS_HD void repair_row(const RepairArgs& a, u64 rec) {
    const u64 i = a.rec_i[rec];
    const i64 t = static_cast<i64>(i) + a.rec_off[rec];
    a.out_j[rec] = kNoSelf;
    if (t < 0) return;
    u64* out = a.rows + rec * a.cap_rows;
    u32 rows = 0, res = 0;  // res: 1 rejoined, 2 pass end, 3 too long
    u32 self = a.P.self_base + static_cast<u32>(t);
    Walker w{a.L[i], 0, 0, 0, 0};
    const bool stopped = drive_bits(a.T, a.lbits, w.c, [&](u32 e, u32 bit) {
        // at a row start after the first row: the pass's end, a start of the
        // list, or (room permitting: out[rows - 1] < cap_rows) one to keep
        const bool rs = w.pos == 0 && w.phase == 0 && rows > 0;
        const bool end = rs && static_cast<u64>(t) + rows >= a.P.n_rows;
        const bool join = rs && !end && bit;
        const bool full = rs && !join && rows >= a.cap_rows;
        res = full ? 3u : (end ? 2u : (join ? 1u : res));
        if (rs && !join && !full) out[rows - 1] = w.c;
        const bool go = !(end || join || full);
        self = rs ? a.P.self_base + static_cast<u32>(t) + rows : self;
        Walker nw = w;
        const int ev = walker_step(nw, e, a.P, self);
        rows += (go && (ev & kEvRowEnd)) ? 1u : 0u;
        commit(w, nw, go);
        return go;
    });
    if (!stopped) {
        atomic_or_u32(a.fail, 8u);
        return;
    }
    if (res == 2) {
        a.out_j[rec] = kPassEnd;
        a.out_m[rec] = rows;
    } else if (res == 1) {
        const u64 j = lower_bound_u64(a.L, i + 1, a.nL, w.c);
        if (j < a.nL && a.L[j] == w.c) {
            a.out_j[rec] = j;
            a.out_m[rec] = rows;
        }
    }
}

// The starts a chosen repair passed: rows - 1 of them when it rejoined the
// list, `rows` (the last is the pass's end) when it reached the pass's end.
AS_HD void repair_copy(const RepairArgs& a, u64 idx) {
    const u32 rec = a.sel[idx];
    const u32 n = a.out_j[rec] == kPassEnd ? a.out_m[rec] : a.out_m[rec] - 1;
    const u64* src = a.rows + u64(rec) * a.cap_rows;
    for (u32 k = 0; k < n; ++k) a.F[a.pos[idx] + k] = src[k];
}

// Bits of the list entries [lo, hi) that lie in the bitmap's range.
struct MarkArgs {
    const u64* L;
    u64 lo;
    u32* w;
    u64 base, bits;
};
AS_HD void mark_list(const MarkArgs& a, u64 idx) {
    const u64 x = a.L[a.lo + idx];
    if (x < a.base || x - a.base >= a.bits) return;
    atomic_or_u32(a.w + ((x - a.base) >> 5), 1u << ((x - a.base) & 31));
}


// F[o + k] = L[s + k] for the splice table's copy segments.
struct SpliceSeg {
    u64 dst, src, len;
};
struct CopyArgs {
    const SpliceSeg* seg;
    u32 n_seg;
    const u64* L;
    u64* F;
};
// One lane per destination element would need a search; segments are long,
// so one lane copies a fixed chunk of one segment.
constexpr u64 kCopyChunk = 256;
AS_HD void splice_copy(const CopyArgs& a, u64 idx) {
    // idx enumerates (segment, chunk) pairs through a flat chunk count per segment
    u64 s = 0, acc = 0;
    for (; s < a.n_seg; ++s) {
        const u64 nch = (a.seg[s].len + kCopyChunk - 1) / kCopyChunk;
        if (idx < acc + nch) break;
        acc += nch;
    }
    if (s == a.n_seg) return;
    const SpliceSeg sg = a.seg[s];
    const u64 k0 = (idx - acc) * kCopyChunk;
    const u64 k1 = k0 + kCopyChunk < sg.len ? k0 + kCopyChunk : sg.len;
    for (u64 k = k0; k < k1; ++k) a.F[sg.dst + k] = a.L[sg.src + k];
}

// ---- 4. exact walks with true indices ----

struct EmitArgs {
    Table T;
    PassDesc P;
    const u64* F;     // n_rows + 1 starts
    u32* n_edge;      // per row (count mode)
    u32* n_tf;
    const u64* e_off; // per row, exclusive (emit mode)
    const u64* t_off;
    u64 e_base, t_base;
    u32* src;
    u32* dst;
    u32* tf_pre;
    u32* tf_post;
    u8* tf_slot;
    u32* fail;
};

template <bool kEmit>
AS_HD void exact_row(const EmitArgs& a, u64 i) {
    const u32 self = a.P.self_base == kNone ? kNone : a.P.self_base + static_cast<u32>(i);
    Walker w{a.F[i], 0, 0, 0, 0};
    u64 ne = 0, nt = 0;
    const bool stopped = drive(a.T, w.c, [&](u32 e) {
        const int ev = walker_step(w, e, a.P, self);
        if (ev & kEvEdge) {
            if (kEmit) {
                const u64 o = a.e_base + a.e_off[i] + ne;
                a.src[o] = static_cast<u32>(i);
                a.dst[o] = a.P.post_offset + w.last;
            }
            ++ne;
        }
        if (ev & kEvTf) {
            if (kEmit) {
                const u64 o = a.t_base + a.t_off[i] + nt;
                a.tf_pre[o] = static_cast<u32>(i);
                a.tf_post[o] = a.P.post_offset + w.last;
                a.tf_slot[o] = static_cast<u8>(e >> kSlotShift);
            }
            ++nt;
        }
        return (ev & kEvRowEnd) == 0;
    });
    if (!stopped) {
        atomic_or_u32(a.fail, 16u);
        return;
    }
    if (!kEmit) {
        if (w.c != a.F[i + 1]) atomic_or_u32(a.fail, 32u);
        a.n_edge[i] = static_cast<u32>(ne);
        if (a.n_tf != nullptr) a.n_tf[i] = static_cast<u32>(nt);
    }
}

// ---- third-factor astrocytes and the out-degree cap ----

struct AstroArgs {
    const u32* tf_post;
    const u8* tf_slot;
    const u32* pool;  // [neuron][take]
    u32 take;
    u32* astro;
    u32* count;  // per astrocyte
    u32* fail;
};

AS_HD void tf_astro(const AstroArgs& a, u64 k) {
    const u32 s = a.pool[u64(a.tf_post[k]) * a.take + a.tf_slot[k]];
    a.astro[k] = s;
    // the cap (src/astrosim_fused.cu:707) binds only past this count; then
    // the host builder applies it in generation order
    if (atomic_add_u32(a.count + s, 1u) + 1u > ScalingConfig::kMaxAstroOutDegree) atomic_or_u32(a.fail, 64u);
}

// Row starts from targets: count[key]++ (then an exclusive scan).
struct HistArgs {
    const u32* key;
    u32* count;
};
AS_HD void histogram(const HistArgs& a, u64 k) { atomic_add_u32(a.count + a.key[k], 1u); }

}  // namespace gb
}  // namespace astrosim
