// as_batch.hip -- the blocked, event-driven method over a batch of networks.
//
// Every synaptic delay is the same d >= 1 steps (1 ms at dt = 0.1 ms), so
// within an aligned block of L steps (L divides d, L <= 32) no cell reads
// anything written in the same block: all inputs of block b come from block
// b - d/L. One INTEGRATE launch therefore advances every cell of every
// member network through a block (or a chunk of it) with its state in
// registers.
//
// Inputs are delivered event-driven instead of by probing every in-edge:
//   * each cell records the steps of its block at which it emitted (a spike,
//     or a nonzero SIC factor) as a bit word, the emitted values in a ring
//     indexed by absolute step, and appends itself to its block's emitter
//     list when the word is nonzero;
//   * a SCATTER launch walks only the emitters' out-edges (a transposed copy
//     of the in-CSR built on the device) and appends an entry (in-CSR slot,
//     source, and for a source that emitted once its step and value) to the
//     target's fixed-capacity bucket for that connection type;
//   * INTEGRATE sorts each bucket by in-CSR slot and sums it per step in that
//     order, which is the reference's CSR-order sum with its exact-zero terms
//     removed (src/astrosim_fused.cu:384-389, 450-463, 482-490), so FP64
//     results are bit-identical to the plain method and to the host oracle.
//     A bucket that overflowed falls back to walking the whole in-row. The
//     per-step sums, the Poisson drive and the noise draws are staged in LDS
//     by a prologue, so the RK4 step loop carries only the cell's state.
//
// Members occupy contiguous ranges of the arena's cell and edge spaces with
// rebased (global) ids. Workgroups are tiles of one member's astrocytes or
// neurons; every random stream uses the member's seed, local cell index and
// local step, so a member's results do not depend on its batch mates.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <limits>
#include <map>
#include <vector>

#include "as_batch.h"
#include "as_model.h"
#include "as_model32.h"

namespace astrosim {
namespace {

constexpr unsigned kBucketCap = 8;
// The chunk length the integrate kernel is also compiled for: the whole
// 1 ms block at dt = 0.1 ms.
constexpr int kDefaultChunk = 10;
constexpr int kCountersPerMember = 8;
enum { T_EXC = 0, T_INH = 1, T_SIC = 2, T_E2A = 3 };

struct Member {
    std::uint64_t seed;
    index_t nb, nn, ne;  // neuron base (global), neurons, excitatory neurons
    index_t ab, na;      // astrocyte base (global), astrocytes
    std::int64_t step0, step_end;  // global steps [step0, step_end) are simulated
    std::uint32_t slot;            // member slot (spike counters)
    std::uint32_t pad;
};

// A workgroup's cells: local ids [first, end) of one member's astrocytes
// (kind 0), excitatory neurons (kind 1) or inhibitory neurons (kind 2). One
// population per tile keeps the population's constants wave-uniform.
struct Tile {
    std::uint32_t member, first, kind, end;
};

// An active member's tiles, written from table position tile0 on: its
// astrocytes, then its excitatory, then its inhibitory neurons, each cut
// into workgroup-sized tiles.
struct TileSpan {
    std::uint32_t member, na, ne, nn, tile0, pad;
};

// Everything that does not depend on the state precision.
struct BatchDev {
    const uint2* in_rs[4];  // (begin, end) per target: neurons (exc, inh, sic), astrocytes (e2a)
    const index_t* in_src[4];
    const uint2* out_rs[4];  // (begin, end) per source: neurons (exc, inh, e2a), astrocytes (sic)
    const uint2* out_ent[4];  // (global target, global in-CSR slot)
    // H bucket sets per type: set b % H holds the inputs block b reads
    unsigned* bcnt[4];  // [H][targets]
    uint4* bent[4];     // [H][targets][kBucketCap] bucket entries (see entry layout below)
    std::uint8_t* ever;
    std::int32_t* refr;
    std::uint32_t* spk_n;  // spikes per neuron since admission
    std::uint32_t* spk_w;  // [H][ncap]
    std::uint32_t* sic_w;  // [H][acap]
    unsigned long long* counters;  // [members][kCountersPerMember]
    unsigned* overflow;
    const Member* members;
    const Tile* tiles;
    std::int64_t* clock;  // block index for graph replays
    std::uint64_t ncap, acap;
    std::uint64_t sstride;  // row stride of the SIC value ring: acap + ncap (FP32c), acap (FP64)
};

// Row stride of the SIC value ring for state type R (FP64 reads acap, the
// field its kernels always held).
template <class R>
__device__ __forceinline__ std::uint64_t sic_stride(const BatchDev& D) {
    return sizeof(R) == 4 ? D.sstride : D.acap;
}

// State arrays in the precision R (double: FP64; float: FP32c, where
// Ca/IP3/h carry compensation terms and tau_IP3 holds 1/tau_IP3).
template <class R>
struct BlkState {
    R *Ca, *IP3, *h;
    R *Ca_lo, *IP3_lo, *h_lo;
    R *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3;
    R *V, *w, *g_ex, *dg_ex, *g_in, *dg_in, *I_sic;
    R *V_reset, *b;
    R *stp_x, *stp_u, *stp_t;
    R* spk_v;  // [Rv][ncap] x*u at the spike step
    R* sic_v;  // [Rv][acap + ncap] SIC factor where nonzero; columns acap + neuron: see neuron_cell
};

struct BlkLaunch {
    std::int64_t t0, bstart;
    int n;             // steps in this launch
    int d, L, H, Rv;   // delay, block length, history blocks, value-ring depth
    std::uint32_t wslot, rslot;  // history slots of block b and block b - d/L
    std::uint32_t bw;            // the bucket set of block b + d/L (b's emissions go there)
    int first, last;   // the launch starts / completes its block
    int chunk;         // most steps one integrate launch advances (LDS staging)
    int sub;           // graph replays: first step of the launch within its block
    int lane_words;    // LDS words (of the state type) of each lane's staging column
};

// Device memory (the global address space). The arrays' pointers reach the
// kernels inside structures, where the compiler cannot tell device memory
// from LDS: accesses through them would be flat instructions, which take a
// 64-bit address per access and share the scalar loads' wait counter.
#define AS_GLOBAL __attribute__((address_space(1)))

// Element at a 32-bit byte offset from a wave-uniform base pointer. One
// offset register per cell then addresses every state array of its type
// (scalar base + vector offset), instead of a 64-bit address per array.
template <class T>
__device__ __forceinline__ AS_GLOBAL T& at(T* base, std::uint32_t byte_off) {
    return *reinterpret_cast<AS_GLOBAL T*>((AS_GLOBAL char*)(base) + byte_off);
}
template <class T>
__device__ __forceinline__ const AS_GLOBAL T& at(const T* base, std::uint32_t byte_off) {
    return *reinterpret_cast<const AS_GLOBAL T*>((const AS_GLOBAL char*)(base) + byte_off);
}
// HIP's vector types are classes: their copies from device memory go
// through these builtin vectors.
using u32x2 = unsigned __attribute__((ext_vector_type(2)));
using u32x4 = unsigned __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 ld(const AS_GLOBAL uint2* p) {
    const u32x2 v = *reinterpret_cast<const AS_GLOBAL u32x2*>(p);
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 ld(const AS_GLOBAL uint4* p) {
    const u32x4 v = *reinterpret_cast<const AS_GLOBAL u32x4*>(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st(AS_GLOBAL uint4* p, const uint4& e) {
    u32x4 v;
    v.x = e.x;
    v.y = e.y;
    v.z = e.z;
    v.w = e.w;
    *reinterpret_cast<AS_GLOBAL u32x4*>(p) = v;
}
// Relaxed device-scope atomic add (atomicAdd's) on device memory.
template <class T>
__device__ __forceinline__ T atomic_add(AS_GLOBAL T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// x, hidden from the optimiser: values derived from it are computed where they
// are used, instead of being hoisted and held in scalar registers across the
// step loop (where the integrate kernel's scalar registers run short).
__device__ __forceinline__ std::uint32_t opaque(std::uint32_t x) {
    asm volatile("" : "+s"(x));
    return x;
}
// The same pointer, typed as device memory.
template <class T>
__device__ __forceinline__ AS_GLOBAL T* gp(T* p) {
    return (AS_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ const AS_GLOBAL T* gp(const T* p) {
    return (const AS_GLOBAL T*)p;
}

// A value every lane of the wave holds, moved to scalar registers. Loads
// through pointers the kernel may also write (the graph clock, the tile and
// member tables) are vector loads, so their results, and everything computed
// from them (steps, stream bases, 64-bit divisions), would otherwise occupy
// vector registers and vector-instruction issue.
__device__ __forceinline__ std::uint32_t uniform(std::uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ std::uint64_t uniform(std::uint64_t x) {
    const std::uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<std::uint32_t>(x));
    const std::uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<std::uint32_t>(x >> 32));
    return (std::uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ std::int64_t uniform(std::int64_t x) {
    return static_cast<std::int64_t>(uniform(static_cast<std::uint64_t>(x)));
}

// Graph replays read the block index from device memory instead of the
// launch arguments: the launch is a whole block.
__device__ __forceinline__ void resolve_launch(BlkLaunch& B, const std::int64_t* clock) {
    if (!clock) return;
    const std::int64_t b = uniform(*gp(clock));
    B.bstart = b * B.L;
    B.t0 = B.bstart + B.sub;
    B.n = B.L - B.sub < B.chunk ? B.L - B.sub : B.chunk;
    B.first = B.sub == 0;
    B.last = B.sub + B.n == B.L;
    B.wslot = static_cast<std::uint32_t>(b % B.H);
    B.rslot = static_cast<std::uint32_t>(((b - B.d / B.L) % B.H + B.H) % B.H);
    B.bw = static_cast<std::uint32_t>((b + B.d / B.L) % B.H);
}

__device__ __forceinline__ unsigned lane_id() { return threadIdx.x & 63u; }


// Row of step t in a ring of depth Rv (t may be negative: the value row of
// t - d before step d is never dereferenced, but stays in range).
__device__ __forceinline__ std::uint32_t ring_row(std::int64_t t, int Rv) {
    const std::int64_t r = t % Rv;
    return static_cast<std::uint32_t>(r < 0 ? r + Rv : r);
}

__device__ __forceinline__ std::uint32_t ring_next(std::uint32_t r, int Rv) {
    return r + 1u == std::uint32_t(Rv) ? 0u : r + 1u;
}

// Bucket set s of connection type T (targets: neurons; astrocytes for T_E2A).
__device__ __forceinline__ AS_GLOBAL unsigned* bucket_counts(const BatchDev& D, int T, std::uint32_t s) {
    return gp(D.bcnt[T]) + std::size_t(s) * (T == T_E2A ? D.acap : D.ncap);
}
__device__ __forceinline__ AS_GLOBAL uint4* bucket_entries(const BatchDev& D, int T, std::uint32_t s) {
    return gp(D.bent[T]) + std::size_t(s) * (T == T_E2A ? D.acap : D.ncap) * kBucketCap;
}

// ---------------------------------------------------------------------------
// Delayed inputs of one target and connection type over a block.
//
// A target's bucket holds one entry per source that emitted in the block d
// steps earlier, in arrival order:
//   x  the edge's in-CSR slot;
//   y  the source's global id (bits 0-25), and, when the source emitted at a
//      single step k of its block (a neuron almost always does), k in bits
//      26-30 and kAttached;
//   zw with kAttached, the emitted value's bits (FP32: z only).
// The reference sums a target's in-row in CSR order (src/astrosim_fused.cu:
// 384-389, 450-463, 482-490); summing the bucket in ascending slot order
// reproduces that sum bit for bit, since the omitted terms are exact zeros.
// Each bucket is sorted once per block. Neuron-sourced inputs keep up to NR
// attached (step, value) pairs in registers, so their step loop touches no
// memory; SIC inputs (astrocytes emit at every step while active) keep
// (source, word) pairs and load the values of step k+1 during step k. A
// target with more entries, or an entry without an attached value (rare),
// takes the in-memory path: entries sorted in place and walked each step,
// or the whole in-row when the bucket overflowed.
// ---------------------------------------------------------------------------

constexpr std::uint32_t kSrcMask = (1u << 26) - 1u;
constexpr std::uint32_t kAttached = 1u << 31;
constexpr int kStepShift = 26;
// Empty network slots sort last: no in-CSR slot reaches 2^32 - 1.
constexpr std::uint32_t kNoSlot = 0xFFFFFFFFu;

// Compare-exchange on bucket entries, ascending by slot.
__device__ __forceinline__ void cas(uint4& a, uint4& b) {
    const bool swap = b.x < a.x;
    const uint4 lo = swap ? b : a, hi = swap ? a : b;
    a = lo;
    b = hi;
}

// Sorting networks (Batcher) for 1, 2 and 4 entries.
template <int N>
__device__ __forceinline__ void sort_entries(uint4 (&e)[N]) {
    static_assert(N == 1 || N == 2 || N == 4, "network size");
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        cas(e[0], e[1]);
    } else {
        cas(e[0], e[1]); cas(e[2], e[3]);
        cas(e[0], e[2]); cas(e[1], e[3]);
        cas(e[1], e[2]);
    }
}

// Loads the first N entries of a bucket (empty slots: kNoSlot) and sorts them.
template <int N>
__device__ __forceinline__ void load_sorted(uint4 (&e)[N], const AS_GLOBAL uint4* ent, unsigned cnt) {
#pragma unroll
    for (int q = 0; q < N; ++q) e[q] = unsigned(q) < cnt ? ld(ent + q) : make_uint4(kNoSlot, 0u, 0u, 0u);
    sort_entries<N>(e);
}

// Sorts a bucket of at most kBucketCap entries in place, by selection in
// memory: more than four entries are rare, and a register sorting network for
// eight would hold 32 registers at the kernel's register peak (fewer resident
// waves). In-CSR slots are distinct, so the order is the same either way.
__device__ __forceinline__ void sort_in_place(AS_GLOBAL uint4* ent, unsigned cnt) {
    for (unsigned r = 0; r + 1u < cnt; ++r) {
        unsigned m = r;
        std::uint32_t km = ld(ent + r).x;
        for (unsigned q = r + 1u; q < cnt; ++q) {
            const std::uint32_t k = ld(ent + q).x;
            if (k < km) {
                km = k;
                m = q;
            }
        }
        if (m != r) {
            const uint4 a = ld(ent + r), b = ld(ent + m);
            st(ent + r, b);
            st(ent + m, a);
        }
    }
}

template <class R>
__device__ __forceinline__ R attached_value(const uint4& e) {
    if constexpr (sizeof(R) == 8) {
        return __hiloint2double(static_cast<int>(e.w), static_cast<int>(e.z));
    } else {
        return __uint_as_float(e.z);
    }
}

__device__ __forceinline__ unsigned attached_step(const uint4& e) { return (e.y >> kStepShift) & 31u; }

// Sorts a bucket of up to kBucketCap entries in place (nothing to do for
// fewer than two).
__device__ __forceinline__ void sort_bucket(AS_GLOBAL uint4* ent, unsigned cnt) {
    if (cnt < 2u || cnt > kBucketCap) return;
    if (cnt <= 4u) {
        uint4 e[4];
        load_sorted<4>(e, ent, cnt);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (unsigned(q) < cnt) st(ent + q, e[q]);
    } else {
        sort_in_place(ent, cnt);
    }
}

// Bits [c0, c1) of a block word.
__device__ __forceinline__ std::uint32_t step_mask(unsigned c0, unsigned c1) {
    const std::uint32_t hi = c1 >= 32u ? 0xFFFFFFFFu : ((1u << c1) - 1u);
    return hi & ~((1u << c0) - 1u);
}

// Each lane stages its inputs in its own contiguous run of LDS words (an odd
// number of 32-bit words apart, so a wave's accesses to one slot fall in
// distinct banks). With the launch's chunk length KC known at compile time
// (0: taken from the launch), loops over a launch's steps unroll and every
// slot is an immediate offset from the lane's base.
template <int KC>
__device__ __forceinline__ unsigned chunk_of(const BlkLaunch& B) {
    return KC > 0 ? unsigned(KC) : static_cast<unsigned>(B.chunk);
}

// f(i) for i in [0, n), n <= the chunk length (uniform).
template <int KC, class F>
__device__ __forceinline__ void for_steps(unsigned n, F f) {
    if constexpr (KC > 0) {
#pragma unroll
        for (unsigned i = 0; i < unsigned(KC); ++i) {
            if (i >= n) break;
            f(i);
        }
    } else {
        for (unsigned i = 0; i < n; ++i) f(i);
    }
}

// A lane's staging slots in LDS. Each wave owns lane_words x 64 words, slot
// by slot: slot i of lane l is word i * 64 + l of its wave's region, so a
// wave's accesses to one slot are 64 consecutive words (no bank conflicts at
// any slot count) and every slot is an immediate offset from the lane's base.
template <class R>
struct Col {
    R* p;  // slot 0 of the lane
    __device__ __forceinline__ R& operator[](unsigned i) const { return p[i * 64u]; }
    __device__ __forceinline__ Col operator+(unsigned i) const { return Col{p + i * 64u}; }
};

// Input sums of one target and connection type for block steps [c0, c1),
// into the lane's LDS slots col[k - c0], in ascending in-CSR slot order: the
// bucket is sorted; an attached entry adds its value at its step, an
// unattached one the source's ring values at its emitted steps, and an
// overflowed bucket walks the whole in-row. ring(k): the source
// population's value-ring row of block step k.
template <int KC, class R, class RingF, class RowF>
__device__ void gather_chunk(Col<R> col, unsigned c0, unsigned c1, const AS_GLOBAL uint4* ent, unsigned cnt,
                             const AS_GLOBAL std::uint32_t* words, RingF ring, RowF row,
                             const AS_GLOBAL index_t* src) {
    for_steps<KC>(c1 - c0, [&](unsigned i) { col[i] = R(0); });
    if (cnt == 0u) return;
    const std::uint32_t mask = step_mask(c0, c1);
    if (cnt <= kBucketCap) {
        for (unsigned r = 0; r < cnt; ++r) {
            const uint4 e = ld(ent + r);
            if (e.y & kAttached) {
                const unsigned k = attached_step(e);
                if ((mask >> k) & 1u) col[k - c0] += attached_value<R>(e);
            } else {
                const index_t q = e.y & kSrcMask;
                for (std::uint32_t w = words[q] & mask; w != 0u; w &= w - 1u) {
                    const unsigned k = __ffs(static_cast<int>(w)) - 1;
                    col[k - c0] += ring(k)[q];
                }
            }
        }
    } else {
        const uint2 rw = row();
        for (index_t e = rw.x; e < rw.y; ++e) {
            const index_t q = src[e];
            for (std::uint32_t w = words[q] & mask; w != 0u; w &= w - 1u) {
                const unsigned k = __ffs(static_cast<int>(w)) - 1;
                col[k - c0] += ring(k)[q];
            }
        }
    }
}

// The in-memory path of the SIC stream: the sum of block step k over a
// sorted bucket of at most kBucketCap entries, or the whole in-row.
template <class R, class RowF>
__device__ R memory_sum(const AS_GLOBAL uint4* ent, unsigned cnt, unsigned k, const AS_GLOBAL std::uint32_t* words,
                        const AS_GLOBAL R* vals, RowF row, const AS_GLOBAL index_t* src) {
    R sum = R(0);
    if (cnt <= kBucketCap) {
        for (unsigned r = 0; r < cnt; ++r) {
            const index_t q = ld(ent + r).y & kSrcMask;
            if ((words[q] >> k) & 1u) sum += vals[q];
        }
    } else {
        const uint2 rw = row();
        for (index_t e = rw.x; e < rw.y; ++e) {
            const index_t q = src[e];
            if ((words[q] >> k) & 1u) sum += vals[q];
        }
    }
    return sum;
}

// SIC inputs: (source, word) pairs in slot order, values prefetched a step ahead.
template <class R, int NR>
struct InputStream {
    std::uint32_t off[NR];   // source id * sizeof(R): the source's value column
    std::uint32_t word[NR];  // the source's emission word (bit k: emitted at block step k)
    R v[NR];                 // prefetched values of one step (zero where not emitted)
    unsigned cnt;            // bucket entries; more than NR: the in-memory path
};

template <class R, int NR>
__device__ __forceinline__ void input_init(InputStream<R, NR>& s, AS_GLOBAL uint4* ent, unsigned cnt,
                                           const AS_GLOBAL std::uint32_t* words) {
    s.cnt = cnt;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        s.off[r] = 0u;
        s.word[r] = 0u;
        s.v[r] = R(0);
    }
    if (cnt == 0u) return;
    if (cnt <= unsigned(NR)) {
        uint4 e[NR];
        load_sorted<NR>(e, ent, cnt);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (e[r].x != kNoSlot) {
                const std::uint32_t q = e[r].y & kSrcMask;
                s.off[r] = q * std::uint32_t(sizeof(R));
                s.word[r] = words[q];
            }
        }
    } else if (cnt <= kBucketCap) {
        sort_in_place(ent, cnt);
    }
}

// Loads the values of block step k from value-ring row `vals`.
template <class R, int NR>
__device__ __forceinline__ void input_prefetch(InputStream<R, NR>& s, unsigned k, const AS_GLOBAL R* vals) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        R x = R(0);
        if ((s.word[r] >> k) & 1u) x = *reinterpret_cast<const AS_GLOBAL R*>((const AS_GLOBAL char*)vals + s.off[r]);
        s.v[r] = x;
    }
}

// The input sum of block step k: the prefetched values in slot order, or the
// in-memory path.
template <class R, int NR, class EntF, class RowF>
__device__ __forceinline__ R input_sum(const InputStream<R, NR>& s, unsigned k, EntF ent,
                                       const AS_GLOBAL std::uint32_t* words, const AS_GLOBAL R* vals, RowF row,
                                       const AS_GLOBAL index_t* src) {
    R sum = R(0);
#pragma unroll
    for (int r = 0; r < NR; ++r) sum += s.v[r];
    if (s.cnt > unsigned(NR)) sum = memory_sum<R>(ent(), s.cnt, k, words, vals, row, src);
    return sum;
}

// ---------------------------------------------------------------------------
// SCATTER: at the end of its block, each wave appends the entries of its
// emitters to their targets' buckets of block b + d/L (set B.bw), inside the
// integrate launch that completes the block: no cell reads that set before
// block b + d/L, and a separate launch per block would wait for the whole
// integrate launch and then for its own tail.
// ---------------------------------------------------------------------------
constexpr unsigned kScatterUnroll = 4;

// `word`: the lane's emission word of the block (0: no emission; a neuron's
// spike steps, an astrocyte's SIC-active steps). The wave's emitters are taken
// one at a time; the wave's lanes walk the emitter's out-rows.
template <class R>
__device__ void scatter_emitters(const BatchDev& D, const R* __restrict__ spk_v, const BlkLaunch& B,
                                 bool is_neuron, std::uint32_t src, std::uint32_t word) {
    std::uint64_t em = __ballot(word != 0u);
    if (em == 0ull) return;
    // A neuron that spiked at one step k of its block hands (k, value) to
    // its targets with the entry.
    std::uint32_t tag = src, vlo = 0u, vhi = 0u;
    if (is_neuron && __popc(word) == 1) {
        const unsigned k = __ffs(static_cast<int>(word)) - 1;
        const R v = gp(spk_v)[std::size_t(ring_row(B.bstart + k, B.Rv)) * D.ncap + src];
        tag |= kAttached | (k << kStepShift);
        if constexpr (sizeof(R) == 8) {
            vlo = static_cast<std::uint32_t>(__double2loint(v));
            vhi = static_cast<std::uint32_t>(__double2hiint(v));
        } else {
            vlo = __float_as_uint(v);
        }
    }
    // A neuron walks its exc and n2a out-edges and the inhibitory rows whose
    // stored LOCAL id equals its own local id (src/astrosim_fused.cu:459-462);
    // an astrocyte walks its a2n out-edges. The emitter's out-rows are walked
    // as one sequence of edge slots, each lane appending kScatterUnroll slots
    // per iteration: their loads, then their returning atomics, then their
    // stores. An append is a chain of dependent memory operations; this keeps
    // several chains in flight per lane and walks all rows in one round
    // instead of one round per row. Bucket order does not matter: buckets are
    // sorted by in-CSR slot when they are read.
    const int T0 = is_neuron ? T_EXC : T_SIC;
    const AS_GLOBAL uint2* ent0 = gp(D.out_ent[T0]);
    const AS_GLOBAL uint2* ent1 = gp(D.out_ent[T_INH]);
    const AS_GLOBAL uint2* ent2 = gp(D.out_ent[T_E2A]);
    AS_GLOBAL unsigned* cnt0 = bucket_counts(D, T0, B.bw);
    AS_GLOBAL uint4* bkt0 = bucket_entries(D, T0, B.bw);
    AS_GLOBAL unsigned* cnt1 = bucket_counts(D, T_INH, B.bw);
    AS_GLOBAL uint4* bkt1 = bucket_entries(D, T_INH, B.bw);
    AS_GLOBAL unsigned* cnt2 = bucket_counts(D, T_E2A, B.bw);
    AS_GLOBAL uint4* bkt2 = bucket_entries(D, T_E2A, B.bw);
    while (em != 0ull) {
        const unsigned l = static_cast<unsigned>(__builtin_ctzll(em));
        em &= em - 1ull;
        const std::uint32_t es = __builtin_amdgcn_readlane(src, l);
        const std::uint32_t et = __builtin_amdgcn_readlane(tag, l);
        const std::uint32_t elo = __builtin_amdgcn_readlane(vlo, l);
        const std::uint32_t ehi = __builtin_amdgcn_readlane(vhi, l);
        const uint2 r0 = ld(gp(D.out_rs[T0]) + es);
        uint2 r1 = make_uint2(0u, 0u), r2 = make_uint2(0u, 0u);
        if (is_neuron) {
            r1 = ld(gp(D.out_rs[T_INH]) + es);
            r2 = ld(gp(D.out_rs[T_E2A]) + es);
        }
        const unsigned n0 = r0.y - r0.x, n01 = n0 + (r1.y - r1.x), n = n01 + (r2.y - r2.x);
        for (unsigned s0 = lane_id(); s0 < n; s0 += 64u * kScatterUnroll) {
            uint2 te[kScatterUnroll];  // (target, in-CSR slot)
            unsigned slot[kScatterUnroll];
            int ty[kScatterUnroll];
#pragma unroll
            for (unsigned u = 0; u < kScatterUnroll; ++u) {
                const unsigned sl = s0 + 64u * u;
                if (sl >= n) continue;
                ty[u] = sl < n0 ? 0 : (sl < n01 ? 1 : 2);
                const std::size_t e = ty[u] == 0 ? r0.x + sl : (ty[u] == 1 ? r1.x + (sl - n0) : r2.x + (sl - n01));
                const AS_GLOBAL uint2* ent = ty[u] == 0 ? ent0 : (ty[u] == 1 ? ent1 : ent2);
                te[u] = ld(ent + e);
            }
#pragma unroll
            for (unsigned u = 0; u < kScatterUnroll; ++u) {
                if (s0 + 64u * u >= n) continue;
                AS_GLOBAL unsigned* cnt = ty[u] == 0 ? cnt0 : (ty[u] == 1 ? cnt1 : cnt2);
                slot[u] = atomic_add(cnt + te[u].x, 1u);
            }
#pragma unroll
            for (unsigned u = 0; u < kScatterUnroll; ++u) {
                if (s0 + 64u * u >= n) continue;
                AS_GLOBAL uint4* bkt = ty[u] == 0 ? bkt0 : (ty[u] == 1 ? bkt1 : bkt2);
                if (slot[u] < kBucketCap)
                    st(bkt + std::size_t(te[u].x) * kBucketCap + slot[u], make_uint4(te[u].y, et, elo, ehi));
                else if (slot[u] == kBucketCap)
                    atomic_add(gp(D.overflow), 1u);
            }
        }
    }
}

__global__ void tick_kernel(std::int64_t* clock) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *clock += 1;
}

// ---------------------------------------------------------------------------
// INTEGRATE: advance every cell over steps [t0, t0 + n) of one block.
// ---------------------------------------------------------------------------

// A uniform value copied into vector registers. The FP64 neuron step loop
// uses more uniform constants than the scalar registers a wave can address,
// and the compiler then parks some in lanes of vector registers, paying a
// v_readlane at every use on every step. Values copied here are read from
// vector registers instead (the kernel has vector registers to spare). The
// asm is pure, so the copy is made once, before the loop.
__device__ __forceinline__ double in_vgpr(double x) {
    double y;
    asm("v_mov_b64 %0, %1" : "=v"(y) : "s"(x));
    return y;
}

__device__ __forceinline__ FixedDivisor in_vgpr(const FixedDivisor& d) {
    return FixedDivisor{in_vgpr(d.b), in_vgpr(d.yd), in_vgpr(d.yld), in_vgpr(d.ny)};
}

// Both cell functions advance one cell over global steps [lo, hi), at most
// B.chunk of them, in two phases. The prologue computes everything that does
// not depend on the cell's own state into the lane's LDS slots col[i]: the
// delayed input sums (in slot order), the Poisson drive folded
// in with the reference's operation order, and the noise draws. The step
// loop then reads those and runs the arithmetic that does; keeping the
// prologue's random-number and gather code out of that loop keeps its
// constants and temporaries out of the loop's registers.

// One astrocyte (global id ga, local id a) (src/astrosim_fused.cu:374-427).
// Returns its emission word when the launch completes its block (bit k: SIC
// factor > 0 at step k), else 0. col: [0, C) the IP3 input of each step,
// [C, 2C) its noise term.
template <int KC, class R>
__device__ std::uint32_t astro_cell(const BatchDev& D, const BlkState<R>& S, const RunParams& P, const Params32& Q,
                           const Member& M, const BlkLaunch& B, index_t ga, index_t a, std::int64_t lo,
                           std::int64_t hi, Col<R> col) {
    const std::uint32_t oa = ga * std::uint32_t(sizeof(R)), oa4 = ga * 4u;
    const unsigned C = chunk_of<KC>(B);
    const unsigned n = static_cast<unsigned>(hi - lo);
    // state first: the loads' latency passes during the prologue
    R ca = at(S.Ca, oa), ip3 = at(S.IP3, oa), hh = at(S.h, oa);
    R ca_lo = R(0), ip3_lo = R(0), hh_lo = R(0);
    if constexpr (sizeof(R) == 4) {
        ca_lo = at(S.Ca_lo, oa);
        ip3_lo = at(S.IP3_lo, oa);
        hh_lo = at(S.h_lo, oa);
    }
    const R Ca_tot = at(S.Ca_tot, oa), IP3_0 = at(S.IP3_0, oa), tau = at(S.tau_IP3, oa), dIP3 = at(S.delta_IP3, oa);
    std::uint32_t wout = B.first ? 0u : at(D.sic_w + std::size_t(B.wslot) * D.acap, oa4);
    {
        const AS_GLOBAL std::uint32_t* nwords = gp(D.spk_w) + std::size_t(B.rslot) * D.ncap;
        AS_GLOBAL uint4* ent = bucket_entries(D, T_E2A, B.wslot) + std::size_t(ga) * kBucketCap;
        const unsigned cnt = at(bucket_counts(D, T_E2A, B.wslot), oa4);
        sort_bucket(ent, cnt);
        auto nring = [&](unsigned k) { return gp(S.spk_v) + std::size_t(ring_row(B.bstart + k - B.d, B.Rv)) * D.ncap; };
        gather_chunk<KC, R>(col, static_cast<unsigned>(lo - B.bstart), static_cast<unsigned>(hi - B.bstart), ent,
                            cnt, nwords, nring, [&] { return ld(gp(D.in_rs[T_E2A]) + ga); }, gp(D.in_src[T_E2A]));
        // random streams tl*kStepStride + a of the member's step tl = t - step0
        // (src/astrosim_fused.cu:393-405): their bases advance by a constant
        // per step; FP32 state draws the noise in FP32 from the same uniforms
        const std::uint64_t stream0 = std::uint64_t(lo - M.step0) * kStepStride + a;
        std::uint64_t zp = rng_base(M.seed ^ kKeyAstroPoisson, stream0);
        std::uint64_t zn = rng_base(M.seed ^ kKeyAstroNoise, stream0);
        for (unsigned j = 0; j < n; ++j) {
            int ev = 0;
            if (P.poiss_astro) ev = rng_poisson_z(zp, P.lam_astro, P.lim_astro);
            R noise = R(0);
            if (P.noise_astro > 0.0) {
                if constexpr (sizeof(R) == 8) noise = P.noise_astro * rng_normal_z(zn);
                else noise = static_cast<float>(P.noise_astro) * normal32(zn);
            }
            zp += kStepStride * kRngStreamMul;
            zn += kStepStride * kRngStreamMul;
            R& in = col[j];
            if constexpr (sizeof(R) == 8) {
                double ip3_input = P.w_n2a * in;
                if (ev > 0) ip3_input += P.w_poiss_astro * static_cast<double>(ev);
                in = ip3_input;
            } else {
                float ip3_input = Q.w_n2a * in;
                if (ev > 0) ip3_input += Q.w_poiss_astro * static_cast<float>(ev);
                in = ip3_input;
            }
            col[C + j] = noise;
        }
    }
    std::uint32_t pr = ring_row(lo, B.Rv);
    FixedDivisor tau_div{};  // FP64: the cell's tau_IP3 divisor, made once per launch
    if constexpr (sizeof(R) == 8) tau_div = make_fixed_divisor(tau);
    bool ever = false;
    for (std::int64_t t = lo; t < hi; ++t) {
        const unsigned k = static_cast<unsigned>(t - B.bstart), j = static_cast<unsigned>(t - lo);
        const R ip3_input = col[j], noise = col[C + j];
        R factor;
        if constexpr (sizeof(R) == 8) {
            astro_advance(P.ac, Ca_tot, IP3_0, tau_div, dIP3, ip3_input, noise, P.h_step, P.h_sixth, P.substeps, ca,
                          ip3, hh);
            factor = sic_factor(P.ac, ca);
        } else {
            astro_advance32(Q.ac, Q, Ca_tot, IP3_0, tau, dIP3, ip3_input, noise, P.substeps, ca, ca_lo, ip3, ip3_lo,
                            hh, hh_lo);
            factor = sic_factor32(Q.ac, ca, ca_lo);
        }
        if (factor > R(0)) {
            ever = true;
            wout |= 1u << k;
            at(S.sic_v + std::size_t(pr) * sic_stride<R>(D), oa) = factor;
        }
        pr = ring_next(pr, B.Rv);
    }
    at(S.Ca, oa) = ca;
    at(S.IP3, oa) = ip3;
    at(S.h, oa) = hh;
    if constexpr (sizeof(R) == 4) {
        at(S.Ca_lo, oa) = ca_lo;
        at(S.IP3_lo, oa) = ip3_lo;
        at(S.h_lo, oa) = hh_lo;
    }
    if (ever) at(D.ever, ga) = 1;
    at(D.sic_w + std::size_t(B.wslot) * D.acap, oa4) = wout;
    if (B.last) at(bucket_counts(D, T_E2A, opaque(B.wslot)), oa4) = 0;
    return B.last ? wout : 0u;
}

// One neuron (global id gj, local id j) (src/astrosim_fused.cu:433-587).
// Returns the number of spikes; *emit: its spike word when the launch
// completes its block, else 0. col: [0, C) the excitatory input of each
// step, [C, 2C) the inhibitory input, then the noise values of the launch's
// noise indices. SIC inputs stay in registers (values loaded a step ahead).
template <int KC, class R, bool exc>
__device__ unsigned neuron_cell(const BatchDev& D, const BlkState<R>& S, const RunParams& P, const Params32& Q,
                                const Member& M, const BlkLaunch& B, index_t gj, index_t j, std::int64_t lo,
                                std::int64_t hi, Col<R> col, std::uint32_t* emit) {
    const std::uint32_t on = gj * std::uint32_t(sizeof(R)), o4 = gj * 4u;
    const unsigned C = chunk_of<KC>(B);
    const unsigned noise_base = 2u * C;
    // FP64 with at most two noise indices per launch (steps_per_noise >= C,
    // the usual case) keeps their values in registers and only the inputs in
    // LDS: 20 words per lane instead of 22 at C = 10 let 4 waves per SIMD share
    // a CU instead of 3. FP32c keeps them in LDS: its LDS already allows the 7
    // waves its registers allow, and two more registers measured slower.
    const bool noise_lds = sizeof(R) == 4 || (C - 1u) / unsigned(P.steps_per_noise) != 0u;
    R noise_first = R(0), noise_next = R(0);
    const AS_GLOBAL std::uint32_t* awords = gp(D.sic_w) + std::size_t(B.rslot) * D.acap;
    auto es = [&] { return bucket_entries(D, T_SIC, opaque(B.wslot)) + std::size_t(gj) * kBucketCap; };
    const double nstd = exc ? P.noise_exc : P.noise_inh;
    const std::int64_t tl0 = lo - M.step0;  // the member's own step at lo
    const std::uint64_t nidx0 = static_cast<std::uint64_t>(tl0 / P.steps_per_noise);
    // The cell's state is loaded first: the loads' latency then passes during
    // the prologue (nothing the prologue writes aliases the state).
    R V = at(S.V, on), w = at(S.w, on), g_ex = at(S.g_ex, on), dg_ex = at(S.dg_ex, on), g_in = at(S.g_in, on),
      dg_in = at(S.dg_in, on);
    int refr = at(D.refr, o4);
    R I_sic = at(S.I_sic, on);
    const R V_reset = at(S.V_reset, on), bb = at(S.b, on);
    std::uint32_t wout = B.first ? 0u : at(D.spk_w + std::size_t(B.wslot) * D.ncap, o4);
    const unsigned cs = at(bucket_counts(D, T_SIC, B.wslot), o4);
    {
        const AS_GLOBAL std::uint32_t* nwords = gp(D.spk_w) + std::size_t(B.rslot) * D.ncap;
        AS_GLOBAL uint4* ee = bucket_entries(D, T_EXC, B.wslot) + std::size_t(gj) * kBucketCap;
        AS_GLOBAL uint4* ei = bucket_entries(D, T_INH, B.wslot) + std::size_t(gj) * kBucketCap;
        const unsigned ce = at(bucket_counts(D, T_EXC, B.wslot), o4), ci = at(bucket_counts(D, T_INH, B.wslot), o4);
        sort_bucket(ee, ce);
        sort_bucket(ei, ci);
        auto nring = [&](unsigned k) { return gp(S.spk_v) + std::size_t(ring_row(B.bstart + k - B.d, B.Rv)) * D.ncap; };
        const unsigned k0 = static_cast<unsigned>(lo - B.bstart), k1 = static_cast<unsigned>(hi - B.bstart);
        // The Poisson drive of each step (stream tl*kStepStride + j) is folded
        // into its excitatory input as w_e*sum, then + w_poiss*events for
        // steps with events (src/astrosim_fused.cu:472-479).
        const unsigned n = static_cast<unsigned>(hi - lo);
        const bool poiss = exc ? P.poiss_exc : P.poiss_inh;
        const double lam = exc ? P.lam_exc : P.lam_inh, lim = exc ? P.lim_exc : P.lim_inh;
        const std::uint64_t zstep = kStepStride * kRngStreamMul;  // stream base advance per step
        const std::uint64_t zp0 = rng_base(M.seed ^ kKeyNeuronPoisson, std::uint64_t(tl0) * kStepStride + j);
        // The event counts of every step with events, handed to record(i, ev).
        // Knuth's method runs in two passes. First every step's first uniform
        // u0 = m0 * 2^-53: most steps stop there with no event; the m0 of the
        // others is kept in LDS slots of the step (stash/unstash). Then one
        // loop continues those steps' products one draw per iteration, so a
        // wave runs the maximum over its lanes of the launch's draws rather
        // than the sum over steps of each step's maximum. Same draws, same
        // products.
        auto poisson = [&](auto stash, auto unstash, auto record) {
            if (poiss && lam > 30.0) {  // the normal approximation: one draw per step
                std::uint64_t z = zp0;
                for (unsigned i = 0; i < n; ++i, z += zstep) {
                    const int ev = rng_poisson_z(z, lam, lim);
                    if (ev > 0) record(i, ev);
                }
                return;
            }
            if (!poiss) return;
            const std::uint64_t lim53 = static_cast<std::uint64_t>(lim * 9007199254740992.0);
            std::uint32_t pending = 0u;
            std::uint64_t z = zp0;
            for_steps<KC>(n, [&](unsigned i) {
                const std::uint64_t m0 = rng_mix(z) >> 11;
                if (m0 > lim53) {
                    pending |= 1u << i;
                    stash(i, m0);
                }
                z += zstep;
            });
            unsigned cur = 32u;  // the step being continued; 32: none
            double product = 0.0;
            int count = 0;
            while (pending != 0u || cur < 32u) {
                if (cur == 32u) {
                    cur = static_cast<unsigned>(__builtin_ctz(pending));
                    pending &= pending - 1u;
                    product = unstash(cur) * (1.0 / 9007199254740992.0);
                    count = 0;
                    z = zp0 + std::uint64_t(cur) * zstep;  // counter 0 of step cur
                }
                ++count;
                z += kRngCounterMul;
                product *= rng_to_unit(rng_mix(z));
                if (!(product > lim)) {
                    record(cur, count);
                    cur = 32u;
                }
            }
        };
        if constexpr (sizeof(R) == 8) {
            // FP64: excitatory sums first, the m0 of pending steps in their
            // (not yet gathered) inhibitory slots, events added at once
            gather_chunk<KC, R>(col, k0, k1, ee, ce, nwords, nring, [&] { return ld(gp(D.in_rs[T_EXC]) + gj); },
                                gp(D.in_src[T_EXC]));
            for_steps<KC>(n, [&](unsigned i) { col[i] = P.w_e * col[i]; });
            poisson([&](unsigned i, std::uint64_t m0) { col[C + i] = static_cast<double>(m0); },  // exact: < 2^53
                    [&](unsigned i) { return col[C + i]; },
                    [&](unsigned i, int ev) {
                        col[i] = col[i] + (exc ? P.w_poiss_exc : P.w_poiss_inh) * static_cast<double>(ev);
                    });
        } else {
            // FP32c: the Poisson passes first, while both input regions are
            // free: a pending step's m0 in its excitatory (low 32 bits) and
            // inhibitory (high bits) slots, then its event count in its
            // inhibitory slot (bit i of evmask: step i has events) until the
            // excitatory sums are gathered. This keeps a lane's LDS to two
            // words per step, which lets more waves share a CU.
            std::uint32_t evmask = 0u;
            poisson(
                [&](unsigned i, std::uint64_t m0) {
                    col[i] = __uint_as_float(static_cast<std::uint32_t>(m0));
                    col[C + i] = __uint_as_float(static_cast<std::uint32_t>(m0 >> 32));
                },
                [&](unsigned i) {
                    return static_cast<double>(std::uint64_t(__float_as_uint(col[i])) |
                                               (std::uint64_t(__float_as_uint(col[C + i])) << 32));
                },
                [&](unsigned i, int ev) {
                    col[C + i] = __int_as_float(ev);  // the step's slots are free again
                    evmask |= 1u << i;
                });
            gather_chunk<KC, R>(col, k0, k1, ee, ce, nwords, nring, [&] { return ld(gp(D.in_rs[T_EXC]) + gj); },
                                gp(D.in_src[T_EXC]));
            for_steps<KC>(n, [&](unsigned i) {
                float ein = Q.w_e_abs * col[i];
                if ((evmask >> i) & 1u)
                    ein = ein + (exc ? Q.w_poiss_exc : Q.w_poiss_inh) * static_cast<float>(__float_as_int(col[C + i]));
                col[i] = ein;
            });
        }
        // inhibitory source ids are local indices used as neuron ids
        // (src/astrosim_fused.cu:459-462); the scatter has already resolved them
        gather_chunk<KC, R>(col + C, k0, k1, ei, ci, nwords, nring, [&] { return ld(gp(D.in_rs[T_INH]) + gj); },
                            gp(D.in_src[T_INH]));
        for_steps<KC>(n, [&](unsigned i) {
            R& iin = col[C + i];
            if constexpr (sizeof(R) == 8) iin = (-P.w_i) * iin;
            else iin = Q.w_i_abs * iin;
        });
        // the noise values of the launch's noise indices tl / steps_per_noise
        // (stream idx*kNoiseStride + j, src/astrosim_fused.cu:496-500); FP32
        // state draws them in FP32 from the same uniforms
        if (nstd > 0.0) {
            const std::uint64_t nlast = static_cast<std::uint64_t>((tl0 + (hi - lo) - 1) / P.steps_per_noise);
            std::uint64_t zn = rng_base(M.seed ^ kKeyNeuronNoise, nidx0 * kNoiseStride + std::uint64_t(j));
            for (std::uint64_t idx = nidx0; idx <= nlast; ++idx, zn += kNoiseStride * kRngStreamMul) {
                R v;
                if constexpr (sizeof(R) == 8) v = nstd * rng_normal_z(zn);
                else v = static_cast<float>(nstd) * normal32(zn);
                const unsigned sl = static_cast<unsigned>(idx - nidx0);
                if (noise_lds) col[noise_base + sl] = v;
                else if (sl == 0u) noise_first = v;
                else noise_next = v;
            }
        }
    }
    // SIC inputs from (source, word) pairs, values loaded a step ahead
    InputStream<R, 1> is;
    std::uint32_t vr = ring_row(lo - B.d, B.Rv);
    input_init(is, es(), cs, awords);
    input_prefetch(is, static_cast<unsigned>(lo - B.bstart), gp(S.sic_v) + std::size_t(vr) * sic_stride<R>(D));
    // FP32c with KC compiled in: a neuron with several SIC sources in the
    // block (several active astrocytes in reach) has its per-step SIC sums
    // computed once per launch: its sources in slot order (the sorted bucket,
    // or the in-row when it overflowed), each source's values of all the
    // launch's steps loaded together and added at its emitted steps, so every
    // step's sum has the in-memory path's terms in the same order. The sums go
    // to the neuron's own column (acap + neuron) of the SIC value ring's rows
    // of those steps, and the neuron's register stream then reads that column
    // as if it were a single source: the step loop loads one value per step,
    // a step ahead, instead of walking the sources in device memory at every
    // step (which held waves for hundreds of microseconds once many
    // astrocytes were active). FP64 has no registers to spare for it.
    if constexpr (KC > 0 && sizeof(R) == 4) {
        if (cs > 1u) {
            const unsigned k0 = static_cast<unsigned>(lo - B.bstart), n = static_cast<unsigned>(hi - lo);
            const std::uint32_t mask = step_mask(k0, k0 + n);
            R acc[KC];
#pragma unroll
            for (unsigned i = 0; i < unsigned(KC); ++i) acc[i] = R(0);
            const bool in_bucket = cs <= kBucketCap;
            index_t e0 = 0u, e1 = cs;
            if (!in_bucket) {
                const uint2 rw = ld(gp(D.in_rs[T_SIC]) + gj);
                e0 = rw.x;
                e1 = rw.y;
            }
            const AS_GLOBAL index_t* isrc = gp(D.in_src[T_SIC]);
            const AS_GLOBAL uint4* ent = es();
            for (index_t e = e0; e < e1; ++e) {
                const index_t q = in_bucket ? (ld(ent + e).y & kSrcMask) : isrc[e];
                const std::uint32_t w = (awords[q] & mask) >> k0;
                if (w == 0u) continue;
                R v[KC];
#pragma unroll
                for (unsigned i = 0; i < unsigned(KC); ++i) {
                    v[i] = R(0);
                    if ((w >> i) & 1u)
                        v[i] = gp(S.sic_v)[std::size_t(ring_row(B.bstart + k0 + i - B.d, B.Rv)) * sic_stride<R>(D) + q];
                }
#pragma unroll
                for (unsigned i = 0; i < unsigned(KC); ++i)
                    if ((w >> i) & 1u) acc[i] += v[i];
            }
            const std::size_t col_n = D.acap + gj;
#pragma unroll
            for (unsigned i = 0; i < unsigned(KC); ++i)
                if (i < n) gp(S.sic_v)[std::size_t(ring_row(B.bstart + k0 + i - B.d, B.Rv)) * sic_stride<R>(D) + col_n] = acc[i];
            is.cnt = 1u;
            is.off[0] = static_cast<std::uint32_t>(col_n * sizeof(R));
            is.word[0] = mask;
            input_prefetch(is, k0, gp(S.sic_v) + std::size_t(vr) * sic_stride<R>(D));
        }
    }
    std::uint32_t pr = ring_row(lo, B.Rv);
    R noise_val = R(0);
    unsigned nspikes = 0;
    // The SIC-update phase (an update at tl >= d_a2n with (tl - d_a2n) %
    // sic_interval == 0) and the noise index with its remainder, advanced
    // per step.
    int sic_wait;
    {
        const std::int64_t u = tl0 - P.d_a2n;
        sic_wait = u < 0 ? static_cast<int>(-u) : static_cast<int>((P.sic_interval - u % P.sic_interval) % P.sic_interval);
    }
    unsigned nslot = 0;
    int nrem = static_cast<int>(tl0 % P.steps_per_noise);
    // the population's constants; FP64 holds its three divisors in vector registers
    NeuronConst pc = exc ? P.ne : P.ni;
    if constexpr (sizeof(R) == 8) {
        pc.Delta_T_div = in_vgpr(pc.Delta_T_div);
        pc.C_m_div = in_vgpr(pc.C_m_div);
        pc.tau_w_div = in_vgpr(pc.tau_w_div);
    }
    // FP32: the spike current's two FP64 constants in vector registers too
    NeuronConst32 pc32 = exc ? Q.ne : Q.ni;
    if constexpr (sizeof(R) == 4) {
        pc32.y_scale_d = in_vgpr(pc32.y_scale_d);
        pc32.y_shift_d = in_vgpr(pc32.y_shift_d);
    }
    for (std::int64_t t = lo; t < hi; ++t) {
        const unsigned k = static_cast<unsigned>(t - B.bstart), i = static_cast<unsigned>(t - lo);
        const std::int64_t tl = t - M.step0;  // the member's own step
        const R exc_input = col[i];
        const R inh_input = col[C + i];
        const bool sic_step = sic_wait == 0;
        sic_wait = sic_step ? P.sic_interval - 1 : sic_wait - 1;
        if (sic_step) {
            const R sic_sum = input_sum(is, k, es, awords, gp(S.sic_v) + std::size_t(vr) * sic_stride<R>(D),
                                        [&] { return ld(gp(D.in_rs[T_SIC]) + gj); }, gp(D.in_src[T_SIC]));
            if constexpr (sizeof(R) == 8) I_sic = P.w_a2n * sic_sum;
            else I_sic = Q.w_a2n * sic_sum;
        }
        {
            const std::uint32_t vn = ring_next(vr, B.Rv);
            if (t + 1 < hi) input_prefetch(is, k + 1u, gp(S.sic_v) + std::size_t(vn) * sic_stride<R>(D));
            vr = vn;
        }
        if (nstd > 0.0 && (t == lo || nrem == 0))  // the noise stream changes every steps_per_noise steps
            noise_val = noise_lds ? col[noise_base + nslot] : (nslot == 0u ? noise_first : noise_next);
        bool spiked;
        if constexpr (sizeof(R) == 8) {
            const NeuronConst& p = pc;
            double I_ext = I_sic;
            if (nstd > 0.0) I_ext += noise_val;
            NeuronRegs r{V, w, g_ex, dg_ex + exc_input * p.psc_init_ex, g_in, dg_in + inh_input * p.psc_init_in, refr};
            spiked = neuron_substeps(p, r, I_ext, V_reset, bb, P.h_step, P.h_sixth, P.substeps,
                                     exc ? P.ex_decay_e : P.ex_decay_i, exc ? P.in_decay_e : P.in_decay_i);
            V = r.V; w = r.w; g_ex = r.g_ex; dg_ex = r.dg_ex; g_in = r.g_in; dg_in = r.dg_in; refr = r.refractory;
        } else {
            const NeuronConst32& p = pc32;
            const float I_ext = I_sic + noise_val;
            NeuronRegs32 r{V, w, g_ex, dg_ex + exc_input * p.psc_init_ex, g_in, dg_in + inh_input * p.psc_init_in, refr};
            spiked = neuron_substeps32(p, Q, r, I_ext, V_reset, bb, P.substeps, exc ? Q.ex_decay_e : Q.ex_decay_i,
                                       exc ? Q.in_decay_e : Q.in_decay_i);
            V = r.V; w = r.w; g_ex = r.g_ex; dg_ex = r.dg_ex; g_in = r.g_in; dg_in = r.dg_in; refr = r.refractory;
        }
        if (spiked) {
            R out = R(1);
            if (P.stp_enabled) {
                R x = at(S.stp_x, on), u = at(S.stp_u, on), tlast = at(S.stp_t, on);
                if constexpr (sizeof(R) == 8) {
                    out = stp_on_spike(x, u, tlast, static_cast<double>(tl) * P.dt, P.stp_U, P.stp_tau_rec,
                                       P.stp_tau_fac);
                } else {
                    const float t_now = static_cast<float>(static_cast<double>(tl) * P.dt);
                    const float elapsed = t_now - tlast;
                    const float x_decay = __expf(-elapsed / static_cast<float>(P.stp_tau_rec));
                    const float u_decay =
                        P.stp_tau_fac < 1e-10 ? 0.0f : __expf(-elapsed / static_cast<float>(P.stp_tau_fac));
                    const float U = static_cast<float>(P.stp_U);
                    x = 1.0f + (x - x * u - 1.0f) * x_decay;
                    u = U + u * (1.0f - U) * u_decay;
                    tlast = t_now;
                    out = x * u;
                }
                at(S.stp_x, on) = x;
                at(S.stp_u, on) = u;
                at(S.stp_t, on) = tlast;
            }
            at(S.spk_v + std::size_t(pr) * D.ncap, on) = out;
            wout |= 1u << k;
            ++nspikes;
        }
        pr = ring_next(pr, B.Rv);
        if (++nrem == P.steps_per_noise) {
            nrem = 0;
            ++nslot;
        }
    }
    at(S.V, on) = V;
    at(S.w, on) = w;
    at(S.g_ex, on) = g_ex;
    at(S.dg_ex, on) = dg_ex;
    at(S.g_in, on) = g_in;
    at(S.dg_in, on) = dg_in;
    at(D.refr, o4) = refr;
    if (nspikes != 0u) at(D.spk_n, o4) += nspikes;
    at(S.I_sic, on) = I_sic;
    at(D.spk_w + std::size_t(B.wslot) * D.ncap, o4) = wout;
    if (B.last) {
        const std::uint32_t ws = opaque(B.wslot);
        at(bucket_counts(D, T_EXC, ws), o4) = 0;
        at(bucket_counts(D, T_INH, ws), o4) = 0;
        at(bucket_counts(D, T_SIC, ws), o4) = 0;
    }
    *emit = B.last ? wout : 0u;
    return nspikes;
}

// Everything an integrate launch reads besides the block descriptor, kept in
// device memory: uniform loads go through the scalar cache on demand instead
// of occupying (and spilling) scalar registers as kernel arguments.
template <class R>
struct IntegrateCtx {
    BatchDev D;
    BlkState<R> S;
    RunParams P;
    Params32 Q;
};

// Minimum waves per SIMD the integrate kernel's register budget targets, per
// state precision. FP32c: the 7 its LDS staging allows at the default 10-step
// chunk (22 words per lane), reached without spills (68 VGPRs). FP64: 3; it
// uses 128 VGPRs anyway, which with its 20 LDS words per lane gives 4 waves,
// and a bound of 4 measured 0.6% slower (a different register assignment).
#ifndef AS_INTEGRATE_WAVES_F64
#define AS_INTEGRATE_WAVES_F64 3
#endif
#ifndef AS_INTEGRATE_WAVES_F32
#define AS_INTEGRATE_WAVES_F32 7
#endif
template <class R>
constexpr int kIntegrateWaves = sizeof(R) == 8 ? AS_INTEGRATE_WAVES_F64 : AS_INTEGRATE_WAVES_F32;

// KC: the launch's chunk length when it is compiled in (the default block of
// 10 steps), 0 otherwise.
template <int MAXT, class R, int KC>
__global__ __launch_bounds__(MAXT, kIntegrateWaves<R>) void integrate_kernel(const IntegrateCtx<R>* __restrict__ ctx, BlkLaunch B) {
    // each lane's staging slots: B.lane_words slots of R in its wave's region
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const Col<R> col{reinterpret_cast<R*>(lds_raw) + (threadIdx.x >> 6) * (unsigned(B.lane_words) * 64u) +
                     (threadIdx.x & 63u)};
    const BatchDev& D = ctx->D;
    const BlkState<R>& S = ctx->S;
    const RunParams& P = ctx->P;
    const Params32& Q = ctx->Q;
    resolve_launch(B, D.clock);
    // the tile and its member are the same for the whole workgroup
    Tile tile = D.tiles[blockIdx.x];
    tile = Tile{uniform(tile.member), uniform(tile.first), uniform(tile.kind), uniform(tile.end)};
    Member M = D.members[tile.member];
    M.seed = uniform(M.seed);
    M.nb = uniform(M.nb);
    M.nn = uniform(M.nn);
    M.ne = uniform(M.ne);
    M.ab = uniform(M.ab);
    M.na = uniform(M.na);
    M.step0 = uniform(M.step0);
    M.step_end = uniform(M.step_end);
    M.slot = uniform(M.slot);
    // the part of this launch's steps inside the member's run
    const std::int64_t lo = B.t0 > M.step0 ? B.t0 : M.step0;
    const std::int64_t end = B.t0 + B.n;
    const std::int64_t hi = end < M.step_end ? end : M.step_end;
    const index_t local = tile.first + threadIdx.x;
    if (tile.kind == 0) {
        const bool run = local < tile.end && lo < hi;
        const std::uint32_t emit = run ? astro_cell<KC, R>(D, S, P, Q, M, B, M.ab + local, local, lo, hi, col) : 0u;
        if (B.last) scatter_emitters<R>(D, S.spk_v, B, false, M.ab + local, emit);
        return;
    }
    const bool run = local < tile.end && lo < hi;
    std::uint32_t emit = 0u;
    unsigned nspikes = 0;
    if (run) {
        // one instantiation per population keeps only its constants live
        if (tile.kind == 1)
            nspikes = neuron_cell<KC, R, true>(D, S, P, Q, M, B, M.nb + local, local, lo, hi, col, &emit);
        else
            nspikes = neuron_cell<KC, R, false>(D, S, P, Q, M, B, M.nb + local, local, lo, hi, col, &emit);
    }
    if (B.last) scatter_emitters<R>(D, S.spk_v, B, true, M.nb + local, emit);
    // diagnostic spike total of the member: one atomic per wave with spikes
    unsigned long long c = nspikes;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if (lane_id() == 0 && c != 0ull)
        atomic_add(gp(D.counters) + std::size_t(M.slot) * kCountersPerMember + (blockIdx.x % kCountersPerMember), c);
}

// ---------------------------------------------------------------------------
// Member admission kernels.
// ---------------------------------------------------------------------------
// in_rs pairs from a raw row_start, shifted to the member's edge range.
__global__ void rows_kernel(const index_t* rs, index_t n, index_t ebase, uint2* out) {
    const index_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = make_uint2(ebase + rs[t], ebase + rs[t + 1]);
}

// One workgroup per active member writes its tiles (the order is that of
// the host loops that formerly built the table).
__global__ void tiles_kernel(const TileSpan* spans, std::uint32_t threads, Tile* tiles) {
    const TileSpan sp = spans[blockIdx.x];
    const std::uint32_t ta = (sp.na + threads - 1) / threads;
    const std::uint32_t te = (sp.ne + threads - 1) / threads;
    const std::uint32_t ti = (sp.nn - sp.ne + threads - 1) / threads;
    for (std::uint32_t t = threadIdx.x; t < ta + te + ti; t += blockDim.x) {
        Tile x;
        x.member = sp.member;
        if (t < ta) {
            x.first = t * threads; x.kind = 0u; x.end = sp.na;
        } else if (t < ta + te) {
            x.first = (t - ta) * threads; x.kind = 1u; x.end = sp.ne;
        } else {
            x.first = sp.ne + (t - ta - te) * threads; x.kind = 2u; x.end = sp.nn;
        }
        tiles[sp.tile0 + t] = x;
    }
}

__global__ void rebase_kernel(index_t* p, std::size_t n, index_t base) {
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] += base;
}

// Counts out-edges per local source (sources outside [sbase, sbase+ns) skip).
__global__ void count_sources_kernel(const index_t* src, std::size_t n_edges, index_t sbase, index_t ns,
                                     index_t* cnt) {
    const std::size_t e = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n_edges) {
        const index_t s = src[e] - sbase;
        if (s < ns) atomicAdd(&cnt[s], 1u);
    }
}

// out_rs of a member's source space [sbase, sbase + n_space) from the
// exclusive scan of the counts (sources >= ns have no out-edges).
__global__ void out_rows_kernel(const index_t* scan, index_t ns, index_t n_space, index_t sbase, index_t ebase,
                                uint2* out_rs) {
    const index_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_space) out_rs[sbase + s] = s < ns ? make_uint2(ebase + scan[s], ebase + scan[s + 1]) : make_uint2(0, 0);
}

__global__ void fill_out_kernel(const uint2* in_rs, const index_t* src, index_t tbase, index_t n_targets,
                                index_t sbase, index_t ns, index_t ebase, index_t* cursor, uint2* out) {
    const index_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_targets) return;
    const uint2 r = in_rs[tbase + t];
    for (index_t e = r.x; e < r.y; ++e) {
        const index_t s = src[e] - sbase;
        if (s < ns) out[ebase + atomicAdd(&cursor[s], 1u)] = make_uint2(tbase + t, e);
    }
}

template <class R>
__global__ void fill_kernel(R* p, R v, std::size_t n) {
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct Geometry {
    int d = 10, L = 10, H = 2, Rv = 20;
};

Geometry geometry(const RunParams& rp) {
    if (rp.d_e != rp.d_i || rp.d_e != rp.d_a2n)
        throw Error(AS_ERR_INVALID, "the blocked method needs equal exc/inh/a2n delays");
    Geometry g;
    g.d = rp.d_e;
    g.L = 1;
    for (int l = std::min(g.d, 32); l >= 1; --l)
        if (g.d % l == 0) {
            g.L = l;
            break;
        }
    g.H = g.d / g.L + 1;
    g.Rv = g.H * g.L;
    return g;
}

// First-fit allocator of [begin, end) ranges in one id space.
class RangeAllocator {
public:
    explicit RangeAllocator(std::uint64_t cap) {
        if (cap > 0) free_.push_back({0, cap});
    }
    bool can(std::uint64_t n) const {
        if (n == 0) return true;
        for (auto& r : free_)
            if (r.second - r.first >= n) return true;
        return false;
    }
    std::uint64_t take(std::uint64_t n) {
        if (n == 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second - it->first >= n) {
                const std::uint64_t b = it->first;
                it->first += n;
                if (it->first == it->second) free_.erase(it);
                return b;
            }
        }
        throw Error(AS_ERR_NOMEM, "batch capacity exhausted");
    }
    void give(std::uint64_t b, std::uint64_t n) {
        if (n == 0) return;
        free_.push_back({b, b + n});
        std::sort(free_.begin(), free_.end());
        std::vector<std::pair<std::uint64_t, std::uint64_t>> m;
        for (auto& r : free_) {
            if (!m.empty() && m.back().second == r.first) m.back().second = r.second;
            else m.push_back(r);
        }
        free_.swap(m);
    }
    std::uint64_t largest() const {
        std::uint64_t best = 0;
        for (auto& r : free_) best = std::max(best, r.second - r.first);
        return best;
    }

private:
    std::vector<std::pair<std::uint64_t, std::uint64_t>> free_;
};

struct BatchLayout {
    std::size_t in_rs[4], in_src[4], out_rs[4], out_ent[4], bcnt[4], bent[4];
    std::size_t astro[10], ever, neuron[12], refr, spk_n;
    std::size_t spk_w, spk_v, sic_w, sic_v, counters, overflow;
    std::size_t members, tiles, clock, ctx, scratch0, scratch1, scan_tmp;
    std::size_t scan_tmp_bytes, tiles_cap, spans;
    std::size_t total;
};

std::size_t scan_temp_bytes(std::uint64_t n) {
    std::size_t bytes = 0;
    index_t* p = nullptr;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, p, p, int(n + 1));
    return std::max<std::size_t>(bytes, 256);
}

BatchLayout batch_layout(const BatchCapacity& c, const RunParams& rp) {
    const Geometry g = geometry(rp);
    const std::size_t rsize = rp.precision == AS_FP64 ? 8 : 4;
    Layout L;
    BatchLayout o{};
    const std::uint64_t nn = std::max<std::uint64_t>(1, c.neurons), na = std::max<std::uint64_t>(1, c.astros);
    for (int i = 0; i < 4; ++i) {
        const std::uint64_t nt = i == T_E2A ? na : nn;   // targets
        const std::uint64_t ns = i == T_SIC ? na : nn;   // sources
        const std::uint64_t ne = std::max<std::uint64_t>(1, c.edges[i]);
        o.in_rs[i] = L.add(nt * sizeof(uint2));
        o.in_src[i] = L.add(ne * sizeof(index_t));
        o.out_rs[i] = L.add(ns * sizeof(uint2));
        o.out_ent[i] = L.add(ne * sizeof(uint2));
        o.bcnt[i] = L.add(std::size_t(g.H) * nt * sizeof(unsigned));
        o.bent[i] = L.add(std::size_t(g.H) * nt * kBucketCap * sizeof(uint4));
    }
    const int n_astro_arrays = rsize == 4 ? 10 : 7;
    for (int i = 0; i < n_astro_arrays; ++i) o.astro[i] = L.add(na * rsize);
    o.ever = L.add(na);
    for (int i = 0; i < 12; ++i) o.neuron[i] = L.add(nn * rsize);
    o.refr = L.add(nn * sizeof(std::int32_t));
    o.spk_n = L.add(nn * sizeof(std::uint32_t));
    o.spk_w = L.add(std::size_t(g.H) * nn * 4);
    o.spk_v = L.add(std::size_t(g.Rv) * nn * rsize);
    o.sic_w = L.add(std::size_t(g.H) * na * 4);
    o.sic_v = L.add(std::size_t(g.Rv) * (rsize == 4 ? na + nn : na) * rsize);
    o.counters = L.add(std::size_t(c.members) * kCountersPerMember * sizeof(unsigned long long));
    o.overflow = L.add(sizeof(unsigned));
    o.members = L.add(std::size_t(c.members) * sizeof(Member));
    o.tiles_cap = (nn + 63) / 64 + (na + 63) / 64 + 3 * std::size_t(c.members);
    o.tiles = L.add(o.tiles_cap * sizeof(Tile));
    o.spans = L.add(std::size_t(c.members) * sizeof(TileSpan));
    o.clock = L.add(sizeof(std::int64_t));
    o.ctx = L.add(2 * sizeof(IntegrateCtx<double>));  // eager and graph-replay contexts
    const std::uint64_t sm = std::max(nn, na) + 1;
    o.scratch0 = L.add(sm * sizeof(index_t));
    o.scratch1 = L.add(sm * sizeof(index_t));
    o.scan_tmp_bytes = scan_temp_bytes(sm);
    o.scan_tmp = L.add(o.scan_tmp_bytes);
    o.total = L.total;
    return o;
}

struct MemberHost {
    bool used = false, active = false;
    bool staging = false;             // uploaded by add_async, not yet active
    std::int64_t pending_steps = 0;   // its steps, counted from its activation
    hipEvent_t up_done = nullptr;     // its upload's completion
    Member m{};
    std::uint64_t ebase[4] = {0, 0, 0, 0}, ecount[4] = {0, 0, 0, 0};
    std::int64_t free_at_block = -1;  // ranges released once the global block reaches this
};

template <class R>
class BatchImpl final : public Batch {
public:
    BatchImpl(const RunParams& rp, const BatchCapacity& cap, hipStream_t s, hipStream_t io)
        : P_(rp), Q_(make_params32(rp)), cap_(cap), stream_(s), geo_(geometry(rp)),
          nalloc_(cap.neurons), aalloc_(cap.astros),
          ealloc_{RangeAllocator(cap.edges[0]), RangeAllocator(cap.edges[1]), RangeAllocator(cap.edges[2]),
                  RangeAllocator(cap.edges[3])},
          hosts_(cap.members), threads_(rp.block) {
        lay_ = batch_layout(cap, rp);
        if (hipMalloc(&base_, lay_.total) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            throw Error(AS_ERR_NOMEM, "device allocation of " + std::to_string(lay_.total) + " bytes failed");
        }
        char* b = static_cast<char*>(base_);
        auto dptr = [&](std::size_t off) { return static_cast<void*>(b + off); };
        const std::uint64_t nn = std::max<std::uint64_t>(1, cap.neurons), na = std::max<std::uint64_t>(1, cap.astros);
        for (int i = 0; i < 4; ++i) {
            D_.in_rs[i] = (const uint2*)dptr(lay_.in_rs[i]);
            D_.in_src[i] = (const index_t*)dptr(lay_.in_src[i]);
            D_.out_rs[i] = (const uint2*)dptr(lay_.out_rs[i]);
            D_.out_ent[i] = (const uint2*)dptr(lay_.out_ent[i]);
            D_.bcnt[i] = (unsigned*)dptr(lay_.bcnt[i]);
            D_.bent[i] = (uint4*)dptr(lay_.bent[i]);
        }
        R* A[10] = {};
        for (int i = 0; i < (sizeof(R) == 4 ? 10 : 7); ++i) A[i] = (R*)dptr(lay_.astro[i]);
        S_.Ca = A[0]; S_.IP3 = A[1]; S_.h = A[2];
        S_.Ca_tot = A[3]; S_.IP3_0 = A[4]; S_.tau_IP3 = A[5]; S_.delta_IP3 = A[6];
        S_.Ca_lo = A[7]; S_.IP3_lo = A[8]; S_.h_lo = A[9];
        R* N[12];
        for (int i = 0; i < 12; ++i) N[i] = (R*)dptr(lay_.neuron[i]);
        S_.V = N[0]; S_.w = N[1]; S_.g_ex = N[2]; S_.dg_ex = N[3]; S_.g_in = N[4]; S_.dg_in = N[5];
        S_.I_sic = N[6]; S_.stp_x = N[7]; S_.stp_u = N[8]; S_.stp_t = N[9]; S_.V_reset = N[10]; S_.b = N[11];
        S_.spk_v = (R*)dptr(lay_.spk_v);
        S_.sic_v = (R*)dptr(lay_.sic_v);
        D_.ever = (std::uint8_t*)dptr(lay_.ever);
        D_.refr = (std::int32_t*)dptr(lay_.refr);
        D_.spk_n = (std::uint32_t*)dptr(lay_.spk_n);
        D_.spk_w = (std::uint32_t*)dptr(lay_.spk_w);
        D_.sic_w = (std::uint32_t*)dptr(lay_.sic_w);
        D_.counters = (unsigned long long*)dptr(lay_.counters);
        D_.overflow = (unsigned*)dptr(lay_.overflow);
        D_.members = (const Member*)dptr(lay_.members);
        D_.tiles = (const Tile*)dptr(lay_.tiles);
        spans_ = (TileSpan*)dptr(lay_.spans);
        own_io_ = io == nullptr;
        if (own_io_) {
            AS_HIP_CHECK(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
            AS_HIP_CHECK(hipStreamCreateWithFlags(&rd_, hipStreamNonBlocking));
        } else {
            up_ = rd_ = io;
        }
        AS_HIP_CHECK(hipEventCreateWithFlags(&enqueued_, hipEventDisableTiming));
        pin_slot_.resize(cap.members);
        pin_prefetch_.resize(cap.members);
        prefetched_.assign(cap.members, nullptr);
        AS_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&pin_members_), std::size_t(cap.members) * sizeof(Member)));
        AS_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&pin_spans_), std::size_t(cap.members) * sizeof(TileSpan)));
        AS_HIP_CHECK(hipEventCreateWithFlags(&tables_copied_, hipEventDisableTiming));
        D_.clock = nullptr;
        clock_ = (std::int64_t*)dptr(lay_.clock);
        D_.ncap = nn;
        D_.acap = na;
        D_.sstride = sizeof(R) == 4 ? na + nn : na;
        scratch0_ = (index_t*)dptr(lay_.scratch0);
        scratch1_ = (index_t*)dptr(lay_.scratch1);
        scan_tmp_ = dptr(lay_.scan_tmp);
        ctx_eager_ = reinterpret_cast<IntegrateCtx<R>*>(dptr(lay_.ctx));
        ctx_graph_ = ctx_eager_ + 1;
        {
            IntegrateCtx<R> host_ctx[2];
            host_ctx[0] = {D_, S_, P_, Q_};
            host_ctx[1] = host_ctx[0];
            host_ctx[1].D.clock = clock_;
            AS_HIP_CHECK(hipMemcpyAsync(ctx_eager_, host_ctx, sizeof host_ctx, hipMemcpyHostToDevice, stream_));
            AS_HIP_CHECK(hipStreamSynchronize(stream_));
        }
        AS_HIP_CHECK(hipMemsetAsync(D_.overflow, 0, sizeof(unsigned), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.counters, 0, std::size_t(cap.members) * kCountersPerMember * 8, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    ~BatchImpl() override {
        drop_graphs();
        // the block, upload and read streams finish before anything they use is freed
        (void)hipStreamSynchronize(stream_);
        if (up_) (void)hipStreamSynchronize(up_);
        if (rd_) (void)hipStreamSynchronize(rd_);
        if (base_) (void)hipFree(base_);
        if (own_io_ && up_) (void)hipStreamDestroy(up_);
        if (own_io_ && rd_) (void)hipStreamDestroy(rd_);
        for (auto& h : hosts_)
            if (h.up_done) (void)hipEventDestroy(h.up_done);
        for (std::size_t i = 0; i < pin_prefetch_.size(); ++i) {
            if (pin_prefetch_[i].p) (void)hipHostFree(pin_prefetch_[i].p);
            if (prefetched_[i]) (void)hipEventDestroy(prefetched_[i]);
        }
        for (Pinned& b : pin_slot_)
            if (b.p) (void)hipHostFree(b.p);
        for (Pinned& b : pin_free_) (void)hipHostFree(b.p);
        if (pin_read_.p) (void)hipHostFree(pin_read_.p);
        if (enqueued_) (void)hipEventDestroy(enqueued_);
        if (pin_members_) (void)hipHostFree(pin_members_);
        if (pin_spans_) (void)hipHostFree(pin_spans_);
        if (tables_copied_) (void)hipEventDestroy(tables_copied_);
    }

    bool fits(const Network& net) const override {
        if (free_slot() < 0) return false;
        if (!nalloc_.can(net.cfg.n_neurons()) || !aalloc_.can(net.cfg.n_astro)) return false;
        for (int i = 0; i < 4; ++i)
            if (!ealloc_[i].can(net.edges(i))) return false;
        return true;
    }

    int add(const Network& net, std::int64_t n_steps) override {
        if (step_ % geo_.L != 0) throw Error(AS_ERR_INVALID, "members join at block boundaries");
        const int slot = add_async(net, n_steps);
        AS_HIP_CHECK(hipEventSynchronize(hosts_[slot].up_done));
        activate(slot);
        rebuild_tiles();
        return slot;
    }

    void add_many(const Network* const* nets, int n, const std::int64_t* n_steps, int* members) override {
        if (step_ % geo_.L != 0) throw Error(AS_ERR_INVALID, "members join at block boundaries");
        for (int i = 0; i < n; ++i) members[i] = -1;
        auto join = [&](int upto) {
            AS_HIP_CHECK(hipStreamSynchronize(up_));
            for (int i = 0; i < upto; ++i) activate(members[i]);
            if (upto > 0) rebuild_tiles();
        };
        int i = 0;
        try {
            for (; i < n; ++i) members[i] = add_async(*nets[i], n_steps[i]);
        } catch (...) {
            join(i);
            throw;
        }
        join(n);
    }

    int add_async(const Network& net, std::int64_t n_steps) override {
        if (n_steps < 0) throw Error(AS_ERR_INVALID, "n_steps must be >= 0");
        release_due();
        if (!fits(net)) throw Error(AS_ERR_NOMEM, "network does not fit in the batch's free capacity");
        const int slot = free_slot();
        MemberHost& h = hosts_[slot];
        const index_t nn = net.cfg.n_neurons(), na = net.cfg.n_astro;
        Member& m = h.m;
        m.seed = net.seed;
        m.nb = static_cast<index_t>(nalloc_.take(nn));
        m.nn = nn;
        m.ne = net.cfg.n_exc;
        m.ab = static_cast<index_t>(aalloc_.take(na));
        m.na = na;
        m.step0 = step_;
        m.step_end = step_;
        m.slot = static_cast<std::uint32_t>(slot);
        for (int i = 0; i < 4; ++i) {
            h.ecount[i] = net.edges(i);
            h.ebase[i] = ealloc_[i].take(h.ecount[i]);
        }
        h.used = true;
        h.active = false;
        h.staging = true;
        h.pending_steps = n_steps;
        h.free_at_block = -1;
        // the ranges may have been released by a member whose last emissions
        // are still being delivered by enqueued work: upload after it
        AS_HIP_CHECK(hipEventRecord(enqueued_, stream_));
        AS_HIP_CHECK(hipStreamWaitEvent(up_, enqueued_, 0));
        const std::size_t pin_need = pinned_bytes(net.Ca_tot.size()) + pinned_bytes(net.IP3_0.size()) +
                                     pinned_bytes(net.tau_IP3.size()) + pinned_bytes(net.delta_IP3.size()) +
                                     pinned_bytes(net.V_reset.size()) + pinned_bytes(net.b.size());
        pin_slot_[slot] = take_pinned(pin_need);
        upload_member(net, h, up_, pin_slot_[slot]);
        if (h.up_done == nullptr) AS_HIP_CHECK(hipEventCreateWithFlags(&h.up_done, hipEventDisableTiming));
        AS_HIP_CHECK(hipEventRecord(h.up_done, up_));
        return slot;
    }

    int activate_ready(std::vector<int>* out) override {
        if (step_ % geo_.L != 0) return 0;
        int n = 0;
        for (std::size_t i = 0; i < hosts_.size(); ++i) {
            MemberHost& h = hosts_[i];
            if (!h.staging) continue;
            const hipError_t q = hipEventQuery(h.up_done);
            if (q == hipErrorNotReady) continue;
            AS_HIP_CHECK(q);
            activate(int(i));
            if (out) out->push_back(int(i));
            ++n;
        }
        if (n > 0) rebuild_tiles();
        return n;
    }

    void remove(int member) override {
        MemberHost& h = host(member);
        if (!h.active) return;
        h.active = false;
        const std::size_t slot = std::size_t(member);
        if (pin_prefetch_[slot].p) {  // prefetched but never read: the staging returns to the pool
            AS_HIP_CHECK(hipEventSynchronize(prefetched_[slot]));
            give_pinned(pin_prefetch_[slot]);
            pin_prefetch_[slot] = Pinned();
        }
        // emissions of the member's last block are still scattered d/L
        // blocks later (into its own ranges): release after that
        h.free_at_block = (step_ + geo_.L - 1) / geo_.L + geo_.d / geo_.L + 1;
        rebuild_tiles();
    }

    void enqueue_steps(std::int64_t k) override {
        flush_tables();
        while (k > 0) {
            if (step_ % geo_.L == 0) release_due();
            const std::int64_t bidx = step_ / geo_.L;
            BlkLaunch B = launch_for(bidx);
            B.t0 = step_;
            B.n = static_cast<int>(std::min<std::int64_t>({k, B.bstart + geo_.L - step_, std::int64_t(B.chunk)}));
            B.first = step_ == B.bstart;
            B.last = step_ + B.n == B.bstart + geo_.L;
            BatchDev D = D_;
            D.clock = nullptr;
            timing_begin();
            launch_integrate(D, B);
            timing_end();
            step_ += B.n;
            k -= B.n;
        }
        AS_HIP_CHECK(hipGetLastError());
        if (step_ % geo_.L == 0) release_due();
    }

    void run_blocks_graph(std::int64_t blocks) override {
        if (blocks <= 0) return;
        if (step_ % geo_.L != 0) throw Error(AS_ERR_INVALID, "graph replays start at a block boundary");
        release_due();
        flush_tables();
        hipGraphExec_t exec = graph_for(blocks);
        const std::int64_t b0 = step_ / geo_.L;
        AS_HIP_CHECK(hipMemcpyAsync(clock_, &b0, sizeof b0, hipMemcpyHostToDevice, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));  // b0 lives on the host stack
        timing_begin();
        AS_HIP_CHECK(hipGraphLaunch(exec, stream_));
        timing_end(blocks);
        step_ += blocks * geo_.L;
        release_due();
    }

    // Times `blocks_each` real blocks per (workgroup size, LDS chunk) pair
    // and keeps the fastest: each candidate workgroup size with the longest
    // chunk its LDS budget allows and with about half of it (half the
    // staging per lane, so more resident waves, against a second launch
    // per block). Results do not depend on either.
    int autotune(const int* cands, int n, int blocks_each) override {
        if (n <= 0) return threads_;
        if (step_ % geo_.L != 0) throw Error(AS_ERR_INVALID, "autotuning starts at a block boundary");
        double best_ms = 1e300;
        int best = threads_, best_cap = chunk_cap_;
        for (int i = 0; i < n; ++i) {
            const int c = cands[i];
            if (c < 64 || c > 1024 || c % 64 != 0) continue;
            set_threads(c);
            set_chunk_cap(0);
            const int full = chunk_len();
            const int caps[2] = {0, full >= 4 ? (full + 1) / 2 : 0};
            for (int k = 0; k < 2; ++k) {
                if (k == 1 && caps[1] == 0) break;
                set_chunk_cap(caps[k]);
                hipEvent_t e0, e1;
                AS_HIP_CHECK(hipEventCreate(&e0));
                AS_HIP_CHECK(hipEventCreate(&e1));
                AS_HIP_CHECK(hipEventRecord(e0, stream_));
                enqueue_steps(std::int64_t(blocks_each) * geo_.L);
                AS_HIP_CHECK(hipEventRecord(e1, stream_));
                AS_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                AS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = c;
                    best_cap = caps[k];
                }
            }
        }
        set_threads(best);
        set_chunk_cap(best_cap);
        return best;
    }

    int chunk_steps() const override { return chunk_len(); }

    int block_threads() const override { return threads_; }

    std::uint64_t spike_count(int member) override {
        const MemberHost& h = host(member);
        unsigned long long c[kCountersPerMember];
        AS_HIP_CHECK(hipMemcpyAsync(c, D_.counters + std::size_t(h.m.slot) * kCountersPerMember, sizeof c,
                                    hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        std::uint64_t t = 0;
        for (auto v : c) t += v;
        return t;
    }

    void reset_ever_active(int member) override {
        const MemberHost& h = host(member);
        AS_HIP_CHECK(hipMemsetAsync(D_.ever + h.m.ab, 0, h.m.na, stream_));
    }

    void read_state(int member, int field, double* out) override {
        const MemberHost& h = host(member);
        const R* src = nullptr;
        const R* lo = nullptr;
        std::size_t base = h.m.nb, n = h.m.nn;
        switch (field) {
            case 0: src = S_.Ca; lo = S_.Ca_lo; base = h.m.ab; n = h.m.na; break;
            case 1: src = S_.IP3; lo = S_.IP3_lo; base = h.m.ab; n = h.m.na; break;
            case 2: src = S_.h; lo = S_.h_lo; base = h.m.ab; n = h.m.na; break;
            case 3: src = S_.V; break;
            case 4: src = S_.w; break;
            case 5: src = S_.g_ex; break;
            case 6: src = S_.dg_ex; break;
            case 7: src = S_.g_in; break;
            case 8: src = S_.dg_in; break;
            case 9: src = S_.I_sic; break;
            case 10: src = S_.stp_x; break;
            case 11: src = S_.stp_u; break;
            case 12: src = S_.stp_t; break;
            default: throw Error(AS_ERR_INVALID, "unknown state field " + std::to_string(field));
        }
        std::vector<R> hv(n), lv;
        if (n) AS_HIP_CHECK(hipMemcpyAsync(hv.data(), src + base, n * sizeof(R), hipMemcpyDeviceToHost, stream_));
        if (sizeof(R) == 4 && lo && n) {
            lv.resize(n);
            AS_HIP_CHECK(hipMemcpyAsync(lv.data(), lo + base, n * sizeof(R), hipMemcpyDeviceToHost, stream_));
        }
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        for (std::size_t i = 0; i < n; ++i) out[i] = double(hv[i]) + (lv.empty() ? 0.0 : double(lv[i]));
    }

    void read_refractory(int member, std::int32_t* out) override {
        const MemberHost& h = host(member);
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.refr + h.m.nb, h.m.nn * 4, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_spike_counts(int member, std::uint32_t* out) override {
        const MemberHost& h = host(member);
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.spk_n + h.m.nb, h.m.nn * 4, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_ever_active(int member, std::uint8_t* out) override {
        const MemberHost& h = host(member);
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.ever + h.m.ab, h.m.na, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    SpikeWords spike_words(int member, std::int64_t first, int n) override {
        const MemberHost& h = host(member);
        const std::int64_t done = member_steps(member);
        const std::int64_t cur_block = step_ > 0 ? (step_ - 1) / geo_.L : 0;
        const std::int64_t g0 = h.m.step0 + first;
        if (n < 0 || n > kRecordMaxSteps || first < 0 || first + n > done ||
            (n > 0 && g0 / geo_.L <= cur_block - geo_.H))
            throw Error(AS_ERR_RANGE, "spike steps not held (the blocked method keeps the last H blocks)");
        SpikeWords w;
        w.kind = 1;
        w.base = D_.spk_w + h.m.nb;
        w.n_neurons = h.m.nn;
        w.n_steps = n;
        for (int i = 0; i < n; ++i) {
            const std::int64_t s = g0 + i, bl = s / geo_.L;
            w.off[i] = std::uint64_t(bl % geo_.H) * D_.ncap;
            w.bit[i] = static_cast<std::uint32_t>(s - bl * geo_.L);
        }
        return w;
    }

    void read_raster(int member, std::int64_t first, int n, std::uint8_t* out) override {
        const MemberHost& h = host(member);
        const std::int64_t done = member_steps(member);
        const std::int64_t cur_block = step_ > 0 ? (step_ - 1) / geo_.L : 0;
        const std::int64_t g0 = h.m.step0 + first;
        if (n < 0 || first < 0 || first + n > done || (n > 0 && g0 / geo_.L <= cur_block - geo_.H))
            throw Error(AS_ERR_RANGE, "raster steps not held (the blocked method keeps the last H blocks)");
        std::vector<std::uint32_t> words(h.m.nn);
        std::int64_t have = -1;
        for (int i = 0; i < n; ++i) {
            const std::int64_t s = g0 + i;
            const std::int64_t bl = s / geo_.L;
            if (bl != have) {
                AS_HIP_CHECK(hipMemcpyAsync(words.data(), D_.spk_w + std::size_t(bl % geo_.H) * D_.ncap + h.m.nb,
                                            h.m.nn * 4, hipMemcpyDeviceToHost, stream_));
                AS_HIP_CHECK(hipStreamSynchronize(stream_));
                have = bl;
            }
            const unsigned bit = static_cast<unsigned>(s - bl * geo_.L);
            for (index_t j = 0; j < h.m.nn; ++j) out[std::size_t(i) * h.m.nn + j] = (words[j] >> bit) & 1u;
        }
    }

    std::int64_t member_steps(int member) const override {
        const MemberHost& h = host(member);
        if (h.staging) return 0;
        return std::max<std::int64_t>(0, std::min(step_, h.m.step_end) - h.m.step0);
    }
    std::int64_t member_steps_at(int member, std::int64_t g) const override {
        const MemberHost& h = host(member);
        if (h.staging) return 0;
        return std::max<std::int64_t>(0, std::min(g, h.m.step_end) - h.m.step0);
    }

    // Copies a member's spike counters, Ca and (FP32c) Ca_lo into pinned
    // staging `base` on stream s.
    void copy_final(const MemberHost& h, char* base, hipStream_t s) {
        const std::size_t n = h.m.na;
        constexpr std::size_t kCountBytes = kCountersPerMember * sizeof(unsigned long long);
        AS_HIP_CHECK(hipMemcpyAsync(base, D_.counters + std::size_t(h.m.slot) * kCountersPerMember, kCountBytes,
                                    hipMemcpyDeviceToHost, s));
        if (n) AS_HIP_CHECK(hipMemcpyAsync(base + kCountBytes, S_.Ca + h.m.ab, n * sizeof(R), hipMemcpyDeviceToHost, s));
        if (sizeof(R) == 4 && n)
            AS_HIP_CHECK(hipMemcpyAsync(base + kCountBytes + pinned_bytes(n), S_.Ca_lo + h.m.ab, n * sizeof(R),
                                        hipMemcpyDeviceToHost, s));
    }

    // The results from staging filled by copy_final.
    void decode_final(const MemberHost& h, const char* base, std::uint64_t* spikes, double* ca) const {
        const std::size_t n = h.m.na;
        constexpr std::size_t kCountBytes = kCountersPerMember * sizeof(unsigned long long);
        const auto* c = reinterpret_cast<const unsigned long long*>(base);
        const R* hv = reinterpret_cast<const R*>(base + kCountBytes);
        const R* lv = reinterpret_cast<const R*>(base + kCountBytes + pinned_bytes(n));
        std::uint64_t t = 0;
        for (int i = 0; i < kCountersPerMember; ++i) t += c[i];
        *spikes = t;
        for (std::size_t i = 0; i < n; ++i) ca[i] = double(hv[i]) + (sizeof(R) == 4 ? double(lv[i]) : 0.0);
    }

    static std::size_t final_bytes(std::size_t n_astro) {
        return kCountersPerMember * sizeof(unsigned long long) + 2 * pinned_bytes(n_astro);
    }

    void prefetch_final(int member) override {
        const MemberHost& h = host(member);
        const std::size_t slot = std::size_t(member);
        if (pin_prefetch_[slot].p) return;  // already enqueued
        pin_prefetch_[slot] = take_pinned(final_bytes(h.m.na));
        copy_final(h, static_cast<char*>(pin_prefetch_[slot].p), stream_);
        if (!prefetched_[slot]) AS_HIP_CHECK(hipEventCreateWithFlags(&prefetched_[slot], hipEventDisableTiming));
        AS_HIP_CHECK(hipEventRecord(prefetched_[slot], stream_));
    }

    void read_final(int member, std::uint64_t* spikes, double* ca) override {
        const MemberHost& h = host(member);
        const std::size_t slot = std::size_t(member);
        if (pin_prefetch_[slot].p) {
            AS_HIP_CHECK(hipEventSynchronize(prefetched_[slot]));
            decode_final(h, static_cast<const char*>(pin_prefetch_[slot].p), spikes, ca);
            give_pinned(pin_prefetch_[slot]);
            pin_prefetch_[slot] = Pinned();
            return;
        }
        const std::size_t need = final_bytes(h.m.na);
        if (pin_read_.bytes < need) {
            if (pin_read_.p) AS_HIP_CHECK(hipHostFree(pin_read_.p));
            pin_read_ = Pinned();
            const std::size_t bytes = std::max<std::size_t>(need + need / 2, std::size_t(1) << 20);
            AS_HIP_CHECK(hipHostMalloc(&pin_read_.p, bytes, hipHostMallocDefault));
            pin_read_.bytes = bytes;
        }
        copy_final(h, static_cast<char*>(pin_read_.p), rd_);
        AS_HIP_CHECK(hipStreamSynchronize(rd_));
        decode_final(h, static_cast<const char*>(pin_read_.p), spikes, ca);
    }

    std::int64_t member_target(int member) const override { return host(member).m.step_end - host(member).m.step0; }
    bool member_active(int member) const override { return host(member).active; }
    int active_members() const override {
        int n = 0;
        for (auto& h : hosts_) n += h.active ? 1 : 0;
        return n;
    }
    std::uint64_t member_overflows() override {
        unsigned v = 0;
        AS_HIP_CHECK(hipMemcpyAsync(&v, D_.overflow, sizeof v, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        return v;
    }

    std::int64_t global_step() const override { return step_; }
    int block_len() const override { return geo_.L; }
    hipStream_t stream() const override { return stream_; }
    std::uint64_t device_bytes() const override { return lay_.total; }
    std::uint64_t free_neurons() const override { return nalloc_.largest(); }

    void set_timing(bool on) override { timing_ = on; }
    void kernel_time(double* total_ms, std::int64_t* launches) override {
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        double t = 0.0;
        for (auto& e : events_) {
            float ms = 0.0f;
            AS_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
            t += ms;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        events_.clear();
        *total_ms = t;
        *launches = timed_launches_;
        timed_launches_ = 0;
    }

private:
    const MemberHost& host(int member) const {
        if (member < 0 || member >= int(hosts_.size()) || !hosts_[member].used)
            throw Error(AS_ERR_INVALID, "no such batch member");
        return hosts_[member];
    }
    MemberHost& host(int member) { return const_cast<MemberHost&>(static_cast<const BatchImpl*>(this)->host(member)); }

    // An uploaded member joins now (a block boundary) for its pending steps.
    void activate(int slot) {
        MemberHost& h = hosts_[slot];
        h.m.step0 = step_;
        h.m.step_end = step_ + h.pending_steps;
        h.staging = false;
        h.active = true;
        give_pinned(pin_slot_[slot]);  // its upload has completed
        pin_slot_[slot] = Pinned();
    }

    int free_slot() const {
        for (std::size_t i = 0; i < hosts_.size(); ++i)
            if (!hosts_[i].used) return int(i);
        return -1;
    }

    BlkLaunch launch_for(std::int64_t bidx) const {
        BlkLaunch B{};
        B.bstart = bidx * geo_.L;
        B.t0 = B.bstart;
        B.n = geo_.L;
        B.d = geo_.d;
        B.L = geo_.L;
        B.H = geo_.H;
        B.Rv = geo_.Rv;
        B.wslot = static_cast<std::uint32_t>(bidx % geo_.H);
        B.rslot = static_cast<std::uint32_t>(((bidx - geo_.d / geo_.L) % geo_.H + geo_.H) % geo_.H);
        B.bw = static_cast<std::uint32_t>((bidx + geo_.d / geo_.L) % geo_.H);
        B.first = 1;
        B.last = 1;
        B.chunk = chunk_len();
        B.lane_words = lane_words(B.chunk);
        return B;
    }

    // LDS words (of R) per lane that an integrate launch of c steps stages:
    // excitatory and inhibitory inputs per step and the noise values of up to
    // (c-1)/steps_per_noise + 2 noise indices, except where the kernel keeps
    // those in registers (FP64 with at most two per launch).
    int lds_words(int c) const {
        const int extra = (c - 1) / P_.steps_per_noise;
        return 2 * c + (sizeof(R) == 4 || extra > 0 ? extra + 2 : 0);
    }
    // A lane's slots in LDS (slot-major in its wave's region: no padding is
    // needed against bank conflicts).
    int lane_words(int c) const { return lds_words(c); }
    // Most steps per integrate launch: the staged words fit 208 bytes per
    // lane (3 resident waves per SIMD share 160 KB) and 64 KB per workgroup.
    int chunk_len() const {
        const int per_lane = std::min(208, 65536 / threads_);
        int c = chunk_cap_ > 0 ? std::min(chunk_cap_, geo_.L) : geo_.L;
        while (c > 1 && lane_words(c) * int(sizeof(R)) > per_lane) --c;
        return c;
    }

    void launch_integrate(const BatchDev& D, const BlkLaunch& B) {
        if (n_tiles_ == 0) return;
        const IntegrateCtx<R>* ctx = D.clock ? ctx_graph_ : ctx_eager_;
        const std::size_t lds = std::size_t(B.lane_words) * std::size_t(threads_) * sizeof(R);
        if (threads_ <= 256) {
            if (B.chunk == kDefaultChunk)
                hipLaunchKernelGGL((integrate_kernel<256, R, kDefaultChunk>), dim3(n_tiles_), dim3(threads_), lds,
                                   stream_, ctx, B);
            else
                hipLaunchKernelGGL((integrate_kernel<256, R, 0>), dim3(n_tiles_), dim3(threads_), lds, stream_, ctx, B);
        } else {
            if (B.chunk == kDefaultChunk)
                hipLaunchKernelGGL((integrate_kernel<1024, R, kDefaultChunk>), dim3(n_tiles_), dim3(threads_), lds,
                                   stream_, ctx, B);
            else
                hipLaunchKernelGGL((integrate_kernel<1024, R, 0>), dim3(n_tiles_), dim3(threads_), lds, stream_, ctx,
                                   B);
        }
    }

    void timing_begin() {
        if (!timing_) return;
        AS_HIP_CHECK(hipEventCreate(&pending_));
        AS_HIP_CHECK(hipEventRecord(pending_, stream_));
    }
    void timing_end(std::int64_t launches = 1) {
        if (!timing_ || !pending_) return;
        hipEvent_t stop;
        AS_HIP_CHECK(hipEventCreate(&stop));
        AS_HIP_CHECK(hipEventRecord(stop, stream_));
        events_.emplace_back(pending_, stop);
        pending_ = nullptr;
        timed_launches_ += launches;
    }

    void set_threads(int t) {
        if (t == threads_) return;
        threads_ = t;
        rebuild_tiles();
    }
    // Caps the LDS chunk (0: the longest the budget allows); captured
    // graphs hold the old launch shapes.
    void set_chunk_cap(int c) {
        if (c == chunk_cap_) return;
        chunk_cap_ = c;
        drop_graphs();
    }

    // The member table and the tile table of the active members. The host
    // writes the member table and one span per active member into pinned
    // staging (after the previous copies out of it have completed) and a
    // kernel writes the tiles, so a change costs a few kilobytes of copy
    // and no synchronisation however many tiles the batch holds.
    // Membership changed: the tables are rewritten before the next launch
    // (flush_tables), once however many members joined or left meanwhile.
    // Writing them waits for the previous copy out of the pinned staging;
    // deferred to the next launch, that copy precedes the blocks already
    // enqueued instead of following them.
    void rebuild_tiles() {
        tables_dirty_ = true;
        drop_graphs();
    }
    void flush_tables() {
        if (!tables_dirty_) return;
        write_tables();
        tables_dirty_ = false;
    }
    void write_tables() {
        AS_HIP_CHECK(hipEventSynchronize(tables_copied_));
        std::uint32_t n_active = 0;
        std::uint64_t total = 0;
        const std::uint32_t T = static_cast<std::uint32_t>(threads_);
        for (std::size_t i = 0; i < hosts_.size(); ++i) {
            const MemberHost& h = hosts_[i];
            pin_members_[i] = h.m;
            if (!h.active) continue;
            const std::uint64_t nt = (h.m.na + T - 1) / T + (h.m.ne + T - 1) / T + (h.m.nn - h.m.ne + T - 1) / T;
            pin_spans_[n_active++] = TileSpan{std::uint32_t(i), h.m.na, h.m.ne, h.m.nn, std::uint32_t(total), 0u};
            total += nt;
        }
        if (total > lay_.tiles_cap) throw Error(AS_ERR_NOMEM, "tile table capacity exceeded");
        n_tiles_ = static_cast<unsigned>(total);
        AS_HIP_CHECK(hipMemcpyAsync(const_cast<Member*>(D_.members), pin_members_, hosts_.size() * sizeof(Member),
                                    hipMemcpyHostToDevice, stream_));
        if (n_active > 0) {
            AS_HIP_CHECK(hipMemcpyAsync(spans_, pin_spans_, n_active * sizeof(TileSpan), hipMemcpyHostToDevice, stream_));
            hipLaunchKernelGGL(tiles_kernel, dim3(n_active), dim3(256), 0, stream_, spans_, T,
                               const_cast<Tile*>(D_.tiles));
            AS_HIP_CHECK(hipGetLastError());
        }
        AS_HIP_CHECK(hipEventRecord(tables_copied_, stream_));
    }

    void release_due() {
        const std::int64_t blk = step_ / geo_.L;
        for (auto& h : hosts_) {
            if (h.used && !h.active && h.free_at_block >= 0 && blk >= h.free_at_block) {
                nalloc_.give(h.m.nb, h.m.nn);
                aalloc_.give(h.m.ab, h.m.na);
                for (int i = 0; i < 4; ++i) ealloc_[i].give(h.ebase[i], h.ecount[i]);
                const hipEvent_t ev = h.up_done;  // reused by the slot's next member
                h = MemberHost();
                h.up_done = ev;
            }
        }
    }

    void drop_graphs() {
        for (auto& g : graphs_) {
            (void)hipGraphExecDestroy(g.second.second);
            (void)hipGraphDestroy(g.second.first);
        }
        graphs_.clear();
    }

    hipGraphExec_t graph_for(std::int64_t blocks) {
        auto it = graphs_.find(blocks);
        if (it != graphs_.end()) return it->second.second;
        BatchDev D = D_;
        D.clock = clock_;
        const BlkLaunch B = launch_for(0);  // the kernels take the block index from the clock
        AS_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        for (std::int64_t i = 0; i < blocks; ++i) {
            for (int sub = 0; sub < geo_.L; sub += B.chunk) {
                BlkLaunch Bs = B;
                Bs.sub = sub;
                launch_integrate(D, Bs);
            }
            hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(64), 0, stream_, clock_);
        }
        hipGraph_t g = nullptr;
        AS_HIP_CHECK(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t exec = nullptr;
        AS_HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        graphs_[blocks] = {g, exec};
        return exec;
    }

    template <class T>
    void fill(T* p, T v, std::size_t n) { fill_on(stream_, p, v, n); }
    template <class T>
    void fill_on(hipStream_t s, T* p, T v, std::size_t n) {
        if (n == 0) return;
        hipLaunchKernelGGL(fill_kernel<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, v, n);
        AS_HIP_CHECK(hipGetLastError());
    }

    // Pinned host memory (hipHostMalloc): copies from and to it run on their
    // own stream beside other streams' kernels, where a copy through
    // pageable memory waited for them here.
    struct Pinned {
        void* p = nullptr;
        std::size_t bytes = 0, used = 0;
    };
    static std::size_t pinned_bytes(std::size_t n) { return (n * sizeof(R) + 255) & ~std::size_t(255); }
    Pinned take_pinned(std::size_t need) {
        std::size_t best = pin_free_.size();
        for (std::size_t i = 0; i < pin_free_.size(); ++i)
            if (pin_free_[i].bytes >= need && (best == pin_free_.size() || pin_free_[i].bytes < pin_free_[best].bytes))
                best = i;
        Pinned b;
        if (best < pin_free_.size()) {
            b = pin_free_[best];
            pin_free_.erase(pin_free_.begin() + std::ptrdiff_t(best));
        } else {
            b.bytes = std::max<std::size_t>(need, std::size_t(1) << 20);
            AS_HIP_CHECK(hipHostMalloc(&b.p, b.bytes, hipHostMallocDefault));
        }
        b.used = 0;
        return b;
    }
    void give_pinned(const Pinned& b) {
        if (b.p) pin_free_.push_back(b);
    }

    // The copy reads a converted array in the member's pinned staging, which
    // is returned to the pool at its activation (its upload has completed).
    void upload_values(R* dst, const std::vector<double>& v, bool reciprocal, hipStream_t s, Pinned& pb) {
        if (v.empty()) return;
        const std::size_t bytes = pinned_bytes(v.size());
        if (pb.used + bytes > pb.bytes) throw Error(AS_ERR_INVALID, "pinned staging smaller than the network's parameters");
        R* f = reinterpret_cast<R*>(static_cast<char*>(pb.p) + pb.used);
        pb.used += bytes;
        for (std::size_t i = 0; i < v.size(); ++i) f[i] = R(reciprocal ? 1.0 / v[i] : v[i]);
        AS_HIP_CHECK(hipMemcpyAsync(dst, f, v.size() * sizeof(R), hipMemcpyHostToDevice, s));
    }

    // Uploads one member: rebased in-CSRs, out-CSRs built on the device,
    // parameters and initial state (src/astrosim_fused.cu:981-1018).
    void upload_member(const Network& net, MemberHost& h, hipStream_t s, Pinned& staged) {
        const Member& m = h.m;
        const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
        // source-id base and source-space size of each type (global ids)
        const index_t sbase[4] = {m.nb, m.nb, m.ab, m.nb};
        const index_t tbase[4] = {m.nb, m.nb, m.nb, m.ab};
        const index_t nsrc[4] = {m.ne, net.cfg.n_inh, m.na, m.ne};
        const index_t nspace[4] = {m.nn, m.nn, m.na, m.nn};
        // a network generated on this GPU is copied device to device
        const bool dev = net.on_device();
        for (int T = 0; T < 4; ++T) {
            const index_t nt = dev ? net.dev[T].n_targets : c[T]->n_targets();
            const std::size_t ne = net.edges(T);
            const index_t eb = static_cast<index_t>(h.ebase[T]);
            index_t* src = const_cast<index_t*>(D_.in_src[T]) + eb;
            // host CSRs from their pinned copy when there is one
            const index_t* h_src = net.pin_src[T] ? net.pin_src[T] : c[T]->src.data();
            const index_t* h_rs = net.pin_rs[T] ? net.pin_rs[T] : c[T]->row_start.data();
            if (ne) {
                AS_HIP_CHECK(hipMemcpyAsync(src, dev ? net.dev[T].src : h_src, ne * sizeof(index_t),
                                            dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
                if (sbase[T] != 0)
                    hipLaunchKernelGGL(rebase_kernel, dim3(unsigned((ne + 255) / 256)), dim3(256), 0, s, src, ne,
                                       sbase[T]);
            }
            AS_HIP_CHECK(hipMemcpyAsync(scratch0_, dev ? net.dev[T].row_start : h_rs,
                                        (std::size_t(nt) + 1) * sizeof(index_t),
                                        dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
            uint2* in_rs = const_cast<uint2*>(D_.in_rs[T]) + tbase[T];
            if (nt) hipLaunchKernelGGL(rows_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, scratch0_, nt, eb, in_rs);
            // out-CSR over the member's source space
            const index_t ns = nsrc[T];
            AS_HIP_CHECK(hipMemsetAsync(scratch0_, 0, (std::size_t(ns) + 1) * sizeof(index_t), s));
            if (ne)
                hipLaunchKernelGGL(count_sources_kernel, dim3(unsigned((ne + 255) / 256)), dim3(256), 0, s,
                                   D_.in_src[T] + eb, ne, sbase[T], ns, scratch0_);
            std::size_t tmp = lay_.scan_tmp_bytes;
            AS_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp_, tmp, scratch0_, scratch1_, int(ns) + 1, s));
            if (nspace[T])
                hipLaunchKernelGGL(out_rows_kernel, dim3((nspace[T] + 255) / 256), dim3(256), 0, s, scratch1_, ns,
                                   nspace[T], sbase[T], eb, const_cast<uint2*>(D_.out_rs[T]));
            if (nt)
                hipLaunchKernelGGL(fill_out_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, D_.in_rs[T],
                                   D_.in_src[T], tbase[T], nt, sbase[T], ns, eb, scratch1_,
                                   const_cast<uint2*>(D_.out_ent[T]));
            AS_HIP_CHECK(hipGetLastError());
            // the member's bucket counts in every set
            const std::size_t set_stride = T == T_E2A ? D_.acap : D_.ncap;
            for (int set = 0; set < geo_.H; ++set)
                AS_HIP_CHECK(hipMemsetAsync(D_.bcnt[T] + std::size_t(set) * set_stride + tbase[T], 0,
                                            std::size_t(nt) * sizeof(unsigned), s));
            // the scratch arrays are reused by the next type in stream order
        }
        // parameters: FP32c stores 1/tau_IP3
        upload_values(S_.Ca_tot + m.ab, net.Ca_tot, false, s, staged);
        upload_values(S_.IP3_0 + m.ab, net.IP3_0, false, s, staged);
        upload_values(S_.tau_IP3 + m.ab, net.tau_IP3, sizeof(R) == 4, s, staged);
        upload_values(S_.delta_IP3 + m.ab, net.delta_IP3, false, s, staged);
        upload_values(S_.V_reset + m.nb, net.V_reset, false, s, staged);
        upload_values(S_.b + m.nb, net.b, false, s, staged);
        const Biology bio;
        fill_on(s, S_.Ca + m.ab, R(bio.Ca_init), m.na);
        fill_on(s, S_.IP3 + m.ab, R(bio.IP3_init), m.na);
        fill_on(s, S_.h + m.ab, R(bio.h_init), m.na);
        if constexpr (sizeof(R) == 4) {  // rounding residue of the initial values
            fill_on(s, S_.Ca_lo + m.ab, R(bio.Ca_init - double(float(bio.Ca_init))), m.na);
            fill_on(s, S_.IP3_lo + m.ab, R(bio.IP3_init - double(float(bio.IP3_init))), m.na);
            fill_on(s, S_.h_lo + m.ab, R(bio.h_init - double(float(bio.h_init))), m.na);
        }
        fill_on(s, S_.V + m.nb, R(bio.e_V_m), m.ne);
        fill_on(s, S_.V + m.nb + m.ne, R(bio.i_V_m), m.nn - m.ne);
        for (R* p : {S_.w, S_.g_ex, S_.dg_ex, S_.g_in, S_.dg_in, S_.I_sic}) fill_on(s, p + m.nb, R(0), m.nn);
        fill_on(s, S_.stp_x + m.nb, R(1), m.nn);
        fill_on(s, S_.stp_u + m.nb, R(P_.stp_U), m.nn);
        fill_on(s, S_.stp_t + m.nb, R(-1e12), m.nn);
        fill_on(s, D_.refr + m.nb, std::int32_t(0), m.nn);
        fill_on(s, D_.spk_n + m.nb, std::uint32_t(0), m.nn);
        fill_on(s, D_.ever + m.ab, std::uint8_t(0), m.na);
        for (int hs = 0; hs < geo_.H; ++hs) {
            fill_on(s, D_.spk_w + std::size_t(hs) * D_.ncap + m.nb, std::uint32_t(0), m.nn);
            fill_on(s, D_.sic_w + std::size_t(hs) * D_.acap + m.ab, std::uint32_t(0), m.na);
        }
        fill_on(s, D_.counters + std::size_t(m.slot) * kCountersPerMember, 0ull, kCountersPerMember);
    }

    RunParams P_;
    Params32 Q_;
    BatchCapacity cap_;
    hipStream_t stream_;
    Geometry geo_;
    RangeAllocator nalloc_, aalloc_;
    RangeAllocator ealloc_[4];
    std::vector<MemberHost> hosts_;
    int threads_;
    unsigned n_tiles_ = 0;
    bool tables_dirty_ = false;  // membership changed since the tables were written
    BatchLayout lay_{};
    void* base_ = nullptr;
    BatchDev D_{};
    BlkState<R> S_{};
    std::int64_t* clock_ = nullptr;
    IntegrateCtx<R>* ctx_eager_ = nullptr;
    IntegrateCtx<R>* ctx_graph_ = nullptr;
    index_t *scratch0_ = nullptr, *scratch1_ = nullptr;
    void* scan_tmp_ = nullptr;
    TileSpan* spans_ = nullptr;
    Member* pin_members_ = nullptr;  // pinned staging of the member and span tables
    TileSpan* pin_spans_ = nullptr;
    hipEvent_t tables_copied_ = nullptr;
    hipStream_t up_ = nullptr;          // member uploads (add_async)
    hipStream_t rd_ = nullptr;          // reads of finished members (read_final)
    bool own_io_ = true;                // up_ and rd_ are the batch's own (else one shared stream)
    std::vector<Pinned> pin_prefetch_;  // per slot: prefetch_final's staging (p == nullptr: none)
    std::vector<hipEvent_t> prefetched_;  // per slot: its copies' completion
    hipEvent_t enqueued_ = nullptr;     // the batch stream's work at an add_async
    std::vector<Pinned> pin_slot_;  // per slot: the converted parameters of its upload, until activation
    std::vector<Pinned> pin_free_;  // pinned staging buffers not in use
    Pinned pin_read_;               // read_final's destination
    std::int64_t step_ = 0;
    std::map<std::int64_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs_;
    int chunk_cap_ = 0;  // most steps per integrate launch (0: as many as the LDS budget allows)
    bool timing_ = false;
    hipEvent_t pending_ = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    std::int64_t timed_launches_ = 0;
};

// A single simulation on the blocked method: a batch with one member whose
// capacity is exactly its network.
class BlockedSim final : public Engine {
public:
    BlockedSim(const Network& net, const RunParams& rp, hipStream_t s) {
        stream_ = s;
        n_neurons_ = net.cfg.n_neurons();
        n_astro_ = net.cfg.n_astro;
        n_exc_ = net.cfg.n_exc;
        const Network* p = &net;
        batch_ = make_batch(rp, capacity_for(&p, 1), s);
        member_ = batch_->add(net, std::numeric_limits<std::int64_t>::max() / 4);
    }
    void enqueue_steps(std::int64_t k) override {
        batch_->enqueue_steps(k);
        step_ += k;
    }
    std::uint64_t spike_count() override { return batch_->spike_count(member_); }
    void reset_ever_active() override { batch_->reset_ever_active(member_); }
    void read_state(int field, double* out) override { batch_->read_state(member_, field, out); }
    void read_refractory(std::int32_t* out) override { batch_->read_refractory(member_, out); }
    void read_ever_active(std::uint8_t* out) override { batch_->read_ever_active(member_, out); }
    void read_spike_counts(std::uint32_t* out) override { batch_->read_spike_counts(member_, out); }
    void read_raster(std::int64_t first, int n, std::uint8_t* out) override {
        batch_->read_raster(member_, first, n, out);
    }
    SpikeWords spike_words(std::int64_t first, int n) override { return batch_->spike_words(member_, first, n); }
    std::uint64_t device_bytes() const override { return batch_->device_bytes(); }
    void set_timing(bool on) override { batch_->set_timing(on); }
    void kernel_time(double* total_ms, std::int64_t* launches) override { batch_->kernel_time(total_ms, launches); }

private:
    std::unique_ptr<Batch> batch_;
    int member_ = 0;
};

}  // namespace

BatchCapacity capacity_for(const Network* const* nets, int n) {
    BatchCapacity c{};
    c.members = static_cast<std::uint32_t>(std::max(1, n));
    for (int i = 0; i < n; ++i) {
        const Network& net = *nets[i];
        c.neurons += net.cfg.n_neurons();
        c.astros += net.cfg.n_astro;
        for (int k = 0; k < 4; ++k) c.edges[k] += net.edges(k);
    }
    return c;
}

std::uint64_t batch_footprint(const BatchCapacity& cap, const RunParams& rp) { return batch_layout(cap, rp).total; }

std::unique_ptr<Batch> make_batch(const RunParams& rp, const BatchCapacity& cap, hipStream_t s, hipStream_t io) {
    if (cap.neurons + cap.astros >= (std::uint64_t(1) << 32) ||
        std::max({cap.edges[0], cap.edges[1], cap.edges[2], cap.edges[3]}) >= (std::uint64_t(1) << 32))
        throw Error(AS_ERR_INVALID, "batch capacity exceeds 32-bit ids");
    if (std::max(cap.neurons, cap.astros) > kSrcMask)
        throw Error(AS_ERR_INVALID, "batch capacity exceeds 2^26 cells of one kind (bucket entry source ids)");
    if (rp.precision == AS_FP64) return std::unique_ptr<Batch>(new BatchImpl<double>(rp, cap, s, io));
    return std::unique_ptr<Batch>(new BatchImpl<float>(rp, cap, s, io));
}

std::uint64_t blocked_footprint(const Network& net, const RunParams& rp) {
    const Network* p = &net;
    return batch_footprint(capacity_for(&p, 1), rp);
}

std::unique_ptr<Engine> make_blocked_engine(const Network& net, const RunParams& rp, hipStream_t s) {
    return std::unique_ptr<Engine>(new BlockedSim(net, rp, s));
}

}  // namespace astrosim
