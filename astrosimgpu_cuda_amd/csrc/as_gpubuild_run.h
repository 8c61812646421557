// as_gpubuild_run.h -- host orchestration of the parallel connectivity
// generator (the bodies and the method are in as_gpubuild.h), written once
// over a backend that provides memory, launches and the two library
// primitives (exclusive scan, stable key-value sort):
//
//   template <class T> T* alloc(u64 n);       void release(void* p);
//   void h2d(void* d, const void* h, u64 bytes);   // both synchronous
//   void d2h(void* h, const void* d, u64 bytes);
//   void zero(void* d, u64 bytes);
//   template <class F> void run(u64 n, const F& f);  // f(i), i < n
//   void scan_u64(u64* p, u64 n);                    // exclusive, in place
//   void scan_u32(u32* p, u64 n);                    // exclusive, in place
//   void sort_pairs(const u32* kin, const u32* vin, u32* kout, u32* vout, u64 n, int bits);  // stable
//
// as_gpubuild.hip instantiates it with HIP; a host emulation instantiates it
// for tests that run without a GPU.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "as_gpubuild.h"
#include "as_setup.h"

namespace astrosim {
namespace gb {

struct BuildStats {
    u64 counters = 0;      // decoded draws
    u64 chains = 0;
    u32 ambiguous = 0;     // skips decided by the C library
    u32 host_rows[3] = {0, 0, 0};  // rows walked on the host before a pass joined the chains
    u64 host_walked = 0;           // rows walked on the host where a chain met no later one
    u32 records[3] = {0, 0, 0};    // rows with a target near their own position
    u32 repairs[3] = {0, 0, 0};    // autapses that moved the walk off the chains
    int max_shift = 0;
    double seconds = 0.0;
    // seconds since the start at: decode, chains, passes 0-2 stitched,
    // counts, edges emitted, the four inversions
    double mark[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double pools_seconds = 0.0, params_wait_seconds = 0.0, params_seconds = 0.0;
    u64 max_b_draws = 0;      // longest chain walk past the next anchor before it met another chain
    u32 max_repair_rows = 0;  // longest repair walk, in rows
};

// The draw at counter c decoded on the host, with the C library deciding
// close skips exactly as geometric_skip does (src/astrosim_fused.cu:617-624).
inline u32 host_entry(u64 seed, u64 c, double log1m_p, u32 take) {
    const double r = rng_uniform(seed, kSetupStream, c);
    const double u = skip_uniform(r);
    u32 g;
    if (!skip_fast(u, log1m_p, g)) {
        const double q = std::floor(std::log1p(-u) / log1m_p);
        g = q >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<u32>(q);
    }
    return encode_entry(r, g, take);
}

struct FillF { FillArgs a; AS_HD void operator()(u64 i) const { table_fill(a, i); } };
struct ChainAF { ChainArgs a; AS_HD void operator()(u64 i) const { chain_a(a, i); } };
struct ChainBF { ChainArgs a; AS_HD void operator()(u64 i) const { chain_b(a, i); } };
struct PathF { PathArgs a; AS_HD void operator()(u64 i) const { path_write(a, i); } };
struct VerifyF { VerifyArgs a; AS_HD void operator()(u64 i) const { verify_row(a, i); } };
struct RepairF { RepairArgs a; AS_HD void operator()(u64 i) const { repair_row(a, i); } };
struct RepairCopyF { RepairArgs a; AS_HD void operator()(u64 i) const { repair_copy(a, i); } };
struct CopyF { CopyArgs a; AS_HD void operator()(u64 i) const { splice_copy(a, i); } };
struct CountF { EmitArgs a; AS_HD void operator()(u64 i) const { exact_row<false>(a, i); } };
struct EmitF { EmitArgs a; AS_HD void operator()(u64 i) const { exact_row<true>(a, i); } };
struct AstroF { AstroArgs a; AS_HD void operator()(u64 i) const { tf_astro(a, i); } };
struct HistF { HistArgs a; AS_HD void operator()(u64 i) const { histogram(a, i); } };
struct MarkF { MarkArgs a; AS_HD void operator()(u64 i) const { mark_list(a, i); } };
struct WidenF {
    const u32* in;
    u64* out;
    AS_HD void operator()(u64 i) const { out[i] = in[i]; }
};

// Frees every allocation it made when it goes out of scope.
template <class BK>
struct Arena {
    BK& bk;
    std::vector<void*> held;
    explicit Arena(BK& b) : bk(b) {}
    ~Arena() {
        for (void* p : held) bk.release(p);
    }
    template <class T>
    T* get(u64 n) {
        T* p = bk.template alloc<T>(n == 0 ? 1 : n);
        held.push_back(p);
        return p;
    }
    void drop(void* p) {
        auto it = std::find(held.begin(), held.end(), p);
        if (it != held.end()) {
            bk.release(p);
            held.erase(it);
        }
    }
    // Hands an allocation over to the caller (no longer freed here).
    void keep(void* p) {
        auto it = std::find(held.begin(), held.end(), p);
        if (it != held.end()) held.erase(it);
    }
};

struct Declined {
    std::string why;
};

// Tuning: anchors every kW counters; rows with a drawn target within kK of
// their own position are checked for autapses (the index shift the repairs
// accumulate must stay within kK).
constexpr u64 kW = 4096;
constexpr int kK = 12;
constexpr u32 kRecordCap = 1u << 18;
constexpr u32 kRepairRows = 2048;
// Row starts stored per chain past its region (more are walked again when
// the path needs them; a test lowers the count to exercise that walk).
constexpr u32 kCapB = 2048;
constexpr u32 kAmbCap = 1u << 16;

// Builds the four target-major CSRs of `net` (cfg and seed already set) from
// the setup stream whose edge draws start at a counter c0 >= c0_lo. The
// astrocyte pools come from pools_fn(pool) (which fills [neuron][take] and
// returns c0); it is first called once the draws are decoded and the chains
// walked from c0_lo, so the pools can be drawn meanwhile. Returns false with a reason when the
// network is outside what the method handles (the caller then builds it on
// the host); the result, when true, equals build_network's bit for bit.
// With `kept`, the CSRs stay in the backend's memory (net.dev; the host CSRs
// are left empty) and their allocations are appended to *kept for the
// caller to free.
template <class BK, class Pools>
bool build_connectivity(BK& bk, const ScalingConfig& cfg, u64 seed, u64 c0_lo, Pools&& pools_fn, u32 take,
                        Network& net, BuildStats& st, std::string& why, std::vector<void*>* kept = nullptr,
                        u32 cap_b = kCapB) {
    const auto t0 = std::chrono::steady_clock::now();
    auto mark = [&](int k) {
        st.mark[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    Arena<BK> ar(bk);
    try {
        const u32 n_exc = cfg.n_exc, n_inh = cfg.n_inh, n_neu = cfg.n_exc + cfg.n_inh, n_astro = cfg.n_astro;
        const double p = cfg.p_primary;
        if (!(p > 0.0 && p < 1.0)) throw Declined{"p_primary must be in (0, 1)"};
        const double log1m_p = std::log1p(-p);
        const PassDesc P[3] = {{n_exc, 1u, n_exc, 0u, 0u},
                               {n_inh, 1u, n_exc, n_exc, kNone},
                               {n_neu, 0u, n_inh, 0u, n_exc}};

        // ---- regions: expected draws per pass with generous margins ----
        double D[3], S2[3];
        for (int q = 0; q < 3; ++q) {
            const double mu = double(P[q].pc) * p;
            D[q] = double(P[q].n_rows) * (mu * (P[q].tf ? 2.2 : 1.0) + 1.0);
            S2[q] = double(P[q].n_rows) * (mu * (P[q].tf ? 5.0 : 1.0) + 1.0);
        }
        auto margin = [](double var) { return u64(16.0 * std::sqrt(var)) + 4 * kW + 65536; };
        u64 r0[3], rend[3];
        const u64 c0 = c0_lo;  // regions and the table start here; the true start is not earlier
        double start_est = double(c0), var_start = 0.0;
        for (int q = 0; q < 3; ++q) {
            const u64 lo_m = q == 0 ? 0 : margin(var_start);
            const u64 est = u64(start_est);
            r0[q] = est > c0 + lo_m ? est - lo_m : c0;
            if (q == 0) r0[q] = c0;
            r0[q] = c0 + (r0[q] - c0) / 32 * 32;  // bitmaps share the table's alignment
            var_start += S2[q];
            start_est += D[q];
            rend[q] = u64(start_est) + margin(var_start);
        }
        ChainArgs C{};
        u64 G = 0;
        for (int q = 0; q < 3; ++q) {
            const u64 nch = (rend[q] - r0[q] + kW - 1) / kW;
            if (nch >= (u64(1) << 31)) throw Declined{"too many chains"};
            C.pass[q] = ChainPass{P[q], r0[q], static_cast<u32>(nch), static_cast<u32>(G)};
            rend[q] = r0[q] + nch * kW;
            G += nch;
        }
        const u64 base = c0;
        const u64 len = std::max({rend[0], rend[1], rend[2]}) - base + 65536;
        const u64 est_edges = u64(ScalingConfig::kSyn * double(n_neu) * 1.2) + 4096;
        if (est_edges >= (u64(1) << 31)) throw Declined{"edge count beyond the 32-bit sort range"};
        st.counters = len;
        st.chains = G;

        // ---- 1. decode every counter ----
        // whole chunks plus one: a walk's next-chunk request stays inside
        u32* E = ar.template get<u32>((len + kChunk - 1) / kChunk * kChunk + kChunk);
        u32* flags = ar.template get<u32>(8);  // [0] ambiguous count, [1] fail bits, [2] records
        bk.zero(flags, 8 * sizeof(u32));
        u64* amb = ar.template get<u64>(kAmbCap);
        bk.run(len, FillF{FillArgs{E, base, len, seed, log1m_p, take, kAmbCap, flags + 0, amb}});
        u32 hflags[8];
        bk.d2h(hflags, flags, sizeof hflags);
        st.ambiguous = hflags[0];
        if (hflags[0] > kAmbCap) throw Declined{"too many close skips"};
        if (hflags[0] > 0) {
            std::vector<u64> hamb(hflags[0]);
            bk.d2h(hamb.data(), amb, hamb.size() * sizeof(u64));
            for (u64 c : hamb) {
                const u32 e = host_entry(seed, c, log1m_p, take);
                bk.h2d(E + (c - base), &e, sizeof e);
            }
        }
        const Table T{E, base, len};
        mark(0);

        // ---- 2. chains ----
        C.T = T;
        C.W = kW;
        C.capA = static_cast<u32>(kW / 8 + 64);
        C.A = ar.template get<u32>(G * C.capA);
        C.nA = ar.template get<u32>(G);
        C.state = ar.template get<Walker>(G);
        C.status_a = ar.template get<u32>(G);
        C.status_b = ar.template get<u32>(G);
        C.nB = ar.template get<u32>(G);
        C.capB = cap_b;
        C.B = ar.template get<u32>(G * cap_b);
        C.over = ar.template get<Walker>(G);
        C.couple_x = ar.template get<u64>(G);
        C.couple_k = ar.template get<u32>(G);
        C.couple_r = ar.template get<u32>(G);
        for (int q = 0; q < 3; ++q) {
            const u64 words = u64(C.pass[q].n_chains) * kW / 32;
            C.abits[q] = ar.template get<u32>(words);
            bk.zero(C.abits[q], words * sizeof(u32));
        }
        bk.run(G, ChainAF{C});
        bk.run(G, ChainBF{C});
        std::vector<u32> nA(G), sa(G), sb(G), nB(G), ck(G), cr(G);
        bk.d2h(nA.data(), C.nA, G * sizeof(u32));
        bk.d2h(sa.data(), C.status_a, G * sizeof(u32));
        bk.d2h(sb.data(), C.status_b, G * sizeof(u32));
        bk.d2h(nB.data(), C.nB, G * sizeof(u32));
        bk.d2h(ck.data(), C.couple_k, G * sizeof(u32));
        bk.d2h(cr.data(), C.couple_r, G * sizeof(u32));
        {
            std::vector<u64> cx(G);
            bk.d2h(cx.data(), C.couple_x, G * sizeof(u64));
            for (int q = 0; q < 3; ++q)
                for (u64 j = 0; j + 1 < C.pass[q].n_chains; ++j) {
                    const u64 g = C.pass[q].chain0 + j;
                    if (sb[g] == kChainOk)
                        st.max_b_draws = std::max<u64>(st.max_b_draws, cx[g] - (C.pass[q].r0 + (j + 1) * kW));
                }
        }
        mark(1);

        // ---- the pools, and with them the first edge draw's counter ----
        std::vector<index_t> pool;
        const u64 c_first = pools_fn(pool);
        if (c_first < c0_lo) throw Declined{"pools ended before their lower bound"};

        // ---- per pass: stitch, repair, final row starts ----
        u64* F[3] = {nullptr, nullptr, nullptr};
        u64 s = c_first;
        std::map<u64, std::vector<u32>> a_cache;
        for (int q = 0; q < 3; ++q) {
            const PassDesc& Pq = P[q];
            const ChainPass& cp = C.pass[q];
            const bool has_self = Pq.self_base != kNone;
            const u64 need = u64(Pq.n_rows) + 1 + (has_self ? u64(kK) + 2 : 0);

            // the list of row starts chain g recorded (fetched once)
            auto a_list = [&](u64 g) -> const std::vector<u32>& {
                auto it = a_cache.find(g);
                if (it == a_cache.end()) {
                    std::vector<u32> a(nA[g]);
                    bk.d2h(a.data(), C.A + g * C.capA, a.size() * sizeof(u32));
                    it = a_cache.emplace(g, std::move(a)).first;
                }
                return it->second;
            };
            // whether x is a row start some chain of this pass recorded: its chain and rank
            auto recorded = [&](u64 x, u64& g_out, u64& r_out) {
                if (x < cp.r0 || x >= rend[q]) return false;
                const u64 k = (x - cp.r0) / kW;
                const u64 g = cp.chain0 + k;
                if (sa[g] != kChainOk) return false;
                const std::vector<u32>& a = a_list(g);
                const u32 off = static_cast<u32>(x - (cp.r0 + k * kW));
                const auto pos = std::lower_bound(a.begin(), a.end(), off);
                if (pos == a.end() || *pos != off) return false;
                g_out = g;
                r_out = static_cast<u64>(pos - a.begin());
                return true;
            };

            // exact host walk from the pass's first row start until it meets a chain
            std::vector<u64> host_rows;
            u64 x = s, g_in = 0, r_in = 0;
            bool walked_all = false;
            for (;;) {
                if (host_rows.size() == Pq.n_rows) {
                    // the host walk went through the whole pass: x is its end
                    host_rows.push_back(x);
                    walked_all = true;
                    break;
                }
                if (recorded(x, g_in, r_in)) break;
                if (host_rows.size() > (1u << 20)) throw Declined{"pass start did not meet the chains"};
                const u32 self = has_self ? Pq.self_base + static_cast<u32>(host_rows.size()) : kNone;
                host_rows.push_back(x);
                Walker w{x, 0, 0, 0, 0};
                while (!(walker_step(w, host_entry(seed, w.c, log1m_p, take), Pq, self) & kEvRowEnd)) {
                }
                x = w.c;
            }
            const u64 m = host_rows.size();
            st.host_rows[q] = static_cast<u32>(m);
            if (walked_all) {
                u64* Fq = ar.template get<u64>(m);
                bk.h2d(Fq, host_rows.data(), m * sizeof(u64));
                F[q] = Fq;
                s = host_rows.back();
                mark(2 + q);
                continue;
            }

            // the path through the chains
            std::vector<PathEntry> ent;
            std::vector<std::pair<u64, std::vector<u64>>> walked;  // (list position, starts) walked on the host
            u64 total = m;
            u64 g = g_in, r = r_in;
            while (total < need) {
                if (sa[g] != kChainOk)
                    throw Declined{"chain walk failed on the path (status " + std::to_string(sa[g]) + ")"};
                PathEntry pe{static_cast<u32>(g), static_cast<u32>(r), 0u, 0u, total};
                pe.na = static_cast<u32>(std::min<u64>(nA[g] - r, need - total));
                total += pe.na;
                if (total < need && sb[g] == kChainOk) {
                    pe.nb = static_cast<u32>(std::min<u64>(nB[g], need - total));
                    total += pe.nb;
                    ent.push_back(pe);
                    if (total >= need) break;
                    r = cr[g];
                    g = cp.chain0 + ck[g];
                    continue;
                }
                ent.push_back(pe);
                if (total >= need) break;
                // chain g met no later chain before its region ended (or is the
                // last one): the host walks on from its last recorded start, as
                // the chains walk (no own-target check), until it stands on a
                // start a later chain recorded, or has enough rows
                const std::vector<u32>& a = a_list(g);
                u64 xw = cp.r0 + (g - cp.chain0) * kW + a.back();
                std::vector<u64> piece;
                bool met = false;
                while (total + piece.size() < need) {
                    Walker w{xw, 0, 0, 0, 0};
                    while (!(walker_step(w, host_entry(seed, w.c, log1m_p, take), Pq, kNone) & kEvRowEnd)) {
                    }
                    xw = w.c;
                    if (recorded(xw, g, r)) {
                        met = true;
                        break;
                    }
                    piece.push_back(xw);
                }
                st.host_walked += piece.size();
                total += piece.size();
                walked.emplace_back(total - piece.size(), std::move(piece));
                if (!met) break;
            }
            u64* L = ar.template get<u64>(need);
            if (m > 0) bk.h2d(L, host_rows.data(), m * sizeof(u64));
            for (const auto& pw : walked)
                if (!pw.second.empty()) bk.h2d(L + pw.first, pw.second.data(), pw.second.size() * sizeof(u64));
            PathEntry* dent = ar.template get<PathEntry>(ent.size());
            bk.h2d(dent, ent.data(), ent.size() * sizeof(PathEntry));
            bk.zero(flags + 1, sizeof(u32));
            bk.run(ent.size(), PathF{PathArgs{T, C, dent, L, flags + 1}});
            ar.drop(dent);
            bk.d2h(hflags, flags, sizeof hflags);
            if (hflags[1] != 0) throw Declined{"path walk left the table"};

            if (!has_self) {
                F[q] = L;
            } else {
                // rows with a drawn target near their own position
                bk.zero(flags + 1, 2 * sizeof(u32));
                u64* rec_i = ar.template get<u64>(kRecordCap);
                int* rec_off = ar.template get<int>(kRecordCap);
                const u64 nchk = need - 1 - m;
                bk.run(nchk, VerifyF{VerifyArgs{T, Pq, L, m, kK, kRecordCap, flags + 2, rec_i, rec_off, flags + 1}});
                bk.d2h(hflags, flags, sizeof hflags);
                if (hflags[1] != 0) throw Declined{"candidate row starts inconsistent"};
                const u32 nrec = hflags[2];
                st.records[q] = nrec;
                std::vector<u64> hi(nrec), hj(nrec);
                std::vector<int> ho(nrec);
                std::vector<u32> hm(nrec);
                // the list's row starts as a bitmap over the pass's region
                const u64 lwords = u64(cp.n_chains) * kW / 32;
                u32* lw = ar.template get<u32>(lwords);
                bk.zero(lw, lwords * sizeof(u32));
                bk.run(need - m, MarkF{MarkArgs{L, m, lw, cp.r0, lwords * 32}});
                RepairArgs ra{T, Pq, L, need, Bitmap{lw, cp.r0, lwords * 32}, rec_i, rec_off, kRepairRows, nullptr,
                              nullptr, nullptr, nullptr, nullptr, nullptr, flags + 1};
                if (nrec > 0) {
                    ra.out_j = ar.template get<u64>(nrec);
                    ra.out_m = ar.template get<u32>(nrec);
                    ra.rows = ar.template get<u64>(u64(nrec) * kRepairRows);
                    bk.run(nrec, RepairF{ra});
                    bk.d2h(hflags, flags, sizeof hflags);
                    if (hflags[1] != 0) throw Declined{"repair walk left the table"};
                    bk.d2h(hi.data(), rec_i, nrec * sizeof(u64));
                    bk.d2h(ho.data(), rec_off, nrec * sizeof(int));
                    bk.d2h(hj.data(), ra.out_j, nrec * sizeof(u64));
                    bk.d2h(hm.data(), ra.out_m, nrec * sizeof(u32));
                }
                for (u32 k = 0; k < nrec; ++k)
                    if (hj[k] != kNoSelf) st.max_repair_rows = std::max(st.max_repair_rows, hm[k]);
                std::vector<u32> order(nrec);
                for (u32 k = 0; k < nrec; ++k) order[k] = k;
                std::sort(order.begin(), order.end(), [&](u32 a, u32 b) {
                    return hi[a] != hi[b] ? hi[a] < hi[b] : ho[a] < ho[b];
                });
                // compose: the shift d (true index = list index + d) changes only at autapses
                std::vector<SpliceSeg> seg;
                std::vector<u32> sel;
                std::vector<u64> spos;
                const u64 fneed = u64(Pq.n_rows) + 1;
                if (m > 0) seg.push_back(SpliceSeg{0, 0, m});
                u64 fpos = m, icur = m;
                i64 d = 0;
                for (u32 k : order) {
                    if (fpos >= fneed) break;
                    if (hi[k] < icur || ho[k] != d) continue;
                    const i64 t = static_cast<i64>(hi[k]) + d;
                    if (t >= static_cast<i64>(Pq.n_rows)) break;
                    if (hj[k] == kNoSelf) throw Declined{"repair did not rejoin the list"};
                    seg.push_back(SpliceSeg{fpos, icur, hi[k] - icur + 1});
                    fpos += hi[k] - icur + 1;
                    sel.push_back(k);
                    spos.push_back(fpos);
                    st.repairs[q]++;
                    if (hj[k] == kPassEnd) {
                        // the walk went through the pass's last row and wrote its end
                        fpos += hm[k];
                        if (fpos != fneed) throw Declined{"repair end inconsistent"};
                        break;
                    }
                    fpos += hm[k] - 1;
                    d = t + hm[k] - static_cast<i64>(hj[k]);
                    icur = hj[k];
                    st.max_shift = std::max<int>(st.max_shift, static_cast<int>(d < 0 ? -d : d));
                    if (d > kK || d < -kK) throw Declined{"index shift beyond the checked window"};
                }
                if (fpos < fneed) {
                    const u64 len_t = fneed - fpos;
                    if (icur + len_t > need) throw Declined{"candidate list too short"};
                    seg.push_back(SpliceSeg{fpos, icur, len_t});
                    fpos += len_t;
                }
                u64* Fq = ar.template get<u64>(fpos + 1);
                SpliceSeg* dseg = ar.template get<SpliceSeg>(seg.size());
                bk.h2d(dseg, seg.data(), seg.size() * sizeof(SpliceSeg));
                u64 chunks = 0;
                for (const auto& sg : seg) chunks += (sg.len + kCopyChunk - 1) / kCopyChunk;
                bk.run(chunks, CopyF{CopyArgs{dseg, static_cast<u32>(seg.size()), L, Fq}});
                if (!sel.empty()) {
                    u32* dsel = ar.template get<u32>(sel.size());
                    u64* dpos = ar.template get<u64>(spos.size());
                    bk.h2d(dsel, sel.data(), sel.size() * sizeof(u32));
                    bk.h2d(dpos, spos.data(), spos.size() * sizeof(u64));
                    ra.sel = dsel;
                    ra.pos = dpos;
                    ra.F = Fq;
                    bk.run(sel.size(), RepairCopyF{ra});
                }
                F[q] = Fq;
            }
            bk.d2h(&s, F[q] + Pq.n_rows, sizeof(u64));
            mark(2 + q);
        }
        bk.d2h(hflags, flags, sizeof hflags);
        if (hflags[1] != 0) throw Declined{"path walk left the table"};

        // ---- 4. exact walks: counts (and the end-to-start check), then edges ----
        u64 n_e[3], n_t[3];
        u64* eo[3];
        u64* to[3] = {nullptr, nullptr, nullptr};
        for (int q = 0; q < 3; ++q) {
            const u64 R = P[q].n_rows;
            u32* ne = ar.template get<u32>(R);
            u32* nt = P[q].tf ? ar.template get<u32>(R) : nullptr;
            bk.run(R, CountF{EmitArgs{T, P[q], F[q], ne, nt, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr,
                                      nullptr, nullptr, flags + 1}});
            eo[q] = ar.template get<u64>(R + 1);
            bk.zero(eo[q] + R, sizeof(u64));
            bk.run(R, WidenF{ne, eo[q]});
            bk.scan_u64(eo[q], R + 1);
            bk.d2h(&n_e[q], eo[q] + R, sizeof(u64));
            n_t[q] = 0;
            if (nt) {
                to[q] = ar.template get<u64>(R + 1);
                bk.zero(to[q] + R, sizeof(u64));
                bk.run(R, WidenF{nt, to[q]});
                bk.scan_u64(to[q], R + 1);
                bk.d2h(&n_t[q], to[q] + R, sizeof(u64));
            }
            ar.drop(ne);
            if (nt) ar.drop(nt);
        }
        bk.d2h(hflags, flags, sizeof hflags);
        if (hflags[1] != 0) throw Declined{"exact walk disagrees with the row starts"};
        mark(5);
        const u64 e_exc = n_e[0] + n_e[1], e_inh = n_e[2], n_tf = n_t[0] + n_t[1];
        if (e_exc >= (u64(1) << 31) || e_inh >= (u64(1) << 31)) throw Declined{"edge count beyond the sort range"};
        u32* exc_src = ar.template get<u32>(e_exc);
        u32* exc_dst = ar.template get<u32>(e_exc);
        u32* inh_src = ar.template get<u32>(e_inh);
        u32* inh_dst = ar.template get<u32>(e_inh);
        u32* tf_pre = ar.template get<u32>(n_tf);
        u32* tf_post = ar.template get<u32>(n_tf);
        u8* tf_slot = ar.template get<u8>(n_tf);
        for (int q = 0; q < 3; ++q) {
            EmitArgs ea{T, P[q], F[q], nullptr, nullptr, eo[q], to[q], q == 1 ? n_e[0] : 0, q == 1 ? n_t[0] : 0,
                        q == 2 ? inh_src : exc_src, q == 2 ? inh_dst : exc_dst, tf_pre, tf_post, tf_slot, flags + 1};
            bk.run(P[q].n_rows, EmitF{ea});
        }

        // ---- third-factor astrocytes ----
        u32* dpool = ar.template get<u32>(pool.size());
        bk.h2d(dpool, pool.data(), pool.size() * sizeof(u32));
        u32* astro = ar.template get<u32>(n_tf);
        u32* acount = ar.template get<u32>(n_astro);
        bk.zero(acount, u64(n_astro) * sizeof(u32));
        bk.run(n_tf, AstroF{AstroArgs{tf_post, tf_slot, dpool, take, astro, acount, flags + 1}});
        bk.d2h(hflags, flags, sizeof hflags);
        if (hflags[1] & 64u) throw Declined{"an astrocyte reaches the out-degree cap"};
        if (hflags[1] != 0) throw Declined{"exact walk left the table"};
        mark(6);

        // ---- inversion: stable sort by target, row starts from the target counts ----
        auto invert = [&](const u32* key, const u32* val, u64 n, u32 n_targets, HostCSR& out, int which) {
            int bits = 1;
            while (bits < 32 && (u64(1) << bits) < u64(n_targets)) ++bits;
            u32* ko = ar.template get<u32>(n);
            u32* vo = ar.template get<u32>(n);
            if (n > 0) bk.sort_pairs(key, val, ko, vo, n, bits);
            u32* cnt = ar.template get<u32>(u64(n_targets) + 1);
            bk.zero(cnt, (u64(n_targets) + 1) * sizeof(u32));
            bk.run(n, HistF{HistArgs{key, cnt}});
            bk.scan_u32(cnt, u64(n_targets) + 1);
            ar.drop(ko);
            if (kept != nullptr) {
                net.dev[which] = DeviceCSR{cnt, vo, n_targets, n};
                ar.keep(cnt);
                ar.keep(vo);
                kept->push_back(cnt);
                kept->push_back(vo);
                return;
            }
            out.row_start.resize(u64(n_targets) + 1);
            out.src.resize(n);
            bk.d2h(out.row_start.data(), cnt, (u64(n_targets) + 1) * sizeof(u32));
            if (n > 0) bk.d2h(out.src.data(), vo, n * sizeof(u32));
            ar.drop(vo);
            ar.drop(cnt);
        };
        invert(exc_dst, exc_src, e_exc, n_neu, net.inc_exc, 0);
        mark(7);
        invert(inh_dst, inh_src, e_inh, n_neu, net.inc_inh, 1);
        mark(8);
        invert(tf_post, astro, n_tf, n_neu, net.inc_sic, 2);
        mark(9);
        invert(astro, tf_pre, n_tf, n_astro, net.inc_e2a, 3);
        mark(10);
    } catch (const Declined& d) {
        why = d.why;
        if (kept != nullptr) {
            for (void* p : *kept) bk.release(p);
            kept->clear();
            for (auto& d2 : net.dev) d2 = DeviceCSR{};
        }
        return false;
    }
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace gb
}  // namespace astrosim
