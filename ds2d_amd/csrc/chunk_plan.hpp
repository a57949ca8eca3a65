// Chunk plan of the two-step elastic kernel over a batch of uneven windows.
//
// A workgroup of k_fused2_step marches one chunk: columns [x0, x1) of one
// window and one block of rows.  Its time is close to linear in the width:
// on 128 windows of 440-1240 padded columns (2000x600 model, 704 lanes) a chunk
// of w columns took 63 us + 2.80 us per column with the PML and 38 us + 1.90 us
// with the random boundary, i.e. w plus 20-23 columns, the 12 + 12 columns of
// the march outside the chunk.  One workgroup fits a CU (112 KB of LDS), so a
// launch is a list-scheduling problem: the k_fused2_step launch gives XCD x
// the chunks [x*per, (x+1)*per) in order (block b runs chunk
// (b % 8) * per + b / 8), and each of the XCD's CUs takes the next chunk
// when its current one ends.
//
// Fixed-width chunks leave CUs idle at the end of a launch: with 240-column
// chunks on that batch every XCD gets 64 chunks, most of them 240 wide, and the CUs
// that draw two of those end 12 % after the mean.  The balanced plan instead
//   (1) deals whole row-block units (window, row block) to the eight XCDs,
//       largest first, each to the least-loaded XCD (neighbouring chunks of a
//       window share their halo columns in that XCD's L2);
//   (2) gives every XCD the same number K of chunks, a multiple of its CU
//       count, splitting each unit into k equal-width chunks, the k chosen by
//       a local search on the simulated list schedule of the XCD;
//   (3) orders each XCD's chunks widest first (longest-processing-time
//       order: the CUs that drew the narrower first chunks take the next
//       wide ones).
// It is used when its simulated makespan beats the fixed-width list's.
// Results do not depend on the plan: every owned cell is computed by the same
// code from the same inputs whichever chunk holds it.
#pragma once
#include <algorithm>
#include <functional>
#include <queue>
#include <vector>

namespace ds2d {

struct PlanChunk {
    int slot, x0, x1, z0;
};

constexpr int kPlanXcds = 8;       // XCDs of gfx950 (the launch's block -> chunk map assumes 8)
constexpr int kPlanHalo = 24;      // march columns outside a chunk (cost model: w + 24)
constexpr int kPlanMinWidth = 16;  // narrowest chunk the plan makes

// Makespan of running chunks of the given costs in order on ncu CUs, each CU
// taking the next chunk when it becomes free.
inline long long plan_makespan(const std::vector<int>& costs, int ncu) {
    std::priority_queue<long long, std::vector<long long>, std::greater<long long>> free_at;
    for (int i = 0; i < ncu; ++i) free_at.push(0);
    long long end = 0;
    for (int c : costs) {
        const long long t = free_at.top() + c;
        free_at.pop();
        free_at.push(t);
        end = std::max(end, t);
    }
    return end;
}

// Fixed-width chunks of every (window, row block) in slot order, and the
// simulated makespan of the launch's XCD mapping over them.
inline long long plan_fixed(const std::vector<int>& nxp, const std::vector<int>& nzp, int rows, int W, int cpx,
                            std::vector<PlanChunk>& out) {
    out.clear();
    for (int s = 0; s < (int)nxp.size(); ++s)
        for (int x0 = 0; x0 < nxp[s]; x0 += W)
            for (int z0 = 0; z0 < nzp[s]; z0 += rows) out.push_back({s, x0, std::min(x0 + W, nxp[s]), z0});
    const int n = (int)out.size(), per = (n + kPlanXcds - 1) / kPlanXcds;
    long long span = 0;
    for (int x = 0; x < kPlanXcds; ++x) {
        std::vector<int> c;
        for (int i = x * per; i < std::min(n, (x + 1) * per); ++i) c.push_back(out[i].x1 - out[i].x0 + kPlanHalo);
        span = std::max(span, plan_makespan(c, cpx));
    }
    return span;
}

namespace plan_detail {

inline void split_costs(const std::vector<int>& w, const std::vector<int>& k, std::vector<int>& costs) {
    costs.clear();
    for (size_t i = 0; i < w.size(); ++i) {
        const int base = w[i] / k[i], rem = w[i] % k[i];
        for (int j = 0; j < k[i]; ++j) costs.push_back(base + (j < rem ? 1 : 0) + kPlanHalo);
    }
    std::sort(costs.begin(), costs.end(), std::greater<int>());
}

// Chunk counts k (sum K, 1 <= k_i <= w_i / kPlanMinWidth) of one XCD's units;
// returns the simulated makespan, or -1 when K chunks cannot be made.
inline long long choose_counts(const std::vector<int>& w, int K, int cpx, std::vector<int>& k) {
    const int m = (int)w.size();
    std::vector<int> cap(m);
    long long capsum = 0;
    for (int i = 0; i < m; ++i) {
        cap[i] = std::max(1, w[i] / kPlanMinWidth);
        capsum += cap[i];
    }
    if (m > K || capsum < K) return -1;
    // smallest maximum width whose ceil-split needs at most K chunks, then
    // split the widest chunks further until there are exactly K
    int lo = 1, hi = *std::max_element(w.begin(), w.end());
    auto need = [&](int W) {
        long long s = 0;
        for (int i = 0; i < m; ++i) s += std::min(cap[i], (w[i] + W - 1) / W);
        return s;
    };
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (need(mid) <= K) hi = mid; else lo = mid + 1;
    }
    k.assign(m, 1);
    int sum = 0;
    for (int i = 0; i < m; ++i) sum += (k[i] = std::min(cap[i], (w[i] + lo - 1) / lo));
    while (sum < K) {
        int best = -1;
        for (int i = 0; i < m; ++i)
            if (k[i] < cap[i] && (best < 0 || (long long)w[i] * k[best] > (long long)w[best] * k[i])) best = i;
        ++k[best];
        ++sum;
    }
    std::vector<int> costs;
    split_costs(w, k, costs);
    long long span = plan_makespan(costs, cpx);
    // no schedule beats the mean load per CU or the longest chunk
    long long total = 0;
    for (int c : costs) total += c;
    const long long bound = std::max<long long>((total + cpx - 1) / cpx, costs.front());
    // local search: move one chunk from unit j to unit i while it helps
    std::vector<int> wide(m);
    for (int pass = 0; pass < 8 && span > bound; ++pass) {
        bool moved = false;
        // receivers: the units with the widest chunks (at most 8 of them)
        for (int i = 0; i < m; ++i) wide[i] = i;
        const int nw = std::min(m, 8);
        std::partial_sort(wide.begin(), wide.begin() + nw, wide.end(), [&](int a, int b) {
            return (long long)w[a] * k[b] > (long long)w[b] * k[a];
        });
        for (int ii = 0; ii < nw && span > bound; ++ii)
            for (int j = 0; j < m && span > bound; ++j) {
                const int i = wide[ii];
                if (i == j || k[i] >= cap[i] || k[j] <= 1) continue;
                ++k[i];
                --k[j];
                split_costs(w, k, costs);
                const long long v = plan_makespan(costs, cpx);
                if (v < span) {
                    span = v;
                    moved = true;
                } else {
                    --k[i];
                    ++k[j];
                }
            }
        if (!moved) break;
    }
    return span;
}

}  // namespace plan_detail

// Balanced plan (see the top of this file).  cpx: CUs per XCD.  Returns the
// simulated makespan, or -1 when no plan of 1-4 rounds exists.
inline long long plan_balanced(const std::vector<int>& nxp, const std::vector<int>& nzp, int rows, int cpx,
                               std::vector<PlanChunk>& out) {
    struct Unit {
        int slot, z0, w;
    };
    std::vector<Unit> units;
    for (int s = 0; s < (int)nxp.size(); ++s)
        for (int z0 = 0; z0 < nzp[s]; z0 += rows) units.push_back({s, z0, nxp[s]});
    std::vector<int> order(units.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return units[a].w > units[b].w; });
    std::vector<std::vector<int>> bins(kPlanXcds);
    std::vector<long long> load(kPlanXcds, 0);
    for (int u : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        bins[x].push_back(u);
        load[x] += units[u].w + kPlanHalo;
    }
    long long best = -1;
    std::vector<std::vector<int>> bestk(kPlanXcds);
    for (int R = 1; R <= 4; ++R) {
        const int K = R * cpx;
        // lower bound of this R: the busiest XCD's columns plus K chunk overheads over its CUs
        long long lb = 0;
        for (int x = 0; x < kPlanXcds; ++x) {
            long long cols = 0;
            for (int u : bins[x]) cols += units[u].w;
            lb = std::max(lb, (cols + (long long)kPlanHalo * K + cpx - 1) / cpx);
        }
        if (best >= 0 && lb >= best) break;  // larger R only adds overhead
        long long span = 0;
        std::vector<std::vector<int>> ks(kPlanXcds);
        for (int x = 0; x < kPlanXcds && span >= 0; ++x) {
            std::vector<int> w;
            for (int u : bins[x]) w.push_back(units[u].w);
            if (w.empty()) { span = -1; break; }
            const long long v = plan_detail::choose_counts(w, K, cpx, ks[x]);
            span = v < 0 ? -1 : std::max(span, v);
        }
        if (span >= 0 && (best < 0 || span < best)) {
            best = span;
            bestk = ks;
        }
    }
    out.clear();
    if (best < 0) return -1;
    for (int x = 0; x < kPlanXcds; ++x) {
        std::vector<PlanChunk> c;
        for (size_t i = 0; i < bins[x].size(); ++i) {
            const Unit& u = units[bins[x][i]];
            const int k = bestk[x][i], base = u.w / k, rem = u.w % k;
            for (int j = 0, x0 = 0; j < k; ++j) {
                const int wj = base + (j < rem ? 1 : 0);
                c.push_back({u.slot, x0, x0 + wj, u.z0});
                x0 += wj;
            }
        }
        std::stable_sort(c.begin(), c.end(),
                         [](const PlanChunk& a, const PlanChunk& b) { return a.x1 - a.x0 > b.x1 - b.x0; });
        out.insert(out.end(), c.begin(), c.end());
    }
    return best;
}

}  // namespace ds2d
