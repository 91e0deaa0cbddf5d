// tune.hip -- autotuner for the force pass of a batch of replicas.
//
// Replicas are grouped into buckets by log2 of their particle count.  For
// each workgroup size, the tile width of each bucket is chosen by
// coordinate descent on a measured step cost of the whole batch: the pair
// pass (with its energies, as every step forms them) plus kUpdateShare of a
// neighbour-list update of every replica (a step updates the lists of only
// the replicas that rebuild); the best (block, widths) pair is kept.  Each
// cost is the smaller of two timings, and a width replaces the current one
// only when it is at least 1 % cheaper, so timing noise does not pick.  Tile shape never
// changes the arithmetic of a particle's force (its candidates are visited
// cell by cell in sorted order whatever the tile), so tuning cannot change
// results.  Choices are cached in a text file, one line per bucket
// signature: "<signature> <block> <bucket>:<width> ...".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "handles.hpp"

namespace {

std::string signature(const fm::Engine &E, const std::vector<int> &bucket) {
    std::ostringstream os;
    os << "R" << E.nrep();
    std::map<int, int> cnt;
    for (int b : bucket) cnt[b]++;
    for (auto &kv : cnt) os << "_b" << kv.first << "x" << kv.second;
    return os.str();
}

bool load_cache(const char *path, const std::string &sig, int &block, std::map<int, int> &w) {
    if (!path) return false;
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        // one record per line: "<sig> <block> <bucket>:<W> ..."
        std::istringstream is(line);
        std::string s;
        is >> s;
        if (s != sig) continue;
        is >> block;
        std::string kv;
        while (is >> kv) {
            int b = 0, v = 0;
            if (std::sscanf(kv.c_str(), "%d:%d", &b, &v) == 2) w[b] = v;
        }
        return block > 0;
    }
    return false;
}

void save_cache(const char *path, const std::string &sig, int block, const std::map<int, int> &w) {
    if (!path) return;
    std::vector<std::string> keep;
    {
        std::ifstream f(path);
        std::string line;
        while (f && std::getline(f, line)) {
            std::istringstream is(line);
            std::string s;
            is >> s;
            if (s != sig && !s.empty()) keep.push_back(line);
        }
    }
    std::ofstream f(path, std::ios::trunc);
    for (auto &l : keep) f << l << "\n";
    f << sig << " " << block;
    for (auto &kv : w) f << " " << kv.first << ":" << kv.second;
    f << "\n";
}

}  // namespace

extern "C" int fm_batch_tune(fm_batch *b, const char *cache_path) {
    if (!b) return FM_ERR_ARG;
    if (!b->eng) { b->err = "batch build failed"; return FM_ERR_STATE; }
    fm::Engine &E = *b->eng;
    const int R = E.nrep();
    std::vector<int> bucket(R);
    for (int r = 0; r < R; ++r) bucket[r] = (int)std::floor(std::log2((double)std::max(1, E.nparticles(r))));
    const std::string sig = signature(E, bucket);
    std::map<int, int> best_w;
    int best_block = 0;
    auto widths_for = [&](const std::map<int, int> &w) {
        std::vector<int> tw(R, 0);
        for (int r = 0; r < R; ++r) {
            auto it = w.find(bucket[r]);
            tw[r] = (it == w.end()) ? 0 : it->second;
        }
        return tw;
    };
    if (load_cache(cache_path, sig, best_block, best_w)) {
        return E.retile(widths_for(best_w), best_block, b->err);
    }
    std::map<int, int> buckets;
    for (int r = 0; r < R; ++r) buckets[bucket[r]] = 0;
    // list updates per step as a share of updating every replica's lists
    // (0.10-0.17 measured on the 512-replica ensemble of uneven sizes)
    constexpr double kUpdateShare = 0.15;
    auto measure = [&](const std::map<int, int> &w, int blk, double &cost) -> int {
        int rc = E.retile(widths_for(w), blk, b->err);
        if (rc) return rc;
        cost = 1e30;
        for (int trial = 0; trial < 2; ++trial) {
            double upd = 0, pass = 0;
            rc = E.time_parts(4, &upd, &pass, b->err);
            if (rc) return rc;
            cost = std::min(cost, pass + kUpdateShare * upd);
        }
        return 0;
    };
    double best_ms = 1e30;
    const int blocks[3] = {256, 128, 64};  // on a tie the first (the default) stays
    const int cand_w[6] = {0, 2, 4, 8, 16, 32};
    for (int blk : blocks) {
        std::map<int, int> w = buckets;  // 0 = width from occupancy
        double cur = 0;
        int rc = measure(w, blk, cur);
        if (rc) return rc;
        for (auto &kv : buckets) {
            const int bk = kv.first;
            int keep = w[bk];
            for (int cw : cand_w) {
                if (cw == keep) continue;
                w[bk] = cw;
                double t = 0;
                rc = measure(w, blk, t);
                if (rc) return rc;
                if (t < 0.99 * cur) {
                    cur = t;
                    keep = cw;
                }
            }
            w[bk] = keep;
        }
        if (cur < 0.99 * best_ms) {
            best_ms = cur;
            best_block = blk;
            best_w = w;
        }
    }
    save_cache(cache_path, sig, best_block, best_w);
    return E.retile(widths_for(best_w), best_block, b->err);
}
