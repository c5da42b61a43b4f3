// Host check of gpuboscia_amd/csrc/gb_stdsort.h: std_order_sort must leave
// every input in exactly the order std::sort leaves it, equal keys included
// (the candidates of top_k_infeasible: index, distance, ordered by distance
// descending).  Inputs: random sizes up to 700 with keys from small value
// sets (many ties), all-equal, ascending, descending, organ-pipe and
// sawtooth patterns, and a median-of-three adversary that exhausts the
// depth budget (heap sort fallback).
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../gpuboscia_amd/csrc/gb_stdsort.h"

struct Cand {
    int idx;
    double dist;
};

static int failures = 0;

static void check(const std::vector<Cand>& input, const char* what) {
    auto less = [](const Cand& a, const Cand& b) { return a.dist > b.dist; };
    std::vector<Cand> want = input, got = input;
    std::sort(want.begin(), want.end(), less);
    gb::std_order_sort(got.data(), got.data() + got.size(), less);
    for (size_t i = 0; i < want.size(); ++i)
        if (want[i].idx != got[i].idx || want[i].dist != got[i].dist) {
            if (failures < 10)
                std::printf("MISMATCH %s size %zu at %zu: want %d got %d\n", what, input.size(), i, want[i].idx,
                            got[i].idx);
            ++failures;
            return;
        }
}

static std::vector<Cand> from_keys(const std::vector<double>& k) {
    std::vector<Cand> v(k.size());
    for (size_t i = 0; i < k.size(); ++i) v[i] = {(int)i, k[i]};
    return v;
}

int main() {
    std::mt19937_64 rng(12345);
    int cases = 0;
    for (int t = 0; t < 20000; ++t) {
        const int n = (int)(rng() % 701);
        const int distinct = 1 + (int)(rng() % 12);
        std::vector<double> k(n);
        for (int i = 0; i < n; ++i) k[i] = 0.5 * (double)(rng() % distinct) / distinct;
        check(from_keys(k), "ties");
        ++cases;
    }
    for (int t = 0; t < 2000; ++t) {
        const int n = (int)(rng() % 701);
        std::vector<double> k(n);
        std::uniform_real_distribution<double> u(0.0, 0.5);
        for (int i = 0; i < n; ++i) k[i] = u(rng);
        check(from_keys(k), "distinct");
        ++cases;
    }
    for (int n : {0, 1, 2, 3, 15, 16, 17, 31, 32, 33, 64, 100, 255, 256, 257, 512, 700, 4096}) {
        std::vector<double> k(n);
        for (int i = 0; i < n; ++i) k[i] = 0.25;
        check(from_keys(k), "equal");
        for (int i = 0; i < n; ++i) k[i] = i;
        check(from_keys(k), "ascending");
        for (int i = 0; i < n; ++i) k[i] = n - i;
        check(from_keys(k), "descending");
        for (int i = 0; i < n; ++i) k[i] = std::min(i, n - 1 - i);
        check(from_keys(k), "organ pipe");
        for (int i = 0; i < n; ++i) k[i] = i % 7;
        check(from_keys(k), "sawtooth");
        cases += 5;
    }
    // Median-of-three adversary (Musser): drives the partition to split off
    // two elements per level, so the depth budget runs out and the heap sort
    // finishes the range.
    for (int n : {64, 250, 512, 1000, 2048}) {
        std::vector<double> k(n);
        const int h = n / 2;
        for (int i = 0; i < h; ++i) {
            k[i] = (i % 2 == 0) ? i + 1 : h + i + (h % 2 == 0 ? 0 : 1);
            k[h + i] = 2 * (i + 1);
        }
        // the comparator is "greater", so negate to feed the adversary
        for (double& v : k) v = -v;
        check(from_keys(k), "adversary");
        ++cases;
    }
    if (failures) {
        std::printf("%d of %d cases differ from std::sort\n", failures, cases);
        return 1;
    }
    std::printf("%d cases: all checks passed\n", cases);
    return 0;
}
